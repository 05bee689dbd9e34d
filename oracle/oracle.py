"""ctypes view of the C oracle (liboracle.so) and a runner for the compiled reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline as a checker.  The product never imports this module.
"""
import ctypes
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")


def build(verbose=False):
    """Builds liboracle.so and, when the reference tree is present, _ref/ref_harness."""
    r = subprocess.run(["make", "-C", HERE, "all"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if verbose or r.returncode:
        print(r.stdout)
    if r.returncode:
        raise RuntimeError("oracle build failed")


class _G(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int64), ("nu", ctypes.c_int64), ("m", ctypes.c_int64),
                ("offV", ctypes.c_void_p), ("edgV", ctypes.c_void_p), ("offU", ctypes.c_void_p),
                ("edgU", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        P, I64, INT = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        PG = ctypes.POINTER(_G)
        sig = {
            "orc_side": (None, [PG, ctypes.POINTER(INT), ctypes.POINTER(I64)]),
            "orc_rank_adeg": (None, [PG, P, P]),
            "orc_rank_acocore": (None, [PG, P, P, P]),
            "orc_rank_wedges": (I64, [PG, P]),
            "orc_rank_graph": (None, [PG, INT, P, P, P, P, P]),
            "orc_count_vert": (INT, [PG, INT, INT, P]),
            "orc_count_edge": (INT, [PG, INT, INT, P]),
            "orc_count_total": (INT, [PG, INT, INT, P]),
            "orc_edge_to_idxs": (None, [PG, INT, P]),
            "orc_peel_vert": (None, [PG, INT, P, P]),
            "orc_peel_edge": (None, [PG, INT, P, P]),
            "orc_brute_vert": (ctypes.c_uint64, [PG, INT, P]),
            "orc_brute_edge": (None, [PG, INT, P]),
            "orc_brute_peel_vert": (None, [PG, INT, P]),
            "orc_brute_peel_edge": (None, [PG, INT, P]),
            "orc_sparsify": (INT, [PG, INT, I64, I64, ctypes.POINTER(I64), ctypes.POINTER(I64),
                                   ctypes.POINTER(I64), P, P, P, P]),
            "orc_hash64": (ctypes.c_uint64, [ctypes.c_uint64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


RANK = {"ACOCORE": 2, "ADEG": 4}


def _g(graph):
    """Any object with nv, nu, m and the four uint32 numpy arrays."""
    arrs = [np.ascontiguousarray(getattr(graph, k), dtype=np.uint32) for k in
            ("offsetsV", "edgesV", "offsetsU", "edgesU")]
    g = _G(graph.nv, graph.nu, graph.m, *[a.ctypes.data for a in arrs])
    g._keep = arrs
    return g


def side(graph):
    uv, w = ctypes.c_int(), ctypes.c_int64()
    g = _g(graph)
    lib().orc_side(ctypes.byref(g), ctypes.byref(uv), ctypes.byref(w))
    return bool(uv.value), w.value


def rank(graph, rank_type="ADEG"):
    """(order, rank_of, core) -- core is the ACOCORE level per vertex (None for ADEG)."""
    g = _g(graph)
    n = graph.nv + graph.nu
    order = np.zeros(max(n, 1), np.uint32)
    rank_of = np.zeros(max(n, 1), np.uint32)
    core = None
    if rank_type == "ADEG":
        lib().orc_rank_adeg(ctypes.byref(g), order.ctypes.data, rank_of.ctypes.data)
    else:
        core = np.zeros(max(n, 1), np.uint32)
        lib().orc_rank_acocore(ctypes.byref(g), order.ctypes.data, rank_of.ctypes.data, core.ctypes.data)
        core = core[:n]
    return order[:n], rank_of[:n], core


def rank_wedges(graph, rank_of):
    g = _g(graph)
    r = np.ascontiguousarray(rank_of, dtype=np.uint32)
    return int(lib().orc_rank_wedges(ctypes.byref(g), r.ctypes.data))


def rank_graph(graph, use_v, order, rank_of):
    g = _g(graph)
    n = graph.nv + graph.nu
    roff = np.zeros(n + 1, np.uint32)
    redg = np.zeros(max(2 * graph.m, 1), np.uint32)
    pos = np.zeros(max(2 * graph.m, 1), np.uint32)
    o = np.ascontiguousarray(order, dtype=np.uint32)
    r = np.ascontiguousarray(rank_of, dtype=np.uint32)
    lib().orc_rank_graph(ctypes.byref(g), int(use_v), o.ctypes.data, r.ctypes.data, roff.ctypes.data,
                         redg.ctypes.data, pos.ctypes.data)
    return roff, redg[:2 * graph.m], pos[:2 * graph.m]


def count_vert(graph, use_v, rank_type="ADEG"):
    g = _g(graph)
    ns = graph.nu if use_v else graph.nv
    out = np.zeros(max(ns, 1), np.uint64)
    assert lib().orc_count_vert(ctypes.byref(g), int(use_v), RANK[rank_type], out.ctypes.data) == 0
    return out[:ns]


def count_edge(graph, use_v, rank_type="ADEG"):
    g = _g(graph)
    out = np.zeros(max(graph.m, 1), np.uint64)
    assert lib().orc_count_edge(ctypes.byref(g), int(use_v), RANK[rank_type], out.ctypes.data) == 0
    return out[:graph.m]


def count_total(graph, use_v, rank_type="ADEG"):
    g = _g(graph)
    t = np.zeros(1, np.uint64)
    assert lib().orc_count_total(ctypes.byref(g), int(use_v), RANK[rank_type], t.ctypes.data) == 0
    return int(t[0])


def edge_to_idxs(graph, use_v):
    g = _g(graph)
    out = np.zeros(max(graph.m, 1), np.uint32)
    lib().orc_edge_to_idxs(ctypes.byref(g), int(use_v), out.ctypes.data)
    return out[:graph.m]


def peel_vert(graph, use_v, counts):
    g = _g(graph)
    ns = graph.nu if use_v else graph.nv
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    tip = np.zeros(max(ns, 1), np.uint64)
    lib().orc_peel_vert(ctypes.byref(g), int(use_v), c.ctypes.data, tip.ctypes.data)
    return tip[:ns]


def peel_edge(graph, use_v, counts):
    g = _g(graph)
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    wing = np.zeros(max(graph.m, 1), np.uint64)
    lib().orc_peel_edge(ctypes.byref(g), int(use_v), c.ctypes.data, wing.ctypes.data)
    return wing[:graph.m]


def brute_vert(graph, use_v):
    g = _g(graph)
    ns = graph.nu if use_v else graph.nv
    out = np.zeros(max(ns, 1), np.uint64)
    tot = lib().orc_brute_vert(ctypes.byref(g), int(use_v), out.ctypes.data)
    return out[:ns], int(tot)


def brute_edge(graph, use_v):
    g = _g(graph)
    out = np.zeros(max(graph.m, 1), np.uint64)
    lib().orc_brute_edge(ctypes.byref(g), int(use_v), out.ctypes.data)
    return out[:graph.m]


def brute_peel_vert(graph, use_v):
    g = _g(graph)
    ns = graph.nu if use_v else graph.nv
    out = np.zeros(max(ns, 1), np.uint64)
    lib().orc_brute_peel_vert(ctypes.byref(g), int(use_v), out.ctypes.data)
    return out[:ns]


def brute_peel_edge(graph, use_v):
    g = _g(graph)
    out = np.zeros(max(graph.m, 1), np.uint64)
    lib().orc_brute_peel_edge(ctypes.byref(g), int(use_v), out.ctypes.data)
    return out[:graph.m]


class CSR:
    """A plain graph record (nv, nu, m and the four uint32 arrays)."""

    def __init__(self, nv, nu, offsetsV, edgesV, offsetsU, edgesU):
        self.nv, self.nu = int(nv), int(nu)
        self.offsetsV, self.edgesV = np.asarray(offsetsV, np.uint32), np.asarray(edgesV, np.uint32)
        self.offsetsU, self.edgesU = np.asarray(offsetsU, np.uint32), np.asarray(edgesU, np.uint32)
        self.m = int(self.edgesV.shape[0])


SPARSE = {"COLOR": 1, "EDGE": 2}


def sparsify(graph, sparse_type="EDGE", denom=25, seed=0):
    """eSparseBipartite / clrSparseBipartite + delZeroDeg; returns a CSR."""
    g = _g(graph)
    m = max(graph.m, 1)
    oV, eV = np.zeros(graph.nv + 1, np.uint32), np.zeros(m, np.uint32)
    oU, eU = np.zeros(graph.nu + 1, np.uint32), np.zeros(m, np.uint32)
    nv, nu, mo = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    assert lib().orc_sparsify(ctypes.byref(g), SPARSE[sparse_type], int(denom), int(seed), ctypes.byref(nv),
                              ctypes.byref(nu), ctypes.byref(mo), oV.ctypes.data, eV.ctypes.data, oU.ctypes.data,
                              eU.ctypes.data) == 0
    assert mo.value >= 0, "V and U sides kept different edge counts"
    return CSR(nv.value, nu.value, oV[:nv.value + 1], eV[:mo.value], oU[:nu.value + 1], eU[:mo.value])


def hash64(a):
    """hashInt(ulong) of ligra/utils.h:389-397."""
    return int(lib().orc_hash64(int(a) & (2 ** 64 - 1)))


def approx_total(graph, sparse_type="EDGE", denom=25, seed=0, rank_type="ADEG"):
    """(scaled total, exact total of the sparsified graph), scaled as the driver does
    (BipartiteButterflies.C:213-214): the long total times pow(denom, 3 or 4), truncated."""
    h = sparsify(graph, sparse_type, denom, seed)
    t = count_total(h, side(h)[0], rank_type) if h.m else 0
    k = 3 if sparse_type == "COLOR" else 4
    return int(float(t) * float(denom) ** k), t


# ---------------- compiled reference (oracle/_ref) ----------------
def ref_available():
    return os.path.exists(REF_HARNESS)


def run_ref(op, graph_path, out_path, count_type="SERIAL", rank_type="ADEG", peel_type="BATCHS", rounds=1,
            threads=None, extra=(), timeout=600):
    """Runs the compiled reference harness; returns (use_v, side_w, counts, peel, info)."""
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    cmd = [REF_HARNESS, op, graph_path, out_path, "-countType", count_type, "-rankType", rank_type,
           "-peelType", peel_type, "-rounds", str(rounds), *extra]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"ref harness failed ({r.returncode}): {r.stderr[-2000:]}")
    info = json.loads(r.stdout.strip().splitlines()[-1])
    with open(out_path, "rb") as f:
        use_v, side_w, nc = np.frombuffer(f.read(24), dtype=np.int64)
        counts = np.frombuffer(f.read(8 * int(nc)), dtype=np.int64).copy()
        (npl,) = np.frombuffer(f.read(8), dtype=np.int64)
        peel = np.frombuffer(f.read(8 * int(npl)), dtype=np.int64).copy()
    return bool(use_v), int(side_w), counts, peel, info
