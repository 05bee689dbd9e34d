"""GPU parity of sparsification (eSparseBipartite / clrSparseBipartite + delZeroDeg,
apps/hyper/butterfly_utils.h:744-996) and of the approximate TOTAL count with the driver's
scaling (BipartiteButterflies.C:213-214).

The sparsified graphs are compared array for array (bit-exact) with the graphs the compiled
reference wrote into tests/golden/sparse_cases.npz, and with the oracle (pinned to those graphs
by tests/test_oracle_sparse.py) on larger graphs up to the 2^24-edge Chung-Lu graph.
"""
import os

import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O
from tests.graphs import complete_bipartite, empty_graph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    with np.load(os.path.join(GOLDEN, "sparse_cases.npz")) as z:
        keys = sorted({k.rsplit("__", 1)[0] for k in z.files})
        return [(k, z[k + "__sizes"], z[k + "__csr"], z[k + "__totals"]) for k in keys]


def _arrays(h):
    """(nv, nu, m, offsetsV, edgesV, offsetsU, edgesU) of a host or device graph, as numpy."""
    if isinstance(h, pb.DeviceGraph):
        a = [x.numpy() for x in (h.offsetsV, h.edgesV, h.offsetsU, h.edgesU)]
        a = [a[0][:h.nv + 1], a[1][:h.m], a[2][:h.nu + 1], a[3][:h.m]]
    else:
        a = [h.offsetsV, h.edgesV, h.offsetsU, h.edgesU]
    return (h.nv, h.nu, h.m, *[np.asarray(x, np.uint32) for x in a])


def _equal(x, y):
    return x[:3] == y[:3] and all(np.array_equal(p, q) for p, q in zip(x[3:], y[3:]))


def _from_golden(sizes, csr):
    nv, nu, m = (int(s) for s in sizes)
    a = np.asarray(csr, np.uint32)
    return (nv, nu, m, a[:nv + 1], a[nv + 1:nv + 1 + m], a[nv + 1 + m:nv + 2 + m + nu], a[nv + 2 + m + nu:])


def test_sparsify_matches_reference_golden(ctx):
    for key, sizes, csr, totals in _golden():
        name, case = key.split("__")
        st, d, seed = case.split("_")
        d, seed = int(d), int(seed)
        print("case", key, flush=True)
        g = pb.Graph.read(os.path.join(GOLDEN, name + ".adj"))
        want = _from_golden(sizes, csr)
        assert _equal(_arrays(ctx.sparsify(g, st, d, seed)), want), key
        dg = g.to_device(ctx)
        assert _equal(_arrays(ctx.sparsify(dg, st, d, seed)), want), key + " device"
        for h in (g, dg):
            assert ctx.count_total_sparse(h, st, d, seed) == (int(totals[1]), int(totals[0])), key


@pytest.mark.parametrize("st", ["EDGE", "COLOR"])
def test_sparsify_against_oracle(ctx, st):
    graphs = [("chunglu-20000x5000", pb.Graph.chung_lu(20000, 5000, 200000, 2.1, 2.1, 5)),
              ("rmat-14x13", pb.Graph.rmat(14, 13, 150000, 0.5, 0.1, 0.1, 0.3, 6)),
              ("uniform-3000x3000", pb.Graph.uniform(3000, 3000, 90000, 8)),
              ("K_6_9", complete_bipartite(6, 9))]
    for tag, g in graphs:
        dg = g.to_device(ctx)
        for d, seed in ((1, 0), (2, 3), (7, 0), (25, 99)):
            print("case", tag, st, d, seed, flush=True)
            h = O.sparsify(g, st, d, seed)
            want = _arrays(h)
            assert _equal(_arrays(ctx.sparsify(dg, st, d, seed)), want), (tag, d, seed)
            approx, exact = O.approx_total(g, st, d, seed)
            for rt in ("ADEG", "ACOCORE"):
                assert ctx.count_total_sparse(dg, st, d, seed, "GPU", rt) == (approx, exact), (tag, d, seed, rt)


def test_sparse_edge_cases(ctx):
    # no edges: every vertex is dropped
    e = empty_graph(7, 5)
    for st in ("EDGE", "COLOR"):
        h = ctx.sparsify(e, st, 3, 1)
        assert (h.nv, h.nu, h.m) == (0, 0, 0)
        assert ctx.count_total_sparse(e, st, 3, 1) == (0, 0)
    # NONE counts the graph itself
    k = complete_bipartite(4, 5)
    assert ctx.count_total_sparse(k, "NONE", 25, 0) == (60, 60)
    # argument checks
    for st, d, seed in (("EDGE", 0, 0), ("COLOR", -3, 0), ("EDGE", 2, -1), ("EDGE", 1 << 32, 0)):
        with pytest.raises(pb.PbfError) as ei:
            ctx.sparsify(k, st, d, seed)
        assert ei.value.name == "PBF_ERR_ARG"
    with pytest.raises(pb.PbfError):
        ctx.count_total_sparse(k, 3, 2, 0)


def test_sparsify_chunglu_2p24(ctx):
    # the headline graph: Chung-Lu nv = 2^20, nu = 2^18, m = 2^24, gamma 2.1, seed 2
    g = pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)
    dg = g.to_device(ctx)
    for st in ("EDGE", "COLOR"):
        print("case chunglu-2^24", st, flush=True)
        h = O.sparsify(g, st, 25, 0)
        got = ctx.sparsify(dg, st, 25, 0)
        assert _equal(_arrays(got), _arrays(h)), st
        approx, exact = O.approx_total(g, st, 25, 0)
        assert ctx.count_total_sparse(dg, st, 25, 0) == (approx, exact), st
