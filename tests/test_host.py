"""Host-side checks that need no GPU: the C-ABI library loads and exports every function that
include/parbutterfly.h declares, the AdjacencyHypergraph reader's error codes (the reference's
abort / exit(0) cases of butterfly_utils.h:1013-1051 become status codes), the writer / reader
round trip, and the seeded generators' determinism and shapes."""
import ctypes
import os
import re

import numpy as np
import pytest

import parbutterfly_amd as pb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "parbutterfly.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pb.library_path())
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert set(names) <= set(pb.EXPORTED_SYMBOLS), set(names) - set(pb.EXPORTED_SYMBOLS)


def test_abi_version_and_status_strings():
    assert pb._lib.pbf_abi_version() == 2
    for s in range(11):
        assert pb._lib.pbf_status_str(s)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reader_round_trip(tmp_path):
    g = pb.Graph.chung_lu(300, 200, 2000, 2.1, 2.1, 7)
    p = str(tmp_path / "g.adj")
    g.write(p)
    h = pb.Graph.read(p)
    for k in ("offsetsV", "edgesV", "offsetsU", "edgesU"):
        assert np.array_equal(getattr(g, k), getattr(h, k))
    b = str(tmp_path / "g.bcsr")
    g.write_binary(b)
    h = pb.Graph.read(b)
    assert np.array_equal(g.edgesU, h.edgesU) and (h.nv, h.nu, h.m) == (g.nv, g.nu, g.m)


def _status(fn):
    with pytest.raises(pb.PbfError) as e:
        fn()
    return e.value.name


def test_reader_errors(tmp_path):
    # bad header (butterfly_utils.h:1013-1016)
    p = _write(tmp_path, "a.adj", "AdjacencyGraph\n1\n1\n1\n1\n0\n0\n0\n0\n")
    assert _status(lambda: pb.Graph.read(p)) == "PBF_ERR_FORMAT"
    # mv != mu (:1024)
    p = _write(tmp_path, "b.adj", "AdjacencyHypergraph\n1\n1\n1\n2\n0\n0\n0\n0\n0\n")
    assert _status(lambda: pb.Graph.read(p)) == "PBF_ERR_FORMAT"
    # length mismatch
    p = _write(tmp_path, "c.adj", "AdjacencyHypergraph\n1\n1\n1\n1\n0\n0\n0\n")
    assert _status(lambda: pb.Graph.read(p)) == "PBF_ERR_FORMAT"
    # neighbour out of range (:1039-1041)
    p = _write(tmp_path, "d.adj", "AdjacencyHypergraph\n1\n1\n1\n1\n0\n5\n0\n0\n")
    assert _status(lambda: pb.Graph.read(p)) == "PBF_ERR_RANGE"
    # unsorted adjacency (checked here; the reference assumes it)
    p = _write(tmp_path, "e.adj", "AdjacencyHypergraph\n1\n2\n2\n2\n0\n1\n0\n0\n1\n0\n0\n")
    assert _status(lambda: pb.Graph.read(p)) == "PBF_ERR_UNSORTED"
    # CSRs that disagree
    p = _write(tmp_path, "f.adj", "AdjacencyHypergraph\n1\n1\n2\n1\n0\n1\n0\n1\n0\n")
    assert _status(lambda: pb.Graph.read(p)) == "PBF_ERR_FORMAT"
    # missing file
    assert _status(lambda: pb.Graph.read(str(tmp_path / "none.adj"))) == "PBF_ERR_IO"
    # a valid one-edge graph reads
    p = _write(tmp_path, "g.adj", "AdjacencyHypergraph\n1\n1\n1\n1\n0\n0\n0\n0\n")
    g = pb.Graph.read(p)
    assert (g.nv, g.nu, g.m) == (1, 1, 1)


def _digest(g):
    return tuple(int(np.bitwise_xor.reduce(a.astype(np.uint64) * np.uint64(2654435761) +
                                           np.arange(a.size, dtype=np.uint64))) if a.size else 0
                 for a in (g.offsetsV, g.edgesV, g.offsetsU, g.edgesU))


def test_generators_deterministic_and_shaped():
    a = pb.Graph.uniform(1000, 1000, 10000, 1)
    b = pb.Graph.uniform(1000, 1000, 10000, 1)
    assert _digest(a) == _digest(b) and a.m == 10000 and a.validate() == 0
    c = pb.Graph.uniform(1000, 1000, 10000, 2)
    assert _digest(a) != _digest(c)
    cl = pb.Graph.chung_lu(1 << 12, 1 << 10, 1 << 15, 2.1, 2.1, 2)
    assert cl.m == 1 << 15 and cl.validate() == 0
    dv = np.diff(cl.offsetsV)
    assert dv[:8].mean() > 10 * dv.mean()  # low ids are the hubs
    r = pb.Graph.rmat(10, 10, 8000, 0.5, 0.1, 0.1, 0.3, 3)
    assert r.m == 8000 and r.validate() == 0
    # the first-m-distinct rule: the same prefix of draws for a larger m
    r2 = pb.Graph.rmat(10, 10, 8000, 0.5, 0.1, 0.1, 0.3, 3)
    assert _digest(r) == _digest(r2)


def test_batch_params_distribution():
    ps = [pb.Graph.batch_params(1000 + i) for i in range(200)]
    nv = np.array([p[0] for p in ps]); nu = np.array([p[1] for p in ps]); m = np.array([p[2] for p in ps])
    gam = np.array([p[3] for p in ps])
    assert nv.min() >= 64 and nv.max() <= 65536 and nu.min() >= 64 and nu.max() <= 65536
    assert (gam >= 2.0).all() and (gam <= 3.0).all()
    avg = 2 * m / (nv + nu)
    assert avg.min() >= 1.9 and avg.max() <= 32.5
    g = pb.Graph.batch_graph(1000)
    assert (g.nv, g.nu, g.m) == ps[0][:3]


def test_side_host_matches_definition():
    g = pb.Graph.chung_lu(500, 300, 4000, 2.1, 2.1, 3)
    dv, du = np.diff(g.offsetsV).astype(np.int64), np.diff(g.offsetsU).astype(np.int64)
    wv, wu = int((dv * (dv - 1) // 2).sum()), int((du * (du - 1) // 2).sum())
    assert g.side() == (wv <= wu, min(wv, wu))


def test_validate_error_codes():
    """pbf_validate on well-formed and broken CSR pairs; each broken case holds one fault, placed
    at several positions so that every thread's share of the parallel checks meets one."""
    g = pb.Graph.uniform(3000, 2000, 40000, 11)
    parts = lambda: [g.offsetsV.copy(), g.edgesV.copy(), g.offsetsU.copy(), g.edgesU.copy()]
    assert g.validate() == 0
    for pos in (0, 7919, 20000, 39999):
        # a V entry replaced by a neighbour absent from the list: the U lists are no transpose
        oV, eV, oU, eU = parts()
        v = int(np.searchsorted(oV, pos, side="right") - 1)
        lst = set(eV[oV[v]:oV[v + 1]].tolist())
        lo = int(eV[pos - 1]) + 1 if pos > oV[v] else 0
        hi = int(eV[pos + 1]) if pos + 1 < oV[v + 1] else g.nu
        free = [u for u in range(lo, hi) if u not in lst]
        if free:
            eV[pos] = free[0]
            assert pb.Graph(g.nv, g.nu, oV, eV, oU, eU).validate() == 2, pos  # PBF_ERR_FORMAT
        # out-of-range neighbour
        oV, eV, oU, eU = parts()
        eU[pos % g.m] = g.nv
        assert pb.Graph(g.nv, g.nu, oV, eV, oU, eU).validate() == 3, pos  # PBF_ERR_RANGE
        # two neighbours of one list swapped
        oV, eV, oU, eU = parts()
        x = int(np.searchsorted(oU, pos, side="right") - 1)
        if oU[x + 1] - oU[x] >= 2:
            a = int(oU[x])
            eU[a], eU[a + 1] = eU[a + 1], eU[a]
            assert pb.Graph(g.nv, g.nu, oV, eV, oU, eU).validate() == 4, pos  # PBF_ERR_UNSORTED
    # decreasing offsets, and an offset array that does not end at m
    oV, eV, oU, eU = parts()
    oV[1500] = oV[1501] + 1
    assert pb.Graph(g.nv, g.nu, oV, eV, oU, eU).validate() == 4
    oV, eV, oU, eU = parts()
    oU[-1] -= 1
    assert pb.Graph(g.nv, g.nu, oV, eV, oU, eU).validate() == 2


KONECT_CASES = ["konect_small", "konect_weighted", "konect_chunglu"]


@pytest.mark.parametrize("name", KONECT_CASES)
def test_konect_matches_reference(name, tmp_path):
    # the graph and the AdjacencyHypergraph text the reference's KONECTtoHyperAdj wrote for it
    import subprocess
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = pb.Graph.read_konect(os.path.join(gold, name + ".tsv"))
    ref = pb.Graph.read(os.path.join(gold, name + ".adj"))
    assert (g.nv, g.nu, g.m) == (ref.nv, ref.nu, ref.m)
    for k in ("offsetsV", "edgesV", "offsetsU", "edgesU"):
        assert np.array_equal(getattr(g, k), getattr(ref, k)), k
    conv = os.path.join(os.path.dirname(pb.library_path()), "bin", "konect_to_hyperadj")
    out = tmp_path / "out.adj"
    subprocess.run([conv, os.path.join(gold, name + ".tsv"), str(out)], check=True)
    with open(os.path.join(gold, name + ".adj"), "rb") as f:
        assert out.read_bytes() == f.read()


def test_konect_errors(tmp_path):
    cases = {"one column": ("% x\n1\n2\n3\n", "PBF_ERR_FORMAT"), "empty": ("% only comments\n", "PBF_ERR_FORMAT"),
             "negative id": ("1 2\n-3 4\n", "PBF_ERR_RANGE"), "id beyond 32 bits": ("1 4294967296\n", "PBF_ERR_RANGE")}
    for tag, (text, err) in cases.items():
        p = tmp_path / (tag.replace(" ", "_") + ".tsv")
        p.write_text(text)
        with pytest.raises(pb.PbfError) as e:
            pb.Graph.read_konect(str(p))
        assert e.value.name == err, tag
    with pytest.raises(pb.PbfError) as e:
        pb.Graph.read_konect(str(tmp_path / "missing.tsv"))
    assert e.value.name == "PBF_ERR_IO"
    # a partial last row is dropped, "+" signs and leading zeros are read as atol reads them
    p = tmp_path / "partial.tsv"
    p.write_text("%c\n+7 005 1\n3 5 2\n7\n")
    g = pb.Graph.read_konect(str(p))
    assert (g.nv, g.nu, g.m) == (2, 1, 2)
