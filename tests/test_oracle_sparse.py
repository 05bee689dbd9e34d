"""The oracle's sparsification (eSparseBipartite / clrSparseBipartite + delZeroDeg,
apps/hyper/butterfly_utils.h:744-996) and the driver's scaling (BipartiteButterflies.C:213-214),
pinned against the graphs and totals the compiled reference wrote into
tests/golden/sparse_cases.npz, and against what the definitions fix: one colour keeps every
edge, colour sparsification keeps exactly the edges whose ends share a colour, and the renumbering
is the order-preserving one."""
import os

import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cases():
    with np.load(os.path.join(GOLDEN, "sparse_cases.npz")) as z:
        keys = sorted({k.rsplit("__", 1)[0] for k in z.files})
        return [(k, z[k + "__sizes"], z[k + "__csr"], z[k + "__totals"]) for k in keys]


CASES = _cases()


def _split(sizes, csr):
    nv, nu, m = (int(x) for x in sizes)
    a = np.asarray(csr, np.uint32)
    return O.CSR(nv, nu, a[:nv + 1], a[nv + 1:nv + 1 + m], a[nv + 1 + m:nv + 2 + m + nu], a[nv + 2 + m + nu:])


def _same(h, e):
    return ((h.nv, h.nu, h.m) == (e.nv, e.nu, e.m) and np.array_equal(h.offsetsV, e.offsetsV)
            and np.array_equal(h.edgesV, e.edgesV) and np.array_equal(h.offsetsU, e.offsetsU)
            and np.array_equal(h.edgesU, e.edgesU))


def test_sparse_golden_present():
    assert len(CASES) == 40


@pytest.mark.parametrize("key,sizes,csr,totals", CASES, ids=[c[0] for c in CASES])
def test_oracle_sparsify_matches_reference(key, sizes, csr, totals):
    name, case = key.split("__")
    st, d, seed = case.split("_")
    g = pb.Graph.read(os.path.join(GOLDEN, name + ".adj"))
    h = O.sparsify(g, st, int(d), int(seed))
    assert _same(h, _split(sizes, csr))
    approx, exact = O.approx_total(g, st, int(d), int(seed))
    assert (exact, approx) == (int(totals[0]), int(totals[1]))


def _edges(g):
    v = np.repeat(np.arange(g.nv, dtype=np.int64), np.diff(g.offsetsV.astype(np.int64)))
    return v, g.edgesV.astype(np.int64)


def _drop_isolated(nv, nu, v, u):
    """delZeroDeg of an edge list: ids renumbered in order over the vertices that keep an edge."""
    kv, ku = np.unique(v), np.unique(u)
    rv = np.full(nv, -1, np.int64); rv[kv] = np.arange(len(kv))
    ru = np.full(nu, -1, np.int64); ru[ku] = np.arange(len(ku))
    return pb.Graph.from_edges(len(kv), len(ku), rv[v], ru[u])


@pytest.mark.parametrize("st", ["EDGE", "COLOR"])
def test_one_colour_keeps_every_edge(st):
    # chung-lu graphs have isolated vertices, which delZeroDeg drops
    g = pb.Graph.chung_lu(400, 300, 3000, 2.1, 2.1, 4)
    v, u = _edges(g)
    want = _drop_isolated(g.nv, g.nu, v, u)
    assert want.nv < g.nv or want.nu < g.nu
    assert _same(O.sparsify(g, st, 1, 9), want)
    assert O.approx_total(g, st, 1, 9)[0] == O.count_total(g, g.side()[0])


def test_colour_rule_and_edge_rate():
    g = pb.Graph.uniform(3000, 2000, 60000, 3)
    v, u = _edges(g)
    d, seed = 5, 11
    # COLOR: exactly the edges whose ends have equal hashInt(seed + id) % d
    cv = np.array([O.hash64(seed + x) % d for x in range(g.nv)])
    cu = np.array([O.hash64(seed + x) % d for x in range(g.nu)])
    keep = cv[v] == cu[u]
    assert _same(O.sparsify(g, "COLOR", d, seed), _drop_isolated(g.nv, g.nu, v[keep], u[keep]))
    # EDGE: about m / d edges stay (the mix is not exactly uniform: within 10 %), and the V and
    # U sides agree (checked inside sparsify)
    h = O.sparsify(g, "EDGE", d, seed)
    assert abs(h.m - g.m / d) < 0.1 * g.m / d
    # a different seed keeps a different set
    assert not _same(O.sparsify(g, "EDGE", d, seed + 1), h)

