"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the GPU's checker.

* golden vectors written by the compiled reference (tests/golden/make_golden.py) -- counts from
  the reference's SERIAL and BATCHS modes, tip / wing numbers from PBATCHS, the reference's own
  ADEG / ACOCORE orders and wedge counts;
* brute force from the definitions (pairwise common neighbours; textbook one-at-a-time peels);
* closed forms (K_{a,b}, stars) and invariants (sum VERT = 2 total, sum EDGE = 4 total,
  ranking invariance).
A planted-mistake test shows that these checks catch a single off-by-one in a count.
"""
import glob
import os

import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O
from tests.graphs import complete_bipartite, small_graphs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
               if os.path.exists(p[:-4] + ".adj"))


def _load(name):
    g = pb.Graph.read(os.path.join(GOLDEN, name + ".adj"))
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        rec = {k: z[k] for k in z.files}
    return g, rec


def test_golden_present():
    assert len(CASES) >= 5


@pytest.mark.parametrize("name", CASES)
def test_oracle_counts_match_reference(name):
    g, rec = _load(name)
    use_v, sw = O.side(g)
    assert use_v == bool(rec["use_v"][0]) and sw == int(rec["side_wedges"][0])
    for rt in ("ADEG", "ACOCORE"):
        v = O.count_vert(g, use_v, rt)
        e = O.count_edge(g, use_v, rt)
        t = O.count_total(g, use_v, rt)
        for ct in ("SERIAL", "BATCHS"):
            assert np.array_equal(v, rec[f"vert_{ct}_{rt}"].astype(np.uint64)), (ct, rt)
            assert np.array_equal(e, rec[f"edge_{ct}_{rt}"].astype(np.uint64)), (ct, rt)
            assert t == int(rec[f"total_{ct}_{rt}"][0]), (ct, rt)


@pytest.mark.parametrize("name", CASES)
def test_oracle_peels_match_reference(name):
    g, rec = _load(name)
    use_v, _ = O.side(g)
    v = rec["vert_SERIAL_ADEG"].astype(np.uint64)
    assert np.array_equal(O.peel_vert(g, use_v, v), rec["tip_PBATCHS"].astype(np.uint64))
    e = rec["edge_SERIAL_ADEG"].astype(np.uint64)
    assert np.array_equal(O.peel_edge(g, use_v, e), rec["wing_PBATCHS"].astype(np.uint64))


@pytest.mark.parametrize("name", CASES)
def test_oracle_ranking_against_reference_order(name):
    g, rec = _load(name)
    n = g.nv + g.nu
    deg = np.concatenate([np.diff(g.offsetsV), np.diff(g.offsetsU)]).astype(np.int64)
    for rt in ("ADEG", "ACOCORE"):
        ref_order = rec[f"order_{rt}"].astype(np.int64)
        assert sorted(ref_order.tolist()) == list(range(n))
        ref_rank = np.empty(n, np.uint32)
        ref_rank[ref_order] = np.arange(n, dtype=np.uint32)
        # rankWedgeF restated: W under the reference's own order equals the reference's W
        assert O.rank_wedges(g, ref_rank) == int(rec[f"W_{rt}"][0]), rt
        # the key is non-increasing along the reference's order
        if rt == "ADEG":
            key = np.array([0 if d == 0 else int(d).bit_length() for d in deg])
        else:
            key = O.rank(g, "ACOCORE")[2].astype(np.int64)
        ks = key[ref_order]
        assert (np.diff(ks) <= 0).all(), rt
        # our order is the same sort with ties by ascending id
        order, rank_of, _ = O.rank(g, rt)
        assert np.array_equal(key[order], ks)
        assert (np.diff(key[order]) <= 0).all()


@pytest.mark.parametrize("case", range(len(small_graphs())))
def test_oracle_against_brute_force(case):
    tag, g = small_graphs()[case]
    use_v, _ = O.side(g)
    bv, tot = O.brute_vert(g, use_v)
    be = O.brute_edge(g, use_v)
    for rt in ("ADEG", "ACOCORE"):
        assert np.array_equal(O.count_vert(g, use_v, rt), bv), tag
        assert np.array_equal(O.count_edge(g, use_v, rt), be), tag
        assert O.count_total(g, use_v, rt) == tot, tag
    assert int(bv.sum()) == 2 * tot and int(be.sum()) == 4 * tot
    assert np.array_equal(O.peel_vert(g, use_v, bv), O.brute_peel_vert(g, use_v)), tag
    if g.m <= 250:
        assert np.array_equal(O.peel_edge(g, use_v, be), O.brute_peel_edge(g, use_v)), tag


def test_closed_forms():
    for a, b in [(2, 2), (2, 5), (4, 3), (6, 6)]:
        g = complete_bipartite(a, b)
        use_v, _ = O.side(g)
        tot = (a * (a - 1) // 2) * (b * (b - 1) // 2)
        assert O.count_total(g, use_v) == tot
        ns, no = (b, a) if use_v else (a, b)
        assert (O.count_vert(g, use_v) == (ns - 1) * (no * (no - 1) // 2)).all()
        assert (O.count_edge(g, use_v) == (a - 1) * (b - 1)).all()
        assert (O.peel_vert(g, use_v, O.count_vert(g, use_v)) == (ns - 1) * (no * (no - 1) // 2)).all()
        assert (O.peel_edge(g, use_v, O.count_edge(g, use_v)) == (a - 1) * (b - 1)).all()
    star = pb.Graph.from_edges(1, 9, np.zeros(9, np.uint32), np.arange(9, dtype=np.uint32))
    use_v, _ = O.side(star)
    assert O.count_total(star, use_v) == 0


def test_side_closed_form():
    # wV = nv * C(b, 2) vs wU = nu * C(a, 2) on K_{a,b}: a=3 (V), b=5 (U): wV = 30, wU = 15
    g = complete_bipartite(3, 5)
    assert O.side(g) == (False, 15)
    g = complete_bipartite(5, 3)
    assert O.side(g) == (True, 15)
    g = complete_bipartite(4, 4)  # tie -> counts on U (use_v true)
    assert O.side(g) == (True, 4 * 6)


def test_edge_to_idxs_is_permutation():
    g = pb.Graph.chung_lu(500, 300, 4000, 2.1, 2.1, 3)
    for use_v in (True, False):
        eti = O.edge_to_idxs(g, use_v)
        assert sorted(eti.tolist()) == list(range(g.m))
        # eti maps the S-side position of (s, o) to the O-side position of the same edge
        offS, edgS = (g.offsetsU, g.edgesU) if use_v else (g.offsetsV, g.edgesV)
        edgO = g.edgesV if use_v else g.edgesU
        owner = np.repeat(np.arange(len(offS) - 1), np.diff(offS))
        assert np.array_equal(edgO[eti], owner.astype(np.uint32))


def test_planted_mistake_is_caught():
    """A single wrong count (as a dropped / extra term would give) fails the checks above."""
    g = pb.Graph.chung_lu(60, 40, 200, 2.1, 2.1, 1)
    use_v, _ = O.side(g)
    bv, tot = O.brute_vert(g, use_v)
    bad = O.count_vert(g, use_v).copy()
    idx = int(np.argmax(bad))
    bad[idx] -= 1
    assert not np.array_equal(bad, bv)
    assert int(bad.sum()) != 2 * tot
    bad_e = O.count_edge(g, use_v).copy()
    bad_e[0] += 1
    assert int(bad_e.sum()) != 4 * tot


def _side_cases():
    with np.load(os.path.join(GOLDEN, "side_cases.npz")) as z:
        return {k: z[k] for k in z.files}


def test_oracle_side_wraps_as_reference():
    # cmpWedgeCounts takes C(d,2) in the 32-bit degree type: the reference's side choice and
    # num_wedges on a graph where that wraps and flips the choice, and on the headline graph
    from tests.graphs import hub_wrap_graph
    rec = _side_cases()
    g = hub_wrap_graph()
    assert O.side(g) == (bool(rec["hub_wrap__use_v"][0]), int(rec["hub_wrap__side_wedges"][0])) == (True, 33767)
    assert g.side() == O.side(g)  # pbf_side_host
    use_v = True
    v = O.count_vert(g, use_v, "ADEG")
    assert np.array_equal(v, rec["hub_wrap__vert"].astype(np.uint64))
    assert np.array_equal(O.count_edge(g, use_v, "ADEG"), rec["hub_wrap__edge"].astype(np.uint64))
    assert O.count_total(g, use_v, "ADEG") == int(rec["hub_wrap__total"][0]) == 1000 * 999 // 2
    h = pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)
    k = "chunglu_2p20x2p18_m2p24_g21_s2"
    assert O.side(h) == h.side() == (bool(rec[k + "__use_v"][0]), int(rec[k + "__side_wedges"][0]))


def _level_minimum(g, use_v, tip, k):
    """Least butterfly count, in the subgraph induced by the S vertices with tip >= k and all of
    O, over those S vertices (the oracle's count of that subgraph)."""
    offS, edgS = (g.offsetsU, g.edgesU) if use_v else (g.offsetsV, g.edgesV)
    owner = np.repeat(np.arange(len(offS) - 1, dtype=np.uint32), np.diff(offS.astype(np.int64)))
    keep = tip[owner] >= k
    s_ids, o_ids = owner[keep], edgS[keep].astype(np.uint32)
    vv, uu = (o_ids, s_ids) if use_v else (s_ids, o_ids)
    h = pb.Graph.from_edges(g.nv, g.nu, np.ascontiguousarray(vv), np.ascontiguousarray(uu))
    return int(O.count_vert(h, use_v)[tip >= k].min())


def test_tip_levels_are_exact_minima():
    # when the peel reaches level k, the live S vertices are those with tip >= k and each one's D
    # is its butterfly count among them, so k is the least such count at every distinct level
    # (the property the full-size GPU test checks).  A tip raised by one breaks it: the vertex
    # would join a subgraph in which, by the maximality of tips, some vertex has fewer
    g = pb.Graph.chung_lu(1500, 800, 10000, 2.1, 2.1, 4)
    use_v, _ = O.side(g)
    tip = O.peel_vert(g, use_v, O.count_vert(g, use_v))
    levels = np.unique(tip)
    assert len(levels) > 100
    for k in levels:
        assert _level_minimum(g, use_v, tip, int(k)) == int(k), int(k)
    bad = tip.copy()
    j = int(np.argmin(tip))
    bad[j] += 1
    assert _level_minimum(g, use_v, bad, int(bad[j])) != int(bad[j])


def test_wing_levels_are_exact_minima():
    # the edge-peel counterpart: at every distinct level k, the edges with wing >= k hold at least
    # k butterflies each among themselves and some edge exactly k; a wing raised by one breaks it
    # (by the maximality of wings)
    g = pb.Graph.chung_lu(800, 600, 8000, 2.1, 2.1, 8)
    use_v, _ = O.side(g)
    wing = O.peel_edge(g, use_v, O.count_edge(g, use_v))
    offO, edgO = (g.offsetsV, g.edgesV) if use_v else (g.offsetsU, g.edgesU)
    owner = np.repeat(np.arange(len(offO) - 1, dtype=np.uint32), np.diff(offO.astype(np.int64)))
    vv, uu = (owner, edgO) if use_v else (edgO, owner)

    def least(w, k):
        sel = w >= k
        h = pb.Graph.from_edges(g.nv, g.nu, np.ascontiguousarray(vv[sel]), np.ascontiguousarray(uu[sel]))
        return int(O.count_edge(h, O.side(h)[0]).min())

    levels = np.unique(wing)
    assert len(levels) > 50
    for k in levels:
        assert least(wing, int(k)) == int(k), int(k)
    bad = wing.copy()
    j = int(np.argmin(wing))
    bad[j] += 1
    assert least(bad, int(bad[j])) != int(bad[j])
