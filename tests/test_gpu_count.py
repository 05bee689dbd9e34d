"""GPU parity of counting (VERT / EDGE / TOTAL) and ranking against the CPU oracle.

Every comparison is bit-exact in uint64.  The oracle (oracle/oracle.c) is itself pinned to the
compiled reference and to brute force by tests/test_oracle.py.
"""
import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O
from tests.graphs import complete_bipartite, empty_graph, hbm_graph, small_graphs, wide_centre_graph

pytestmark = pytest.mark.gpu

RANKS = ["ADEG", "ACOCORE"]


def _check_counts(ctx, g, tag, rank_type):
    use_v, _ = ctx.side(g)
    assert (use_v, _) == O.side(g), tag
    v = ctx.count(g, use_v, "VERT", "GPU", rank_type)
    ov = O.count_vert(g, use_v, rank_type)
    assert np.array_equal(pb.as_u64(v), ov), f"{tag} {rank_type} VERT"
    e = ctx.count(g, use_v, "EDGE", "GPU", rank_type)
    oe = O.count_edge(g, use_v, rank_type)
    assert np.array_equal(pb.as_u64(e), oe), f"{tag} {rank_type} EDGE"
    t = ctx.count(g, use_v, "TOTAL", "GPU", rank_type)
    assert t == O.count_total(g, use_v, rank_type), f"{tag} {rank_type} TOTAL"
    assert int(ov.sum()) == 2 * t and int(oe.sum()) == 4 * t


@pytest.mark.parametrize("rank_type", RANKS)
def test_small_graphs(ctx, rank_type):
    for tag, g in small_graphs():
        print("case", tag, rank_type, flush=True)
        _check_counts(ctx, g, tag, rank_type)


@pytest.mark.parametrize("rank_type", RANKS)
def test_config1_uniform(ctx, rank_type):
    g = pb.Graph.uniform(1000, 1000, 10000, 1)
    _check_counts(ctx, g, "config1", rank_type)


@pytest.mark.parametrize("rank_type", RANKS)
def test_chung_lu_medium(ctx, rank_type):
    # 2^16 x 2^14, 2^20 edges, gamma 2.1: hubs with thousands of neighbours
    g = pb.Graph.chung_lu(1 << 16, 1 << 14, 1 << 20, 2.1, 2.1, 2)
    use_v, _ = ctx.side(g)
    v = pb.as_u64(ctx.count(g, use_v, "VERT", "GPU", rank_type))
    assert np.array_equal(v, O.count_vert(g, use_v, rank_type))
    t = ctx.count(g, use_v, "TOTAL", "GPU", rank_type)
    assert 2 * t == int(v.sum())


def test_chung_lu_split_hubs(ctx):
    # hubs above 32768 neighbours take the multi-workgroup (split) path; checked on every mode
    g = pb.Graph.chung_lu(1 << 18, 1 << 16, 1 << 22, 2.1, 2.1, 5)
    assert max(np.diff(g.offsetsV).max(), np.diff(g.offsetsU).max()) > 32768
    use_v, _ = ctx.side(g)
    v = pb.as_u64(ctx.count(g, use_v, "VERT"))
    assert np.array_equal(v, O.count_vert(g, use_v))
    e = pb.as_u64(ctx.count(g, use_v, "EDGE"))
    assert np.array_equal(e, O.count_edge(g, use_v))
    assert ctx.count(g, use_v, "TOTAL") == int(v.sum()) // 2


def test_rmat_medium_edge(ctx):
    g = pb.Graph.rmat(14, 14, 200000, 0.5, 0.1, 0.1, 0.3, 3)
    use_v, _ = ctx.side(g)
    e = pb.as_u64(ctx.count(g, use_v, "EDGE", "GPU", "ACOCORE"))
    assert np.array_equal(e, O.count_edge(g, use_v, "ACOCORE"))


def test_device_memory_path(ctx):
    g = pb.Graph.chung_lu(3000, 2000, 30000, 2.3, 2.2, 5)
    dg = g.to_device(ctx)
    use_v, w = ctx.side(dg)
    assert (use_v, w) == g.side()
    v = ctx.count(dg, use_v, "VERT")
    assert np.array_equal(pb.as_u64(v), O.count_vert(g, use_v))
    e = ctx.count(dg, use_v, "EDGE")
    assert np.array_equal(pb.as_u64(e), O.count_edge(g, use_v))
    assert ctx.count(dg, use_v, "TOTAL") == O.count_total(g, use_v)


def test_closed_forms(ctx):
    for a, b in [(2, 2), (3, 4), (5, 3), (7, 7)]:
        g = complete_bipartite(a, b)
        use_v, _ = ctx.side(g)
        tot = ctx.count(g, use_v, "TOTAL")
        assert tot == (a * (a - 1) // 2) * (b * (b - 1) // 2)
        v = pb.as_u64(ctx.count(g, use_v, "VERT"))
        # S = U (size b) when use_v; each S vertex lies in (other-S-count choose 1) * C(|O|,2)
        ns, no = (b, a) if use_v else (a, b)
        assert (v == (ns - 1) * (no * (no - 1) // 2)).all()
        e = pb.as_u64(ctx.count(g, use_v, "EDGE"))
        assert (e == (a - 1) * (b - 1)).all()


def test_empty_and_isolated(ctx):
    g = empty_graph(5, 3)
    use_v, _ = ctx.side(g)
    assert ctx.count(g, use_v, "TOTAL") == 0
    assert (pb.as_u64(ctx.count(g, use_v, "VERT")) == 0).all()
    # star plus isolated vertices: no butterflies
    v = np.zeros(6, np.uint32)
    u = np.arange(6, dtype=np.uint32)
    g = pb.Graph.from_edges(4, 9, v, u)
    use_v, _ = ctx.side(g)
    assert ctx.count(g, use_v, "TOTAL") == 0
    assert (pb.as_u64(ctx.count(g, use_v, "EDGE")) == 0).all()


def test_rank_matches_oracle(ctx):
    for rank_type in RANKS:
        g = pb.Graph.chung_lu(5000, 3000, 40000, 2.1, 2.1, 11)
        order, rank_of, w = ctx.rank(g, rank_type)
        oo, orr, _ = O.rank(g, rank_type)
        assert np.array_equal(order, oo), rank_type
        assert np.array_equal(rank_of, orr), rank_type
        assert w == O.rank_wedges(g, orr), rank_type


def test_edge_to_idxs(ctx):
    g = pb.Graph.chung_lu(2000, 1500, 20000, 2.1, 2.1, 4)
    for use_v in (True, False):
        assert np.array_equal(ctx.edge_to_idxs(g, use_v), O.edge_to_idxs(g, use_v))


def test_deterministic_repeats(ctx):
    g = pb.Graph.chung_lu(1 << 14, 1 << 12, 1 << 18, 2.1, 2.1, 9)
    use_v, _ = ctx.side(g)
    a = pb.as_u64(ctx.count(g, use_v, "VERT"))
    for _ in range(3):
        assert np.array_equal(a, pb.as_u64(ctx.count(g, use_v, "VERT")))
    ea = pb.as_u64(ctx.count(g, use_v, "EDGE"))
    assert np.array_equal(ea, pb.as_u64(ctx.count(g, use_v, "EDGE")))


def test_side_ranking(ctx):
    # SIDE: count side first, then O, each by id; counts equal every other ranking's
    for tag, g in small_graphs() + [("cl", pb.Graph.chung_lu(3000, 2000, 30000, 2.2, 2.1, 8))]:
        use_v, _ = ctx.side(g)
        v = pb.as_u64(ctx.count(g, use_v, "VERT", "GPU", "SIDE"))
        assert np.array_equal(v, O.count_vert(g, use_v)), tag
        e = pb.as_u64(ctx.count(g, use_v, "EDGE", "GPU", "SIDE"))
        assert np.array_equal(e, O.count_edge(g, use_v)), tag
        assert ctx.count(g, use_v, "TOTAL", "GPU", "SIDE") == O.count_total(g, use_v), tag
        order, rank_of, w = ctx.rank(g, "SIDE")
        ids_s = np.arange(g.nv, g.nv + g.nu) if use_v else np.arange(g.nv)
        ids_o = np.arange(g.nv) if use_v else np.arange(g.nv, g.nv + g.nu)
        assert np.array_equal(order, np.concatenate([ids_s, ids_o]).astype(np.uint32)), tag
        # W under SIDE = wedges centred on O with both endpoints on S
        offs = g.offsetsV if use_v else g.offsetsU
        d = np.diff(offs).astype(np.int64)
        assert w == int((d * (d - 1) // 2).sum()), tag


def test_hbm_fallback_reuse(ctx):
    # the HBM path must leave its pool slice zero, used lists included: a later call with
    # another vertex count lays its counters over that storage
    big, small = hbm_graph(45000, 5000), hbm_graph(10000, 10000)
    for g in (big, small):
        use_v, _ = ctx.side(g)
        assert use_v
    res = []
    for g in (big, small, big, small, big):
        v = pb.as_u64(ctx.count(g, True, "VERT", "GPU", "SIDE"))
        assert ctx.timing()["hub_vertices"] >= 1
        res.append(v)
    ob, os_ = O.count_vert(big, True), O.count_vert(small, True)
    for k, v in enumerate(res):
        assert np.array_equal(v, ob if k % 2 == 0 else os_), k


def test_side_wraps_as_reference(ctx):
    # cmpWedgeCounts' C(d,2) wraps in 32 bits for a degree above 65536; on this graph that
    # flips the side choice, and the reference's choice and counts are in side_cases.npz
    import os
    from tests.graphs import hub_wrap_graph
    with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "side_cases.npz")) as z:
        rec = {k: z[k] for k in z.files}
    g = hub_wrap_graph()
    dg = g.to_device(ctx)
    exp = (bool(rec["hub_wrap__use_v"][0]), int(rec["hub_wrap__side_wedges"][0]))
    assert ctx.side(g) == ctx.side(dg) == exp
    use_v = exp[0]
    for rt in RANKS:
        assert np.array_equal(pb.as_u64(ctx.count(dg, use_v, "VERT", "GPU", rt)), rec["hub_wrap__vert"].astype(np.uint64))
        assert np.array_equal(pb.as_u64(ctx.count(dg, use_v, "EDGE", "GPU", rt)), rec["hub_wrap__edge"].astype(np.uint64))
        assert ctx.count(dg, use_v, "TOTAL", "GPU", rt) == int(rec["hub_wrap__total"][0])
    # a batch takes each graph's side by the same rule
    vert, tot = ctx.count_batch([g, complete_bipartite(3, 4)], "ADEG", mode="union")
    assert np.array_equal(vert[0], rec["hub_wrap__vert"].astype(np.uint64)) and int(tot[0]) == int(rec["hub_wrap__total"][0])
    h = pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)
    k = "chunglu_2p20x2p18_m2p24_g21_s2"
    assert ctx.side(h.to_device(ctx)) == (bool(rec[k + "__use_v"][0]), int(rec[k + "__side_wedges"][0]))


def test_edge_wide_centres(ctx):
    # per-edge pass 2 credits a chunk in 64 bits when one of its centres could pass 2^32 in a
    # 32-bit slot counter; here every chunk through the U hub does (closed form: edge (0, 0) in
    # k butterflies, every other edge in one)
    k = 69999
    g = wide_centre_graph(k)
    for rt in RANKS:
        use_v, _ = ctx.side(g)
        e = pb.as_u64(ctx.count(g, use_v, "EDGE", "GPU", rt))
        assert e.shape == (3 * k + 1,)
        assert int(e.max()) == k and int((e == k).sum()) == 1 and int((e == 1).sum()) == 3 * k, rt
        assert np.array_equal(e, O.count_edge(g, use_v, rt)), rt
        assert ctx.count(g, use_v, "TOTAL", "GPU", rt) == k
