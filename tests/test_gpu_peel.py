"""GPU parity of tip (vertex) and wing (edge) peeling against the CPU oracle, bit-exact."""
import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O
from tests.graphs import complete_bipartite, small_graphs, two_hubs_with_leaves

pytestmark = pytest.mark.gpu


def test_peel_small(ctx):
    for tag, g in small_graphs():
        print("case", tag, flush=True)
        use_v, _ = ctx.side(g)
        v = O.count_vert(g, use_v)
        tip = pb.as_u64(ctx.peel(g, use_v, v))
        assert np.array_equal(tip, O.peel_vert(g, use_v, v)), tag
        e = O.count_edge(g, use_v)
        wing = pb.as_u64(ctx.peel_edge(g, use_v, e))
        assert np.array_equal(wing, O.peel_edge(g, use_v, e)), tag


def test_peel_closed_form(ctx):
    for a, b in [(3, 4), (6, 5)]:
        g = complete_bipartite(a, b)
        use_v, _ = ctx.side(g)
        ns, no = (b, a) if use_v else (a, b)
        v = pb.as_u64(ctx.count(g, use_v, "VERT"))
        tip = pb.as_u64(ctx.peel(g, use_v, v))
        assert (tip == (ns - 1) * (no * (no - 1) // 2)).all()
        e = pb.as_u64(ctx.count(g, use_v, "EDGE"))
        wing = pb.as_u64(ctx.peel_edge(g, use_v, e))
        assert (wing == (a - 1) * (b - 1)).all()


@pytest.mark.parametrize("hub,leaves", [(300, 2), (3000, 2)])
def test_wing_walks_unstaged_lists(ctx, hub, leaves):
    # two V hubs on the same `hub` U vertices, each U vertex with `leaves` private V leaves: a
    # hub edge is walked from its U end (the smaller two-hop work), so N(b) is a hub's list, too
    # long to stage in LDS for the wavefront walks (hub 300 > 256) or the workgroup walks
    # (hub 3000 > 2048); closed form: every hub edge lies in hub - 1 butterflies, a leaf edge in none
    g = two_hubs_with_leaves(hub, leaves)
    use_v, _ = ctx.side(g)
    e = pb.as_u64(ctx.count(g, use_v, "EDGE"))
    wing = pb.as_u64(ctx.peel_edge(g, use_v, e))
    assert np.array_equal(wing, O.peel_edge(g, use_v, e))
    assert sorted(set(wing.tolist())) == [0, hub - 1] and int((wing == hub - 1).sum()) == 2 * hub
    v = pb.as_u64(ctx.count(g, use_v, "VERT"))
    tip = pb.as_u64(ctx.peel(g, use_v, v))
    assert np.array_equal(tip, O.peel_vert(g, use_v, v))


def test_peel_medium(ctx):
    g = pb.Graph.rmat(11, 11, 30000, 0.5, 0.1, 0.1, 0.3, 3)
    use_v, _ = ctx.side(g)
    v = pb.as_u64(ctx.count(g, use_v, "VERT", "GPU", "ACOCORE"))
    tip = pb.as_u64(ctx.peel(g, use_v, v))
    assert np.array_equal(tip, O.peel_vert(g, use_v, v))
    e = pb.as_u64(ctx.count(g, use_v, "EDGE", "GPU", "ACOCORE"))
    wing = pb.as_u64(ctx.peel_edge(g, use_v, e))
    assert np.array_equal(wing, O.peel_edge(g, use_v, e))


def test_peel_chung_lu_hubs(ctx):
    # hubs with large two-hop neighbourhoods exercise the workgroup and HBM-histogram walks
    g = pb.Graph.chung_lu(4000, 3000, 60000, 2.1, 2.1, 8)
    use_v, _ = ctx.side(g)
    v = pb.as_u64(ctx.count(g, use_v, "VERT"))
    tip = pb.as_u64(ctx.peel(g, use_v, v))
    assert np.array_equal(tip, O.peel_vert(g, use_v, v))
    # about 800 wing levels: the edge peel compacts its adjacency several times on the way
    e = pb.as_u64(ctx.count(g, use_v, "EDGE"))
    wing = pb.as_u64(ctx.peel_edge(g, use_v, e))
    assert np.array_equal(wing, O.peel_edge(g, use_v, e))


def test_peel_device_memory(ctx):
    g = pb.Graph.rmat(9, 9, 5000, 0.5, 0.1, 0.1, 0.3, 5)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    e = ctx.count(dg, use_v, "EDGE", "GPU", "ACOCORE")
    wing = ctx.peel_edge(dg, use_v, e)
    assert np.array_equal(pb.as_u64(wing), O.peel_edge(g, use_v, pb.as_u64(e)))


def test_tip_peel_split_walks(ctx):
    # every S vertex has more than 65536 two-hop paths, so the rounds walk their frontier
    # vertices across many workgroups (dense rows in HBM); one per round in the uniform graph
    g = pb.Graph.uniform(1000, 300, 200000, 4)
    use_v, _ = ctx.side(g)
    assert use_v
    v = O.count_vert(g, use_v)
    tip = pb.as_u64(ctx.peel(g, use_v, v))
    assert np.array_equal(tip, O.peel_vert(g, use_v, v))
    # K_{100,1500}: all 100 S vertices peel in one round, beyond the split rows of a round
    k = complete_bipartite(100, 1500)
    use_v, _ = ctx.side(k)
    assert not use_v
    v = pb.as_u64(ctx.count(k, use_v, "VERT"))
    tip = pb.as_u64(ctx.peel(k, use_v, v))
    assert (tip == 99 * (1500 * 1499 // 2)).all()
    # a power-law graph: hubs split, the rest on the LDS walks
    h = pb.Graph.chung_lu(20000, 2000, 200000, 2.1, 2.1, 6)
    use_v, _ = ctx.side(h)
    v = O.count_vert(h, use_v)
    tip = pb.as_u64(ctx.peel(h, use_v, v))
    assert np.array_equal(tip, O.peel_vert(h, use_v, v))


def test_tip_peel_split_round_beyond_kept_prefixes(ctx):
    # 32 hubs, every pair sharing t degree-2 centres: each hub is in 31 * C(t, 2) butterflies,
    # all 32 peel in one round as split rows, and that round has 32 * ceil(31 t / 256) = 65,888
    # chunks, past the 2^16 whose scanned prefixes the chunk phase keeps (the rest are scanned
    # again by the unit phase)
    t = 17000
    hubs = 32
    I, J = np.triu_indices(hubs, 1)
    npairs = len(I)
    centres = np.arange(npairs * t, dtype=np.uint32)
    v = np.concatenate([np.repeat(I.astype(np.uint32), t), np.repeat(J.astype(np.uint32), t)])
    u = np.concatenate([centres, centres])
    g = pb.Graph.from_edges(hubs, npairs * t, v, u)
    use_v, _ = ctx.side(g)
    assert not use_v  # counts on the hubs
    want = (hubs - 1) * (t * (t - 1) // 2)
    counts = pb.as_u64(ctx.count(g, use_v, "VERT"))
    assert (counts == want).all()
    tip = pb.as_u64(ctx.peel(g, use_v, counts))
    assert (tip == want).all()
    assert ctx.timing()["rounds"] <= 32  # one batch of rounds: the hubs go in the first


@pytest.mark.parametrize("nv, nu, m, seed, nb, bs_log", [
    # n_S = 2^19: 256 blocks of 2^11 items (the whole block-minimum reduction, blocks above 1024)
    (1 << 19, 1 << 19, 1 << 21, 5, 256, 11),
    # 2^17 < n_S < 2^18: 245 blocks of 2^10 items, the last one partial
    (250000, 250000, 1 << 21, 6, 245, 10),
])
def test_tip_peel_many_blocks(ctx, nv, nu, m, seed, nb, bs_log):
    # sparse power-law graphs (gamma 2.5) with 1400-3000 distinct tip levels: the tip peel keeps
    # per-block minima of D over at most 256 blocks of at least 1024 S vertices each
    g = pb.Graph.chung_lu(nv, nu, m, 2.5, 2.5, seed)
    use_v, _ = ctx.side(g)
    ns = g.nu if use_v else g.nv
    b = 10
    while ((ns - 1) >> b) + 1 > 256:
        b += 1
    assert (b, ((ns - 1) >> b) + 1) == (bs_log, nb), "case no longer covers the intended block layout"
    v = pb.as_u64(ctx.count(g, use_v, "VERT"))
    assert np.array_equal(v, O.count_vert(g, use_v, "ADEG"))
    tip = pb.as_u64(ctx.peel(g, use_v, v))
    ref = O.peel_vert(g, use_v, v)
    assert len(np.unique(ref)) > 1000
    assert np.array_equal(tip, ref)


def test_tip_peel_level_beyond_frontier_cap(ctx):
    # 3000 S vertices with no butterfly (a matching) and 4000 in one butterfly each (2000
    # disjoint C4s), plus one K_{3,3}: the levels 0 and 1 hold more items than a round's
    # frontier takes (1024), so each is peeled over several rounds at the same level; peeling
    # one vertex of a C4 leaves its partner at level 1 with its butterfly count untouched
    c4 = 2000
    s_m = np.arange(3000, dtype=np.uint32)  # the matching: V j - U j, j < 3000
    o_m = np.arange(3000, dtype=np.uint32)
    # C4 k: V vertices 3000 + 2k, 3000 + 2k + 1 both on U vertices 3000 + 2k, 3000 + 2k + 1
    s_c = 3000 + np.repeat(np.arange(2 * c4, dtype=np.uint32), 2)
    o_c = 3000 + 2 * (np.arange(4 * c4, dtype=np.uint32) // 4) + (np.arange(4 * c4, dtype=np.uint32) % 2)
    s_k = 3000 + 2 * c4 + np.repeat(np.arange(3, dtype=np.uint32), 3)
    o_k = 3000 + 2 * c4 + np.tile(np.arange(3, dtype=np.uint32), 3)
    nv = nu = 3000 + 2 * c4 + 3
    g = pb.Graph.from_edges(nv, nu, np.concatenate([s_m, s_c, s_k]), np.concatenate([o_m, o_c, o_k]))
    for use_v in (0, 1):
        v = O.count_vert(g, use_v)
        assert (v > 1).sum() == 3 and (v == 1).sum() == 2 * c4  # the K_{3,3}, the C4s
        tip = pb.as_u64(ctx.peel(g, use_v, v))
        assert np.array_equal(tip, O.peel_vert(g, use_v, v))
        assert ctx.timing()["rounds"] >= 8  # 3 rounds at level 0, 4 at level 1, the K_{3,3}
