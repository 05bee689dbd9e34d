"""Parity at the full workload sizes, through properties that do not depend on size.

The oracle cannot count a 2^24-edge graph in the time a test may take, so at full size the
GPU results are checked by:
  * invariance: per-vertex and per-edge counts are identical under ADEG and ACOCORE ranking;
  * sums: sum over S of the vertex counts = 2 TOTAL, sum of the edge counts = 4 TOTAL;
  * independent recounts of sampled vertices and edges from their two-hop neighbourhoods
    (co-occurrence counts in numpy, no ranking involved);
  * for the tip and wing peels, the defining property of the peel levels: when the peel reaches
    a level k, the items still live are exactly those numbered >= k, each one's D is its
    butterfly count among them, and k is the least of those counts; so in the subgraph of the
    vertices (edges) numbered >= k every one lies in at least k butterflies and some in exactly
    k, at every level checked.
"""
import numpy as np
import pytest

import parbutterfly_amd as pb

pytestmark = pytest.mark.gpu


def _sides(g, use_v):
    """(offS, edgS, offO, edgO) as int64 numpy arrays; S = U when use_v."""
    if use_v:
        return (g.offsetsU.astype(np.int64), g.edgesU.astype(np.int64), g.offsetsV.astype(np.int64),
                g.edgesV.astype(np.int64))
    return (g.offsetsV.astype(np.int64), g.edgesV.astype(np.int64), g.offsetsU.astype(np.int64),
            g.edgesU.astype(np.int64))


def _recount(g, use_v, xs):
    """Per-vertex count of each x in xs (S ids) and, for every edge (x, v), its per-edge count
    and O-CSR index, from co-occurrence counts over the two-hop neighbourhood of x."""
    offS, edgS, offO, edgO = _sides(g, use_v)
    ns = len(offS) - 1
    verts, edges = [], []
    for x in xs:
        nx = edgS[offS[x]:offS[x + 1]]
        hop = np.concatenate([edgO[offO[v]:offO[v + 1]] for v in nx]) if len(nx) else np.zeros(0, np.int64)
        cnt = np.bincount(hop, minlength=ns)
        cnt[x] = 0
        verts.append(int((cnt * (cnt - 1) // 2).sum()))
        for v in nx:
            nb = edgO[offO[v]:offO[v + 1]]
            p = int(offO[v] + np.searchsorted(nb, x))
            others = nb[nb != x]
            edges.append((p, int((cnt[others] - 1).sum())))
    return verts, edges


def _sample(g, use_v, k, max_deg, seed):
    offS = _sides(g, use_v)[0]
    deg = np.diff(offS)
    cand = np.nonzero((deg > 0) & (deg <= max_deg))[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, len(cand)), replace=False)


def test_chunglu_2p24(ctx):
    # the headline graph: Chung-Lu nv = 2^20, nu = 2^18, m = 2^24, gamma 2.1, seed 2
    g = pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    assert (use_v, _) == g.side()
    v_ad = pb.as_u64(ctx.count(dg, use_v, "VERT", "GPU", "ADEG"))
    v_ac = pb.as_u64(ctx.count(dg, use_v, "VERT", "GPU", "ACOCORE"))
    assert np.array_equal(v_ad, v_ac)
    assert np.array_equal(pb.as_u64(ctx.count(g, use_v, "VERT", "GPU", "ADEG")), v_ad)  # host path
    e_ad = pb.as_u64(ctx.count(dg, use_v, "EDGE", "GPU", "ADEG"))
    e_ac = pb.as_u64(ctx.count(dg, use_v, "EDGE", "GPU", "ACOCORE"))
    assert np.array_equal(e_ad, e_ac)
    t = ctx.count(dg, use_v, "TOTAL", "GPU", "ADEG")
    assert int(v_ad.sum(dtype=np.uint64)) == 2 * t
    assert int(e_ad.sum(dtype=np.uint64)) == 4 * t
    xs = _sample(g, use_v, 12, 24, seed=5)
    verts, edges = _recount(g, use_v, xs)
    for x, b in zip(xs, verts):
        assert int(v_ad[x]) == b, f"vertex {x}"
    for p, c in edges:
        assert int(e_ad[p]) == c, f"edge {p}"


def test_rmat_2p22_count_and_wing_peel(ctx):
    # R-MAT 2^22 x 2^22, 1.2e8 edges, (a, b, c, d) = (.5, .1, .1, .3), seed 3
    g = pb.Graph.rmat(22, 22, 120_000_000, 0.5, 0.1, 0.1, 0.3, 3)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    e = ctx.count(dg, use_v, "EDGE", "GPU", "ACOCORE")
    E = e.numpy()
    assert np.array_equal(E, ctx.count(dg, use_v, "EDGE", "GPU", "ADEG").numpy())
    t = ctx.count(dg, use_v, "TOTAL", "GPU", "ACOCORE")
    assert int(E.sum(dtype=np.uint64)) == 4 * t
    xs = _sample(g, use_v, 8, 64, seed=7)
    _, edges = _recount(g, use_v, xs)
    for p, c in edges:
        assert int(E[p]) == c, f"edge {p}"
    wing = ctx.peel_edge(dg, use_v, e).numpy()
    assert np.all(wing <= E)
    # wing subgraphs: the top level, and the largest level whose subgraph has <= 4M edges
    offO, edgO = (g.offsetsV, g.edgesV) if use_v else (g.offsetsU, g.edgesU)
    owner = np.repeat(np.arange(len(offO) - 1, dtype=np.uint32), np.diff(offO.astype(np.int64)))
    vv, uu = (owner, edgO) if use_v else (edgO, owner)
    levels = np.unique(wing)[::-1]
    sizes = np.cumsum([int(np.count_nonzero(wing == k)) for k in levels])
    ks = {int(levels[0])} | {int(levels[j]) for j in range(len(levels)) if sizes[j] <= 4_000_000}
    for k in sorted(ks, reverse=True)[:2] + ([min(ks)] if len(ks) > 2 else []):
        sel = wing >= k
        h = pb.Graph.from_edges(g.nv, g.nu, np.ascontiguousarray(vv[sel]), np.ascontiguousarray(uu[sel]))
        huv, _ = ctx.side(h)
        hc = pb.as_u64(ctx.count(h, huv, "EDGE", "GPU", "ADEG"))
        assert int(hc.min()) == k, f"wing level {k}"


def test_chunglu_2p24_tip_peel(ctx):
    # tip numbers of the headline graph (about 2e5 levels, the late ones one hub each, walked
    # across the device); checked by the defining property of tip numbers at levels spread over
    # the whole range: when the peel reaches a level k, the S vertices still live are exactly
    # those with tip >= k, each one's D equals its butterfly count in the subgraph they induce
    # with all of O (D = max(B, k') and D > k' for the previous level k'), and k is the least of
    # those counts.  So in that subgraph every vertex with tip >= k lies in at least k
    # butterflies and some vertex in exactly k: tips too low or too high at the level fail.
    g = pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)
    dg = g.to_device(ctx)
    use_v, _ = ctx.side(dg)
    b = ctx.count(dg, use_v, "VERT", "GPU", "ADEG")
    tip = ctx.peel(dg, use_v, b).numpy()
    B = pb.as_u64(b)
    assert np.all(tip <= B)
    offS, edgS = (g.offsetsU, g.edgesU) if use_v else (g.offsetsV, g.edgesV)
    owner = np.repeat(np.arange(len(offS) - 1, dtype=np.uint32), np.diff(offS.astype(np.int64)))
    levels = np.unique(tip)[::-1]  # descending
    L = len(levels)
    ks = sorted({int(levels[j]) for j in (0, 1000, 3999, L // 100, L // 10, L // 2, L - 1) if j < L}, reverse=True)
    assert len(ks) >= 5
    for k in ks:
        keep = tip[owner] >= k
        s_ids, o_ids = owner[keep], edgS[keep].astype(np.uint32)
        vv, uu = (o_ids, s_ids) if use_v else (s_ids, o_ids)
        h = pb.Graph.from_edges(g.nv, g.nu, np.ascontiguousarray(vv), np.ascontiguousarray(uu))
        hc = pb.as_u64(ctx.count(h, use_v, "VERT", "GPU", "ADEG"))
        sel = tip >= k
        assert int(hc[sel].min()) == k, f"tip level {k}"
