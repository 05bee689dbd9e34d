"""Batched counting (pbf_count_batch): every graph's per-vertex counts and total must equal the
single-graph results, which the oracle pins (bit-exact uint64)."""
import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O
from tests.graphs import complete_bipartite, empty_graph

pytestmark = pytest.mark.gpu


def _mixed(k, seed0=50):
    gs = []
    for j in range(k):
        s = seed0 + j
        rng = np.random.default_rng(s)
        nv, nu = int(rng.integers(8, 3000)), int(rng.integers(8, 3000))
        m = int(min(nv * nu // 2, rng.integers(nv + nu, 12 * (nv + nu))))
        gs.append(pb.Graph.chung_lu(nv, nu, m, 2.0 + rng.random(), 2.0 + rng.random(), s))
    gs.append(complete_bipartite(4, 6))
    gs.append(empty_graph(7, 3))
    gs.append(complete_bipartite(9, 2))
    return gs


def _check(gs, vert, tot, rank_type="ADEG"):
    # counts do not depend on the ranking; the oracle ranks by ADEG or ACOCORE
    ork = "ACOCORE" if rank_type == "ACOCORE" else "ADEG"
    for k, g in enumerate(gs):
        use_v, _ = O.side(g)
        ov = O.count_vert(g, use_v, ork) if g.m else np.zeros(g.nu if use_v else g.nv, np.uint64)
        assert np.array_equal(np.asarray(vert[k], np.uint64), ov), f"graph {k} VERT"
        assert int(tot[k]) == int(ov.sum()) // 2, f"graph {k} total"


@pytest.mark.parametrize("rank_type", ["ADEG", "ACOCORE", "SIDE"])
def test_union_host(ctx, rank_type):
    gs = _mixed(24)
    vert, tot = ctx.count_batch(gs, rank_type, mode="union")
    _check(gs, vert, tot, rank_type)


def test_union_split_capacity(ctx):
    # 8300 copies of K_{2,4097}: the second V vertex of every copy (degree 4097, graph-local
    # rank 1) is split across workgroups; there are more such endpoints than the split
    # capacity (8192), so the rest run as ordinary workgroup endpoints.  Closed form: every V
    # vertex and every graph has C(4097, 2) butterflies.
    dg = complete_bipartite(2, 4097).to_device(ctx)
    k = 8300
    vout, voff, tot = (x.numpy() for x in ctx.count_batch([dg] * k, "ADEG", mode="union"))
    c = 4097 * 4096 // 2
    assert np.array_equal(voff[:k + 1], 2 * np.arange(k + 1, dtype=np.uint64))
    assert np.all(tot[:k] == c)
    assert np.all(vout[:2 * k] == c)


def test_plain_equals_union(ctx):
    gs = _mixed(12, 200)
    v1, t1 = ctx.count_batch(gs, mode="plain")
    v2, t2 = ctx.count_batch(gs, mode="union")
    assert np.array_equal(t1, t2)
    for a, b in zip(v1, v2):
        assert np.array_equal(a, b)
    _check(gs, v1, t1)


def test_generator_graphs(ctx):
    # graphs of the mixed-batch distribution (sizes log-uniform up to 65536 per side)
    gs = [pb.Graph.batch_graph(1000 + k) for k in range(6)]
    vert, tot = ctx.count_batch(gs, mode="union")
    _check(gs, vert, tot)


def test_device_capture_replays(ctx):
    gs = _mixed(16, 400)
    dgs = [g.to_device(ctx) for g in gs]
    ref_v, ref_t = ctx.count_batch(gs, mode="plain")
    outs = ctx.count_batch(dgs, mode="union", capture=True)  # eager pass, then capture
    for rep in range(3):
        vout, voff, tot = ctx.count_batch(dgs, mode="union", capture=True, outputs=outs)  # replay
        off = voff.numpy()
        vv = vout.numpy()
        assert np.array_equal(tot.numpy()[:len(gs)], ref_t), f"replay {rep}"
        for k in range(len(gs)):
            assert np.array_equal(vv[int(off[k]):int(off[k + 1])], ref_v[k]), f"replay {rep} graph {k}"
        # eager work between replays
        g = gs[rep]
        use_v, _ = ctx.side(g)
        assert np.array_equal(pb.as_u64(ctx.count(g, use_v, "VERT")), ref_v[rep])
        ctx.synchronize()
    ctx.trim()  # releases the captured graph with the scratch


def test_autotune_keeps_results(ctx):
    gs = _mixed(10, 600)
    best, ms = ctx.batch_autotune(gs)
    assert 0 <= best < 8 and ms > 0
    vert, tot = ctx.count_batch(gs, mode="union")
    _check(gs, vert, tot)
    ctx.set_tuning()  # defaults for the tests that follow


def test_learned_selection_is_exact(ctx):
    # any variant choice gives the same counts: an untrained network scatters the graphs over
    # the variants, and every part must still match the plain loop
    from parbutterfly_amd import selector as S
    gs = _mixed(14, 800)
    model = dict(params=pb.mlp_init(11), fmu=np.zeros(16, np.float32), fsd=np.ones(16, np.float32),
                 tmu=np.zeros(9, np.float32), tsd=np.ones(9, np.float32))
    verts, tots, pick = S.count_learned(ctx, model, gs)
    assert len(set(pick.tolist())) >= 2
    _check(gs, verts, tots)
