"""Parity at the full workload sizes against the reference itself, and the full batched set.

  * The headline graph (Chung-Lu nv = 2^20, nu = 2^18, m = 2^24, gamma 2.1, seed 2): per-vertex
    and per-edge counts bit-exact against the reference's BATCHS Count / CountE
    (apps/hyper/butterfly_count.h:1634, butterfly_ecount.h:1766), compiled from the reference's
    own sources into oracle/_ref by oracle/Makefile.  Skipped where that build is absent.
  * The 4096-graph mixed batch (seeds 1000 .. 5095): the captured union pass and the captured
    pass after autotuning against the plain per-graph loop for every graph, and against the
    oracle for a seeded sample of 512 graphs plus the 16 with the most edges.
"""
import os
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


@pytest.fixture(scope="module")
def headline():
    return pb.Graph.chung_lu(1 << 20, 1 << 18, 1 << 24, 2.1, 2.1, 2)


@pytest.mark.parametrize("per, op", [("VERT", "vert"), ("EDGE", "edge")])
def test_headline_vs_reference(ctx, headline, per, op):
    if not O.ref_available():
        pytest.skip("oracle/_ref (the compiled reference) is not built")
    g = headline
    dg = g.to_device(ctx)
    use_v, w = ctx.side(dg)
    got = pb.as_u64(ctx.count(dg, use_v, per, "GPU", "ADEG"))
    with tempfile.TemporaryDirectory() as td:
        gp = os.path.join(td, "g.bcsr")
        g.write_binary(gp)
        print(f"reference {op} on {_threads()} threads", flush=True)
        ruv, rw, ref, _, _ = O.run_ref(op, gp, os.path.join(td, "o.bin"), "BATCHS", "ADEG", threads=_threads(),
                                       timeout=600)
    assert (ruv, rw) == (use_v, w)
    assert len(ref) == len(got)
    assert np.array_equal(ref.astype(np.uint64), got)


def test_batch_4096(ctx):
    seeds = list(range(1000, 1000 + 4096))
    with ThreadPoolExecutor(max_workers=_threads()) as ex:
        hg = list(ex.map(pb.Graph.batch_graph, seeds))
    dgs = [g.to_device(ctx) for g in hg]
    print(f"batch: {len(hg)} graphs, {sum(g.m for g in hg)} edges", flush=True)

    def host(out):
        v, o, t = (x.numpy() for x in out)
        return v, o, t[:len(hg)]

    vp, op, tp = host(ctx.count_batch(dgs, "ADEG", mode="plain"))
    outs = ctx.count_batch(dgs, "ADEG", mode="union", capture=True)  # eager pass, then capture
    for rep in range(2):
        outs = ctx.count_batch(dgs, "ADEG", mode="union", capture=True, outputs=outs)  # replays
        v, o, t = host(outs)
        assert np.array_equal(o, op) and np.array_equal(t, tp), f"captured replay {rep}"
        assert np.array_equal(v[:int(op[-1])], vp[:int(op[-1])]), f"captured replay {rep}"
    best, _ = ctx.batch_autotune(dgs)
    outs = ctx.count_batch(dgs, "ADEG", mode="union", capture=True, outputs=outs)
    outs = ctx.count_batch(dgs, "ADEG", mode="union", capture=True, outputs=outs)
    v, o, t = host(outs)
    assert np.array_equal(o, op) and np.array_equal(t, tp), f"autotuned (candidate {best})"
    assert np.array_equal(v[:int(op[-1])], vp[:int(op[-1])]), f"autotuned (candidate {best})"
    ctx.set_tuning()
    rng = np.random.default_rng(4096)
    ms = np.array([g.m for g in hg])
    sample = sorted(set(rng.choice(len(hg), 512, replace=False).tolist()) | set(np.argsort(ms)[-16:].tolist()))
    for k in sample:
        g = hg[k]
        use_v, _ = O.side(g)
        ov = O.count_vert(g, use_v, "ADEG") if g.m else np.zeros(g.nu if use_v else g.nv, np.uint64)
        assert np.array_equal(vp[int(op[k]):int(op[k + 1])], ov), f"graph {k} (seed {seeds[k]})"
        assert int(tp[k]) == int(ov.sum()) // 2, f"graph {k} total"
    del dgs
    ctx.trim()
