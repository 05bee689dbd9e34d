"""The shipped selector (parbutterfly_amd/selector_model.npz) on the 512 held-out graphs of the
mixed-batch distribution (tests/golden/selector_heldout512.npz: their features and the measured
log2 runtimes of all 9 variants, written by tests/golden/make_selector_heldout.py).

  * the fixture's features are what pbf_batch_features computes for those graphs now;
  * the GPU inference (BF16 MFMA kernel) agrees with the CPU restatement of the network;
  * the regret against exhaustive search and the gain over the method without the selector
    (fixed ADEG, default knobs) are recomputed from the measured times, reproduce
    profiles/selector_heldout512.json, and the selector beats the fixed choice;
  * learned batches give every graph's counts bit-exactly (against the oracle) whatever
    variants the network picks.
"""
import json
import os

import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import mlp as M
from oracle import oracle as O
from parbutterfly_amd import selector as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILE = os.path.join(ROOT, "profiles", "selector_heldout512.json")


@pytest.fixture(scope="module")
def heldout():
    F, T, seeds = S.load_fixture()
    assert F.shape == (512, pb.MLP_IN) and T.shape == (512, S.NVAR) and len(seeds) == 512
    return F, T, seeds


def test_fixture_features(ctx, heldout):
    F, _, seeds = heldout
    pick = list(range(0, 512, 16))
    gs = [pb.Graph.batch_graph(int(seeds[k])) for k in pick]
    Fn = ctx.batch_features(gs)
    assert np.allclose(Fn, F[pick], rtol=1e-6, atol=1e-6)


def test_inference_vs_oracle(ctx, heldout):
    F, _, _ = heldout
    model = S.load()
    X = ((F - model["fmu"]) / model["fsd"]).astype(np.float32)
    Y = ctx.mlp_predict(model["params"], X)
    Yo, _ = M.forward(model["params"], X, rounding=True)
    assert np.abs(Y - Yo).max() <= 1e-3 * max(1.0, np.abs(Yo).max())
    # the choice is the argmin; where the two disagree, the predicted times are a near tie
    P, Po = Y * model["tsd"] + model["tmu"], Yo * model["tsd"] + model["tmu"]
    a, b = P.argmin(1), Po.argmin(1)
    diff = np.nonzero(a != b)[0]
    assert len(diff) <= 5
    for k in diff:
        assert abs(Po[k, a[k]] - Po[k, b[k]]) <= 1e-2


def test_regret_and_gain(ctx, heldout):
    F, T, _ = heldout
    r = S.evaluate(ctx, S.load(), F, T)
    with open(PROFILE) as f:
        prof = json.load(f)
    for k in ("mean_regret_selected", "mean_regret_fixed", "total_ms_exhaustive_best", "total_ms_selected",
              "total_ms_fixed", "gain_selected_vs_fixed"):
        assert abs(r[k] - prof[k]) <= 1e-4 * max(1.0, abs(prof[k])), k
    assert r["variant_histogram"] == prof["variant_histogram"]
    # exhaustive search bounds the selector from below; the selector beats the fixed variant
    assert r["total_ms_exhaustive_best"] <= r["total_ms_selected"] < r["total_ms_fixed"]
    assert r["mean_regret_selected"] < 0.25 * r["mean_regret_fixed"]
    assert r["gain_selected_vs_fixed"] >= 1.2


def test_learned_batch_exact(ctx, heldout):
    _, _, seeds = heldout
    gs = [pb.Graph.batch_graph(int(s)) for s in seeds[:48]]
    verts, tots, pick = S.count_learned(ctx, S.load(), gs)
    assert len(pick) == len(gs)
    for k, g in enumerate(gs):
        use_v, _ = O.side(g)
        ov = O.count_vert(g, use_v, "ADEG") if g.m else np.zeros(g.nu if use_v else g.nv, np.uint64)
        assert np.array_equal(np.asarray(verts[k], np.uint64), ov), f"graph {k} (seed {seeds[k]}) VERT"
        assert int(tots[k]) == int(ov.sum()) // 2, f"graph {k} total"


def test_select_abi_and_fallback(ctx, heldout):
    # pbf_select picks the argmin of the model's predictions; without a model, or with a model
    # whose predictions are not finite, every graph gets the fixed variant
    _, _, seeds = heldout
    gs = [pb.Graph.batch_graph(int(s)) for s in seeds[:24]]
    model = S.load()
    v = ctx.select(gs, model)
    assert np.array_equal(v, S.predict(ctx, model, ctx.batch_features(gs)).argmin(1))
    assert np.all(ctx.select(gs, None) == pb.VARIANT_FIXED)
    bad = dict(model)
    bad["tsd"] = np.full(9, np.nan, np.float32)
    assert np.all(ctx.select(gs, bad) == pb.VARIANT_FIXED)
    # fixed fallback and learned runs give identical counts (host and device graphs)
    (vf, tf), pf = ctx.count_batch_learned(gs, None)
    assert np.all(pf == pb.VARIANT_FIXED)
    outs, pl = ctx.count_batch_learned([g.to_device(ctx) for g in gs], model)
    vl, tl = S.gather(outs, len(gs))
    assert np.array_equal(pl, v)
    assert np.array_equal(tf, tl)
    for a, b in zip(vf, vl):
        assert np.array_equal(a, b)
