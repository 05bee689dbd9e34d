"""Selector MLP on the GPU (BF16 MFMA kernels) against the CPU oracle, and the batch features
against host computations pinned by the oracle's rankings."""
import numpy as np
import pytest

import parbutterfly_amd as pb
from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _data(seed, rows):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (rows, M.IN)).astype(np.float32)
    T = rng.normal(0, 1, (rows, M.OUT)).astype(np.float32)
    return X, T


@pytest.mark.parametrize("rows", [128, 77, 1])
def test_gradient_vs_oracle(ctx, rows):
    p = pb.mlp_init(5)
    X, T = _data(rows, rows)
    g, loss = ctx.mlp_grad(p, X, T)
    lo, go = M.loss_and_grad(p, X, T, rounding=True)
    # FP32 accumulation in another order; an intermediate near a BF16 rounding boundary may round
    # the other way, so the tolerance is relative to the gradient's scale
    assert abs(loss - lo) <= 1e-4 * max(1.0, abs(lo))
    scale = np.abs(go).max()
    assert np.abs(g - go).max() <= 2e-2 * scale
    assert np.median(np.abs(g - go)) <= 1e-3 * scale


def test_forward_vs_oracle(ctx):
    p = pb.mlp_init(9)
    X, _ = _data(2, 1000)
    Y = ctx.mlp_predict(p, X)
    Yo, _ = M.forward(p, X, rounding=True)
    assert np.abs(Y - Yo).max() <= 1e-3 * max(1.0, np.abs(Yo).max())


def test_training_fits(ctx):
    # a smooth target the network can fit: the loss falls by an order of magnitude
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (2048, M.IN)).astype(np.float32)
    W = rng.normal(0, 0.5, (M.IN, M.OUT))
    T = (np.tanh(X @ W) + 0.1 * X[:, :M.OUT] ** 2).astype(np.float32)
    p = pb.mlp_init(1)
    l0 = M.loss_and_grad(p, X, T)[0]
    l1 = ctx.mlp_train(p, X, T, epochs=60, lr=3e-3, seed=2)
    assert l1 < 0.1 * l0
    lo = M.loss_and_grad(p, X, T)[0]
    assert abs(lo - l1) < 0.5 * lo + 1e-3  # the trained parameters really give that loss


def test_batch_features(ctx):
    gs = [pb.Graph.chung_lu(300 + 37 * k, 200 + 53 * k, 2000 + 300 * k, 2.2, 2.4, 70 + k) for k in range(6)]
    gs.append(pb.Graph.batch_graph(1234))
    F = ctx.batch_features(gs)
    for k, g in enumerate(gs):
        dv, du = np.diff(g.offsetsV).astype(np.float64), np.diff(g.offsetsU).astype(np.float64)
        wv, wu = int((dv * (dv - 1) // 2).sum()), int((du * (du - 1) // 2).sum())
        use_v = wv <= wu
        wad = O.rank_wedges(g, O.rank(g, "ADEG")[1])
        waco = O.rank_wedges(g, O.rank(g, "ACOCORE")[1])
        exp = [np.log2(g.nv + 1), np.log2(g.nu + 1), np.log2(g.m + 1), np.log2(dv.mean() + 1), np.log2(du.mean() + 1),
               np.log2((dv ** 2).mean() + 1), np.log2((du ** 2).mean() + 1), np.log2(dv.max() + 1),
               np.log2(du.max() + 1), np.log2(wv + 1), np.log2(wu + 1), np.log2(min(wv, wu) + 1), np.log2(wad + 1),
               np.log2(waco + 1), 1.0 if use_v else 0.0, np.log2(g.m / (g.nv * g.nu) + 1e-9)]
        assert np.allclose(F[k], np.array(exp, np.float32), rtol=1e-5, atol=1e-6), k
