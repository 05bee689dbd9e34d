"""Pins the CPU restatement of the selector MLP (oracle/mlp.py): the real-valued gradient
against central finite differences, BF16 rounding against known values, and a planted mistake
that the finite-difference check must catch."""
import numpy as np

from oracle import mlp as M


def _problem(seed, rows=12):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.NPARAM)
    p[M.OW3:M.OB3].reshape(M.HID, M.OUTP)[:, M.OUT:] = 0
    p[M.OB3 + M.OUT:] = 0
    X = rng.normal(0, 1, (rows, M.IN))
    T = rng.normal(0, 1, (rows, M.OUT))
    return p, X, T


def _fd_check(grad_fn, seed=3, probes=60):
    p, X, T = _problem(seed)
    _, g = grad_fn(p, X, T)
    rng = np.random.default_rng(seed + 1)
    live = [k for k in range(M.NPARAM)
            if not ((M.OW3 <= k < M.OB3 and (k - M.OW3) % M.OUTP >= M.OUT) or k >= M.OB3 + M.OUT)]
    worst = 0.0
    for k in rng.choice(live, probes, replace=False):
        h = 1e-6
        q1, q2 = p.copy(), p.copy()
        q1[k] += h
        q2[k] -= h
        num = (M.loss_and_grad(q1, X, T, rounding=False)[0] - M.loss_and_grad(q2, X, T, rounding=False)[0]) / (2 * h)
        worst = max(worst, abs(num - g[k]) / max(1e-6, abs(num)))
    return worst


def test_gradient_matches_finite_differences():
    assert _fd_check(lambda p, X, T: M.loss_and_grad(p, X, T, rounding=False)) < 1e-5


def test_planted_mistake_is_caught():
    def wrong(p, X, T):  # forgets the ReLU mask of the first layer
        W1, b1, W2, b2, W3, b3 = M.unpack(p)
        loss, g = M.loss_and_grad(p, X, T, rounding=False)
        y, (xb, h1, h2) = M.forward(p, X, rounding=False)
        dy = np.zeros((len(X), M.OUTP))
        dy[:, :M.OUT] = 2.0 * (y - T) / (len(X) * M.OUT)
        dh2 = (dy @ W3.T) * (h2 > 0)
        dh1 = dh2 @ W2.T
        g = g.copy()
        g[M.OW1:M.OB1] = (xb.T @ dh1).ravel()
        g[M.OB1:M.OW2] = dh1.sum(0)
        return loss, g
    assert _fd_check(wrong) > 1e-3


def test_bf16_rounding():
    one = np.float32(1.0)
    assert M.bf16(one) == 1.0
    assert M.bf16(np.float32(1 + 2 ** -8)) == 1.0                       # tie -> even
    assert M.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6            # tie -> even (up)
    assert M.bf16(np.float32(1 + 2 ** -8 + 2 ** -20)) == 1 + 2 ** -7     # above the tie
    assert M.bf16(np.float32(-2.0)) == -2.0
    x = np.random.default_rng(0).normal(0, 10, 1000).astype(np.float32)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)                # half an ulp of 7 stored bits
    assert np.all((r.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)
