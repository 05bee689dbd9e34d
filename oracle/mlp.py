"""CPU restatement of the selector MLP (test infrastructure: only tests/, the smoke check and
the bench's CPU baseline use it).

Mirrors parbutterfly_amd/csrc/mlp.hip: 16 -> 64 -> 64 -> 9 with ReLU; inputs, weights and
ReLU outputs rounded to BF16 (round to nearest even) before each product, dY = 2 (Y - T) /
(rows * 9) and the masked dH rounded to BF16 before theirs; everything else in float64 here
(float32 on the GPU).  With rounding=False it is the plain real-valued network, which the
finite-difference test pins.  The network has no counterpart in the reference.
"""
import numpy as np

IN, HID, OUT, OUTP = 16, 64, 9, 16
OW1 = 0
OB1 = OW1 + IN * HID
OW2 = OB1 + HID
OB2 = OW2 + HID * HID
OW3 = OB2 + HID
OB3 = OW3 + HID * OUTP
NPARAM = OB3 + OUTP


def bf16(x):
    """Rounds float32 values to the nearest BF16 (ties to even); returns float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OW1:OB1].reshape(IN, HID), p[OB1:OW2], p[OW2:OB2].reshape(HID, HID), p[OB2:OW3],
            p[OW3:OB3].reshape(HID, OUTP), p[OB3:NPARAM])


def _r(x, rounding):
    return bf16(x) if rounding else np.asarray(x, np.float64)


def forward(p, X, rounding=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xb = _r(X, rounding)
    h1 = _r(np.maximum(xb @ _r(W1, rounding) + b1, 0.0), rounding)
    h2 = _r(np.maximum(h1 @ _r(W2, rounding) + b2, 0.0), rounding)
    y = h2 @ _r(W3, rounding) + np.where(np.arange(OUTP) < OUT, b3, 0.0)
    return y[:, :OUT], (xb, h1, h2)


def loss_and_grad(p, X, T, rounding=True):
    """Mean squared error over rows x 9 outputs and its gradient (parameter layout)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    rows = len(X)
    y, (xb, h1, h2) = forward(p, X, rounding)
    e = y - np.asarray(T, np.float64)
    loss = float((e * e).sum() / (rows * OUT))
    dy = np.zeros((rows, OUTP))
    dy[:, :OUT] = 2.0 * e / (rows * OUT)
    dy = _r(dy, rounding)
    g = np.zeros(NPARAM)
    g[OW3:OB3] = (h2.T @ dy).ravel()
    g[OB3:NPARAM] = dy.sum(0)
    dh2 = _r((dy @ _r(W3, rounding).T) * (h2 > 0), rounding)
    g[OW2:OB2] = (h1.T @ dh2).ravel()
    g[OB2:OW3] = dh2.sum(0)
    dh1 = _r((dh2 @ _r(W2, rounding).T) * (h1 > 0), rounding)
    g[OW1:OB1] = (xb.T @ dh1).ravel()
    g[OB1:OW2] = dh1.sum(0)
    return loss, g
