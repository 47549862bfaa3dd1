"""CPU restatement of the learned selector's network and its gradient.

TEST INFRASTRUCTURE ONLY (the checker of emptyheaded_amd's HIP MLP kernels).
The selector is this project's addition, not a reference function: an
8 -> 32 -> 32 -> 2 ReLU MLP, parameters packed as W1[32][8], b1[32],
W2[32][32], b2[32], W3[2][32], b3[2], weights rounded to FP16 before use,
loss = sum(mask * (out - y)^2) / sum(mask).  Computed here in float64; the
backward pass is pinned against central finite differences in the tests.
"""
import numpy as np

IN, H, OUT = 8, 32, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + OUT * H
NPARAM = OFF_B3 + OUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(OUT, H), p[OFF_B3:NPARAM])


def round_half(p):
    return np.asarray(p, dtype=np.float32).astype(np.float16).astype(np.float64)


def forward(params, norm, x, half=True):
    w = round_half(params) if half else np.asarray(params, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(w)
    norm = np.asarray(norm, np.float64)
    xn = (np.asarray(x, np.float64) - norm[:IN]) * norm[IN:]
    z1 = xn @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    out = h2 @ W3.T + b3
    return out, (xn, z1, h1, z2, h2)


def loss_and_grad(params, norm, x, y, m, half=True):
    out, (xn, z1, h1, z2, h2) = forward(params, norm, x, half)
    w = round_half(params) if half else np.asarray(params, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(w)
    m = np.asarray(m, np.float64)
    cnt = m.sum()
    e = out - np.asarray(y, np.float64)
    loss = float((m * e * e).sum() / cnt)
    g = 2.0 * m * e / cnt                      # d loss / d out
    gW3 = g.T @ h2
    gb3 = g.sum(0)
    g2 = (g @ W3) * (z2 > 0)
    gW2 = g2.T @ h1
    gb2 = g2.sum(0)
    g1 = (g2 @ W2) * (z1 > 0)
    gW1 = g1.T @ xn
    gb1 = g1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad


def numeric_grad(params, norm, x, y, m, eps=1e-6):
    """Central differences of the loss in float64 (no FP16 rounding)."""
    p = np.asarray(params, np.float64).copy()
    g = np.zeros_like(p)
    for i in range(len(p)):
        old = p[i]
        p[i] = old + eps
        lp, _ = loss_and_grad(p, norm, x, y, m, half=False)
        p[i] = old - eps
        lm, _ = loss_and_grad(p, norm, x, y, m, half=False)
        p[i] = old
        g[i] = (lp - lm) / (2 * eps)
    return g
