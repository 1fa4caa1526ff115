"""CPU oracle of the learned-screen network (test infrastructure only).

Float64 restatement of the network the HIP kernels train and evaluate
(32 -> 64 -> 64 -> 1, ReLU), with the same bf16 rounding points:
    xb = bf16(fp32((x - mean) * inv_std))
    h1 = bf16(fp32(relu(W1b xb + b1)))
    h2 = relu(W2b h1 + b2)
    y  = w3b . h2 + b3
(Wb = weights rounded to bf16, biases exact), the loss
    L = mean c(r) r^2,  r = t - y,  c = w_under if r > 0 else 1,
its gradient with the roundings taken as identity, and one Adam step
(beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected).  The network has no
reference counterpart; the oracle is pinned by central finite differences
(with rounding off), Adam's closed-form first step and exact bf16 cases
(tests/test_oracle_mlp.py).

Parameter layout: W1[64][32], b1[64], W2[64][64], b2[64], w3[64], b3,
mean[32], inv_std[32].
"""
from __future__ import annotations

import numpy as np

NIN, NH = 32, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1
OFF_MEAN = NPARAM
OFF_ISTD = NPARAM + NIN
NALL = NPARAM + 2 * NIN


def bf16(x):
    """Round to the nearest bfloat16 (ties to even) via float32 bits."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return dict(W1=w[OFF_W1:OFF_B1].reshape(NH, NIN), b1=w[OFF_B1:OFF_W2], W2=w[OFF_W2:OFF_B2].reshape(NH, NH),
                b2=w[OFF_B2:OFF_W3], w3=w[OFF_W3:OFF_B3], b3=w[OFF_B3], mean=w[OFF_MEAN:OFF_ISTD],
                istd=w[OFF_ISTD:NALL])


def forward(w, X, rounding=True):
    """Returns y and the cache used by backward."""
    p = unpack(w)
    X = np.asarray(X)
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    if rounding:
        xn32 = (X.astype(np.float32) - p["mean"].astype(np.float32)) * p["istd"].astype(np.float32)
        xb = bf16(xn32)
        W1, W2, w3 = bf16(p["W1"]), bf16(p["W2"]), bf16(p["w3"])
    else:
        xb = (X - p["mean"]) * p["istd"]
        W1, W2, w3 = p["W1"], p["W2"], p["w3"]
    z1 = xb @ W1.T + p["b1"]
    h1 = r(np.maximum(z1, 0.0).astype(np.float32)) if rounding else np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + p["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + p["b3"]
    return y, dict(xb=xb, z1=z1, h1=h1, z2=z2, h2=h2, W2=W2, w3=w3)


def loss_grad(w, X, t, w_under=4.0, rounding=True):
    y, c = forward(w, X, rounding)
    n = len(y)
    res = np.asarray(t, dtype=np.float64) - y
    cw = np.where(res > 0.0, w_under, 1.0)
    loss = float(np.sum(cw * res * res) / n)
    dy = -2.0 * cw * res / n
    dz2 = (dy[:, None] * c["w3"][None, :]) * (c["z2"] > 0.0)
    dz1 = (dz2 @ c["W2"]) * (c["z1"] > 0.0)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (dz1.T @ c["xb"]).reshape(-1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (dz2.T @ c["h1"]).reshape(-1)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W3:OFF_B3] = dy @ c["h2"]
    g[OFF_B3] = dy.sum()
    return loss, g


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update of the first NPARAM entries; step is
    the 1-based index of this update.  Returns (w, m, v)."""
    w = np.array(w, dtype=np.float64)
    m = b1 * np.asarray(m, dtype=np.float64) + (1.0 - b1) * g
    v = b2 * np.asarray(v, dtype=np.float64) + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    w[:NPARAM] -= lr * mh / (np.sqrt(vh) + eps)
    return w, m, v
