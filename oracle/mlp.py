"""CPU FP64 restatement of the subgrid-closure MLP -- test infrastructure only.

The network 18 -> 64 -> 64 -> 1 with GELU (z (1 + erf(z / sqrt 2)) / 2) and its parameter order
(W1[18][64] input-major, b1[64], W2[64][64], b2[64], W3[64], b3) are those of the GPU kernels; the reference
has no learned closure (its Smagorinsky inputs m7neq / m8neq, lbm_core4.cu:172-173, are the features).
Only tests/ use this module, as the checker of the GPU forward, gradient and Adam kernels:

  * ``forward`` / ``grad`` in plain FP64 (``mirror=False``); ``grad`` is pinned by central finite
    differences in the tests;
  * the same with ``mirror=True``: every BF16 rounding the kernels make (inputs, weights, hidden activations,
    dZ operands of the weight-gradient products) applied to FP64 values, so that the GPU agrees with it to
    FP32 accumulation error;
  * ``adam`` in FP32 with the kernels' operation order.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NIN, NH = 18, 64
OFF_W1, OFF_B1 = 0, NIN * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NIN, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def init(seed=0):
    """He-style random parameters (FP32 values)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2.0 / NIN), NIN * NH)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1.0 / NH), NH)
    return p


def forward(p, x, mirror=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    if mirror:
        x, W1, W2, W3 = bf16(x), bf16(W1), bf16(W2), bf16(W3)
    z1 = x @ W1 + b1
    h1 = gelu(z1)
    if mirror:
        h1 = bf16(h1)
    z2 = h1 @ W2 + b2
    h2 = gelu(z2)
    if mirror:
        h2 = bf16(h2)
    y = h2 @ W3 + b3
    return y, (x, W1, W2, W3, z1, h1, z2, h2)


def grad(p, x, t, wgt, mirror=False):
    """(gradient [NPARAM], loss) of L = sum_s wgt_s (y_s - t_s)^2."""
    y, (xx, W1, W2, W3, z1, h1, z2, h2) = forward(p, x, mirror)
    e = y - np.asarray(t, np.float64)
    wgt = np.asarray(wgt, np.float64)
    loss = float(np.sum(wgt * e * e))
    dy = 2.0 * wgt * e
    g = np.zeros(NPARAM)
    g[OFF_B3] = dy.sum()
    g[OFF_W3:OFF_B3] = h2.T @ dy
    dz2 = np.outer(dy, W3) * gelu_d(z2)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz2m = bf16(dz2) if mirror else dz2
    g[OFF_W2:OFF_B2] = (h1.T @ dz2m).ravel()
    dz1 = (dz2m @ W2.T) * gelu_d(z1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    dz1m = bf16(dz1) if mirror else dz1
    g[OFF_W1:OFF_B1] = (xx.T @ dz1m).ravel()
    return g, loss


def adam(p, m, v, g, lr, beta1, beta2, eps, bc1, bc2):
    """One Adam step in FP32, in the kernel's operation order; returns new (p, m, v)."""
    f = np.float32
    p, m, v, g = (np.asarray(a, np.float32) for a in (p, m, v, g))
    b1, b2 = f(beta1), f(beta2)
    mm = b1 * m + (f(1) - b1) * g
    vv = b2 * v + (f(1) - b2) * (g * g)
    upd = f(lr) * (mm / f(bc1)) / (np.sqrt(vv / f(bc2)) + f(eps))
    return (p - upd).astype(np.float32), mm.astype(np.float32), vv.astype(np.float32)
