"""CPU oracle of the learned switch's MLP (test infrastructure only).

Per-particle network 8 -> 64 -> 64 -> 1, SiLU, SiLU, sigmoid, restated in
numpy FP64 with the BF16 rounding points of the GPU kernels
(warpsph_amd/csrc/mlp.hip): inputs, weights, the two hidden activations and
the two back-propagated pre-activation gradients that feed the matrix cores
are rounded to BF16 (round to nearest even); every sum is exact FP64 here
and FP32 on the GPU.  ``rounding=False`` gives the plain FP64 network whose
gradient is checked against finite differences (tests/test_oracle_mlp.py).

Parameter layout (FP32 master copy): W1[64][8] | b1[64] | W2[64][64] | b2[64]
| w3[64] | b3.  Loss: mean_p (alpha_p - t_p)^2.  This network has no
counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 8, 64
NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1


def bf16(a) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NH * NIN].reshape(NH, NIN); o += NH * NIN
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, x, rounding=True):
    """alpha [n] for normalised features x [n][8]; also returns the tape."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    W1, W2, w3 = r(W1), r(W2), r(w3)
    xb = r(x)
    z1 = xb @ W1.T + b1
    a1 = r(z1 * _sig(z1))
    z2 = a1 @ W2.T + b2
    a2 = r(z2 * _sig(z2))
    z3 = a2 @ w3 + b3
    al = _sig(z3)
    return al, (xb, W1, W2, w3, z1, a1, z2, a2, al)


def _dsilu(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


def loss_grad(p, x, t, rounding=True):
    """(loss, grad [NPARAM]) of mean (alpha - t)^2."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    n = x.shape[0]
    al, (xb, W1, W2, w3, z1, a1, z2, a2, _) = forward(p, x, rounding)
    e = al - t
    loss = float(np.mean(e * e))
    dz3 = 2.0 * e / n * al * (1.0 - al)
    gw3 = a2.T @ dz3
    gb3 = dz3.sum()
    dz2 = (dz3[:, None] * w3[None, :]) * _dsilu(z2)
    gb2 = dz2.sum(axis=0)
    dz2r = r(dz2)
    gW2 = dz2r.T @ a1
    da1 = dz2r @ W2
    dz1 = da1 * _dsilu(z1)
    gb1 = dz1.sum(axis=0)
    gW1 = r(dz1).T @ xb
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def init_params(seed=0):
    """He-style initialisation (FP32), output bias 0."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, np.sqrt(2.0 / NIN), (NH, NIN))
    W2 = rng.normal(0.0, np.sqrt(2.0 / NH), (NH, NH))
    w3 = rng.normal(0.0, np.sqrt(1.0 / NH), NH)
    p = np.concatenate([W1.ravel(), np.zeros(NH), W2.ravel(), np.zeros(NH), w3, [0.0]])
    return p.astype(np.float32)


def adam(p, m, v, g, lr, b1, b2, eps, step):
    """One Adam step (Kingma & Ba 2015), FP64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
