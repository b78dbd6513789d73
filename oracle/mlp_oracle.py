"""CPU oracle of the learned-nprobe MLP. TEST INFRASTRUCTURE ONLY: used by
tests/ as the checker of the GPU forward/backward kernels, never by the
product. No reference counterpart: the reference has one nprobe per call
(include/katoann/index.h:56-62).

Network: x (F features) -> relu(W1 x + b1) -> relu(W2 h1 + b2) -> W3 h2 + b3,
loss = mean((z - y)^2) or, at quantile tau, the pinball loss
mean(rho_tau(y - z)) with rho_tau(u) = u (tau - [u < 0]). With ``bf16=True`` it follows the GPU's rounding
points exactly, accumulating in float64 where the GPU accumulates in fp32:
products take bf16 operands (the inputs, the weights, the hidden activations
and the gradients fed to the next product), and activations and gradients are
kept in fp32 otherwise. With ``bf16=False`` everything is float64. That
form is pinned by finite differences in tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round to bfloat16 (round to nearest even) and back, as float64."""
    a = np.asarray(x, dtype=np.float32).copy()
    b = a.view(np.uint32)
    finite = np.isfinite(a)
    rnd = ((b >> 16) & 1) + 0x7FFF
    b[finite] = ((b[finite].astype(np.uint64) + rnd[finite]) & 0xFFFF0000).astype(np.uint32)
    return a.astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def forward(p, X, bf=True):
    """Returns (z, cache). p: dict W1 (H x F), b1, W2 (H x H), b2, W3 (1 x H), b3."""
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    c = f32 if bf else (lambda v: np.asarray(v, np.float64))
    Xb = r(X)
    W1, W2, W3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    a1 = c(Xb @ W1.T + p["b1"])
    H1 = np.maximum(a1, 0.0)
    H1b = r(H1)
    a2 = c(H1b @ W2.T + p["b2"])
    H2 = np.maximum(a2, 0.0)
    H2b = r(H2)
    z = c(H2b @ W3.T + p["b3"])[:, 0]
    return z, dict(Xb=Xb, H1=H1, H1b=H1b, H2=H2, H2b=H2b, W1=W1, W2=W2, W3=W3)


def loss_grad(z, y, bf=True, tau=None):
    """Mean squared error and its gradient 2 (z - y) / n; with tau the pinball
    loss and its gradient -(tau - [y < z]) / n (the subgradient at u = 0 taken
    from the u >= 0 side)."""
    c = f32 if bf else (lambda v: np.asarray(v, np.float64))
    n = z.shape[0]
    if tau is not None:
        u = y - z
        w = np.where(u < 0, tau - 1.0, tau)
        return float(np.mean(u * w)), c(-w / n)
    d = z - y
    return float(np.mean(d * d)), c(2.0 * d / n)


def backward(cache, g, bf=True):
    """Gradients of the loss w.r.t. W1, b1, W2, b2, W3, b3 given dL/dz = g."""
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    c = f32 if bf else (lambda v: np.asarray(v, np.float64))
    gb = r(g)[:, None]                                   # B x 1
    dW3 = gb.T @ cache["H2b"]                            # 1 x H
    db3 = np.array([np.sum(g)])
    G2 = c(gb @ cache["W3"]) * (cache["H2"] > 0)         # B x H
    G2b = r(G2)
    dW2 = G2b.T @ cache["H1b"]                           # H x H
    db2 = G2.sum(0)
    G1 = c(G2b @ cache["W2"]) * (cache["H1"] > 0)        # B x H
    G1b = r(G1)
    dW1 = G1b.T @ cache["Xb"]                            # H x F
    db1 = G1.sum(0)
    return dict(W1=dW1, b1=db1, W2=dW2, b2=db2, W3=dW3, b3=db3)


def loss(p, X, y, bf=False, tau=None):
    z, _ = forward(p, X, bf)
    if tau is not None:
        u = y - z
        return float(np.mean(u * np.where(u < 0, tau - 1.0, tau)))
    return float(np.mean((z - y) ** 2))


def init(F=64, H=256, seed=0):
    """He-normal weights, zero biases (the GPU model's initialisation law)."""
    rng = np.random.default_rng(seed)
    return dict(W1=rng.standard_normal((H, F)) * np.sqrt(2.0 / F), b1=np.zeros(H),
                W2=rng.standard_normal((H, H)) * np.sqrt(2.0 / H), b2=np.zeros(H),
                W3=rng.standard_normal((1, H)) * np.sqrt(1.0 / H), b3=np.zeros(1))


def coverage_label(coarse_order, nn_lists):
    """Smallest nprobe whose first lists cover every true neighbour's list:
    1 + max over the neighbours of the list's rank in the query's coarse order."""
    pos = np.empty_like(coarse_order)
    pos[coarse_order] = np.arange(coarse_order.shape[0])
    return int(pos[np.asarray(nn_lists)].max()) + 1
