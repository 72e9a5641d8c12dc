"""FP64 oracle of the chemistry surrogate MLP (TEST INFRASTRUCTURE ONLY:
loaded by tests/ to check the GPU forward, backward and Adam kernels).

Network: x -> [W_l, b_l, GELU] x 3 -> W_4, b_4 (linear output), loss the
mean of squared differences over the batch and the outputs.  GELU is the
exact form 0.5 z (1 + erf(z / sqrt 2)) (Hendrycks & Gimpel 2016).  Weights
W_l are [in][out] (x . W + b).  The surrogate has no counterpart in the
reference.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SQRT2_I = 1.0 / np.sqrt(2.0)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT2_I))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z * SQRT2_I)) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, x):
    """params: list of (W, b); returns (output, cache)."""
    a = x
    cache = [(a, None)]
    for li, (W, b) in enumerate(params):
        z = a @ W + b
        if li < len(params) - 1:
            a = gelu(z)
        else:
            a = z
        cache.append((a, z))
    return a, cache


def loss_and_grads(params, x, y):
    """Mean-square loss over all entries and its gradients (dW, db)."""
    out, cache = forward(params, x)
    n = y.size
    d = out - y
    loss = float(np.sum(d * d) / n)
    g = 2.0 * d / n
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        a_prev = cache[li][0]
        W, _ = params[li]
        grads[li] = (a_prev.T @ g, g.sum(axis=0))
        if li > 0:
            z_prev = cache[li][1]
            g = (g @ W.T) * gelu_d(z_prev)
    return loss, grads


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) in place; returns p."""
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    p[:] = p - lr * mh / (np.sqrt(vh) + eps)
    return p


def bf16_round(a):
    """Round float64 values to the nearest BF16 (ties to even)."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def element_projector(kncf, awt):
    """Projector onto dY with sum_k (a_ek W_e / W_k) dY_k = 0 for every
    element e (H, O, N): P = I - E^T (E E^T)^-1 E, with W_k = sum_e a_ek W_e
    so that the element constraints also keep sum_k dY_k = 0."""
    kncf = np.asarray(kncf, dtype=np.float64)
    awt = np.asarray(awt, dtype=np.float64)
    Wk = awt @ kncf
    E = kncf * awt[:, None] / Wk[None, :]
    return np.eye(E.shape[1]) - E.T @ np.linalg.solve(E @ E.T, E)
