"""CPU pins of the surrogate's FP64 oracle (oracle/mlp_oracle.py) and of
the element projector: gradients against central finite differences, Adam
against its closed form, the projector against the element constraints."""
from __future__ import annotations

import numpy as np

from oracle import mlp_oracle as mo


def _net(rng, sizes):
    return [(rng.normal(0, 1 / np.sqrt(a), (a, b)), rng.normal(0, 0.1, b))
            for a, b in zip(sizes[:-1], sizes[1:])]


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    params = _net(rng, [12, 8, 8, 8, 10])
    x = rng.normal(size=(5, 12))
    y = rng.normal(size=(5, 10))
    _, grads = mo.loss_and_grads(params, x, y)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            flat = arr.reshape(-1)
            for idx in rng.choice(flat.size, min(6, flat.size), replace=False):
                old = flat[idx]
                flat[idx] = old + h
                lp, _ = mo.loss_and_grads(params, x, y)
                flat[idx] = old - h
                lm, _ = mo.loss_and_grads(params, x, y)
                flat[idx] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g.reshape(-1)[idx]) <= 1e-6 * max(1.0, abs(fd)), (li, idx)


def test_gelu_derivative_and_limits():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_d(z))) <= 1e-8
    assert abs(mo.gelu(np.array([10.0]))[0] - 10.0) <= 1e-12
    assert abs(mo.gelu(np.array([-10.0]))[0]) <= 1e-12


def test_adam_closed_form_first_step():
    # after one step from zero moments the update is lr * g / (|g| + eps)
    g = np.array([0.5, -2.0, 1e-3])
    p = np.zeros(3)
    mo.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159])
    r = mo.bf16_round(x)
    # spacing 2^-7 at 1: the halfway case ties to the even 1.0
    assert r.tolist() == [1.0, 1.0, 1.0 + 2 ** -7, -3.140625]


def test_element_projector(thd):
    from phd_amd.surrogate import element_projector
    P = element_projector(thd)
    kn = np.asarray(thd.kncf, float)
    aw = np.asarray(thd.awt, float)
    E = kn * aw[:, None] / (aw @ kn)[None, :]
    d = np.random.default_rng(1).normal(size=(9, 20))
    assert np.max(np.abs(E @ (P @ d))) <= 1e-14
    assert np.max(np.abs((P @ d).sum(axis=0))) <= 1e-14
    assert np.max(np.abs(P @ P - P)) <= 1e-14
    # the oracle's restatement agrees
    assert np.max(np.abs(mo.element_projector(thd.kncf, thd.awt) - P)) <= 1e-15
