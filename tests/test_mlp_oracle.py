"""Pins of the MLP oracle (CPU): exact gradient against finite differences, bfloat16
rounding against representable values and ties, Adam against its first-step closed form."""

import numpy as np
import pytest

from oracle import mlp_oracle as M

DIMS = (5, 7, 6, 1)


def _params(dims, seed):
    rng = np.random.default_rng(seed)
    n = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))
    return rng.normal(0.0, 0.7, n)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_central_differences(seed):
    rng = np.random.default_rng(100 + seed)
    p = _params(DIMS, seed)
    X = rng.normal(size=(9, DIMS[0]))
    t = rng.normal(size=9)
    _, g = M.loss_and_gradient(p, DIMS, X, t)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (M.loss_and_gradient(p + e, DIMS, X, t)[0] - M.loss_and_gradient(p - e, DIMS, X, t)[0]) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_forward_closed_form_one_hidden_unit():
    # 1 -> 1 -> 1: y = w2 * silu(w1 x + b1) + b2
    p = np.array([0.5, -0.25, 2.0, 0.125])
    X = np.array([[1.0], [-3.0]])
    y, _, _ = M.forward(p, (1, 1, 1), X)
    z = 0.5 * X[:, 0] - 0.25
    assert np.allclose(y, 2.0 * z / (1 + np.exp(-z)) + 0.125, rtol=0, atol=1e-15)


def test_bf16_rounding():
    # representable values are kept; halfway cases round to even; carries propagate
    exact = np.array([1.0, -2.5, 0.0078125, 3.0 * 2.0**-20, 65280.0])
    assert np.array_equal(M.to_bf16(exact), exact)
    one_ulp = 2.0**-7  # bf16 spacing at 1
    assert M.to_bf16(1.0 + one_ulp / 2) == 1.0              # tie -> even (1.0)
    assert M.to_bf16(1.0 + 3 * one_ulp / 2) == 1.0 + 2 * one_ulp  # tie -> even (upper)
    assert M.to_bf16(1.0 + 0.51 * one_ulp) == 1.0 + one_ulp
    assert M.to_bf16(2.0 - one_ulp / 4) == 2.0               # carry into the exponent
    rng = np.random.default_rng(3)
    x = rng.normal(size=1000) * 10.0 ** rng.integers(-5, 5, 1000)
    r = M.to_bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0**-8 * (1 + 1e-6))


def test_adam_first_step_is_sign_step():
    rng = np.random.default_rng(4)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m, v = M.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=1e-3, eps=0.0)
    assert np.allclose(p1, p - 1e-3 * np.sign(g), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_planted_backprop_mistake_is_caught():
    """Dropping the SiLU derivative from the back-propagation must fail the finite-difference
    check (guards the check itself)."""
    rng = np.random.default_rng(5)
    p = _params(DIMS, 7)
    X = rng.normal(size=(6, DIMS[0]))
    t = rng.normal(size=6)
    orig = M._silu_grad
    try:
        M._silu_grad = lambda z: np.ones_like(z)
        _, g = M.loss_and_gradient(p, DIMS, X, t)
    finally:
        M._silu_grad = orig
    h = 1e-6
    i = 0  # first-layer weight
    e = np.zeros_like(p)
    e[i] = h
    fd = (M.loss_and_gradient(p + e, DIMS, X, t)[0] - M.loss_and_gradient(p - e, DIMS, X, t)[0]) / (2 * h)
    assert abs(g[i] - fd) > 1e-6
