"""The float64 reference of the warm-start network (oracle/mlp_ref.py), pinned without a GPU:
its analytic gradients against central finite differences of its own loss, and its pieces against
closed forms (GELU values, bf16 rounding), so that the GPU gradient tests compare against a
checked reference."""
import numpy as np

from oracle import mlp_ref


def small_problem(seed=0):
    rng = np.random.default_rng(seed)
    dims = [6, 8, 8, 8, 3 + 5]
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    params = rng.standard_normal(n) * 0.4
    x = rng.uniform(0.0, 1.0, (7, dims[0]))
    tc = rng.standard_normal((7, 3)) * 0.3
    tn = (rng.random((7, 5)) < 0.5).astype(np.uint8)
    return dims, params, x, tc, tn


def total_loss(params, dims, x, tc, tn, wc, wl):
    (lc, ll), _ = mlp_ref.loss_and_grads(params, dims, x, tc, tn, wc, wl, emulate_bf16=False)
    return lc + ll


def test_gradients_match_finite_differences():
    dims, params, x, tc, tn = small_problem()
    wc, wl = 1.0, 0.1
    _, g = mlp_ref.loss_and_grads(params, dims, x, tc, tn, wc, wl, emulate_bf16=False)
    h = 1e-6
    fd = np.empty_like(params)
    for i in range(params.size):
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd[i] = (total_loss(p1, dims, x, tc, tn, wc, wl) - total_loss(p2, dims, x, tc, tn, wc, wl)) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd)))


def test_gelu_closed_forms():
    # tanh-GELU: 0 at 0, odd part x/2 * tanh(...), limits x and 0
    assert mlp_ref.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mlp_ref.gelu(np.array([10.0]))[0] - 10.0) < 1e-12
    assert abs(mlp_ref.gelu(np.array([-10.0]))[0]) < 1e-12
    x = np.linspace(-3, 3, 61)
    fd = (mlp_ref.gelu(x + 1e-6) - mlp_ref.gelu(x - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - mlp_ref.gelu_grad(x))) < 1e-8


def test_bf16_rounding():
    # exactly representable values are kept; ties round to even; 8 bits of mantissa
    assert mlp_ref.bf16(np.array([1.0, -2.5, 0.0]))[1] == -2.5
    one_ulp = 2.0 ** -7
    assert mlp_ref.bf16(np.array([1.0 + one_ulp / 2]))[0] == 1.0              # tie -> even
    assert mlp_ref.bf16(np.array([1.0 + 1.5 * one_ulp]))[0] == 1.0 + 2 * one_ulp   # tie -> even
    v = np.random.default_rng(3).standard_normal(1000)
    assert np.max(np.abs(mlp_ref.bf16(v) - v) / np.abs(v)) <= 2.0 ** -8
