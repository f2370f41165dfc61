"""The MLP oracle itself (CPU): its analytic gradient against central finite
differences of its own loss, the bfloat16 rounding against known values, and
the parameter layout."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even keeps 1
    assert mlp_ref.bf16(1.0 + 2.0 ** -8) == 1.0
    assert mlp_ref.bf16(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7
    assert mlp_ref.bf16(1.0 + 2.0 ** -7 + 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie, rounds to even
    assert mlp_ref.bf16(-3.0) == -3.0
    x = np.random.default_rng(1).normal(size=1000)
    r = mlp_ref.bf16(x)
    assert np.all(np.abs(r - x) <= 2.0 ** -8 * np.abs(x) + 1e-300)


def test_pack_unpack_roundtrip():
    th = np.arange(mlp_ref.NPARAM, dtype=np.float64)
    p = mlp_ref.unpack(th)
    assert p["W1"][1, 0] == 64 and p["b1"][0] == 512 and p["W2"][0, 0] == 576 and p["b4"][0] == 8960
    assert np.array_equal(mlp_ref.pack(p), th)


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = mlp_ref.init_params(3).astype(np.float64)
    theta[512:576] = rng.normal(0, 0.1, 64)   # nonzero biases exercise their gradients
    theta[8960] = 0.3
    X = rng.normal(size=(40, 8))
    T = rng.normal(size=40)
    loss, g = mlp_ref.loss_grad(theta, X, T, bf16_mode=False)
    idx = np.concatenate([rng.choice(512, 12, replace=False), 512 + rng.choice(64, 4, replace=False),
                          576 + rng.choice(4096, 12, replace=False), 4672 + rng.choice(64, 4, replace=False),
                          4736 + rng.choice(4096, 12, replace=False), 8832 + rng.choice(64, 4, replace=False),
                          8896 + rng.choice(64, 6, replace=False), [8960]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mlp_ref.loss_grad(tp, X, T, False)[0] - mlp_ref.loss_grad(tm, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_a_plausible_backward_mistake_is_caught():
    """Dropping the SiLU derivative of one layer changes the gradient far
    beyond the tolerance of the finite-difference check above."""
    rng = np.random.default_rng(2)
    theta = mlp_ref.init_params(5).astype(np.float64)
    X = rng.normal(size=(30, 8))
    T = rng.normal(size=30)
    _, g = mlp_ref.loss_grad(theta, X, T, bf16_mode=False)
    orig = mlp_ref.dsilu
    try:
        mlp_ref.dsilu = lambda z: np.ones_like(z)
        _, gbad = mlp_ref.loss_grad(theta, X, T, bf16_mode=False)
    finally:
        mlp_ref.dsilu = orig
    assert np.linalg.norm(gbad - g) > 1e-2 * np.linalg.norm(g)
