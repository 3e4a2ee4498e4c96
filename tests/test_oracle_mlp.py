"""Pin the CPU reference of the dt0/method MLP: its analytic gradient against
central finite differences (float64, no BF16 rounding), and the BF16
rounding helper against known values."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_ref  # noqa: E402


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.00390625, 1.001953125, 1.0058593750, -3.14159265, 65504.0])
    got = mlp_ref.bf16(vals)
    # spacing at 1 is 2^-7: 1 + 2^-8 is a tie -> even (1.0); 1 + 2^-9 rounds down;
    # 1 + 3*2^-9 is past the midpoint -> 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0
    assert got[3] == 1.0078125
    assert abs(got[4] + 3.140625) < 1e-12
    assert got[5] == 65536.0


def test_gelu_grad_matches_fd():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (mlp_ref.gelu(z + h) - mlp_ref.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp_ref.gelu_grad(z), fd, rtol=1e-7, atol=1e-8)


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    B = 37
    X = rng.standard_normal((B, 16)).astype(np.float32)
    mean = np.zeros(16, np.float32)
    std = np.ones(16, np.float32)
    target = rng.standard_normal(B)
    label = (rng.uniform(size=B) > 0.5).astype(np.float64)
    cost = 8.0 + rng.standard_normal(B)
    never = (rng.uniform(size=B) > 0.6).astype(np.float64)
    theta = mlp_ref.init_params(1).astype(np.float64)
    theta += 0.01 * rng.standard_normal(theta.size)  # nonzero biases
    kw = dict(exact=True, cost=cost, w_cost=0.4, never=never, w_never=0.6)
    L0, g, _ = mlp_ref.loss_and_grad(theta, X, target, label, mean, std, 1.0, 0.7, **kw)
    idx = rng.choice(theta.size, 60, replace=False)
    # the four output biases, entries of the last two output columns of W4,
    # and W1/b1 entries
    P = mlp_ref.NPARAM
    idx = np.concatenate([idx, [P - 1, P - 2, P - 3, P - 4, P - 5, P - 6, P - 9, P - 10,
                                1024, 1100]])
    h = 1e-6
    for i in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mlp_ref.loss(tp, X, target, label, mean, std, 1.0, 0.7, **kw)
              - mlp_ref.loss(tm, X, target, label, mean, std, 1.0, 0.7, **kw)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    assert np.isfinite(L0)
    # without cost and never targets those heads get no gradient (bias and
    # W4 column: the last entry of each W4 row for the never head)
    _, g0, _ = mlp_ref.loss_and_grad(theta, X, target, label, mean, std, 1.0, 0.7, exact=True)
    assert g0[P - 1] == 0.0 and g0[P - 2] == 0.0 and g0[P - 5] == 0.0 and g0[P - 6] == 0.0
