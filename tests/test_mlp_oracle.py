"""Pins the FP64 numpy MLP oracle with central finite differences (CPU)."""
import numpy as np

import mlp_ref


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mlp_ref.NPARAMS) * 0.2
    X = rng.standard_normal((7, 16))
    T = rng.standard_normal((7, 2))
    _, g = mlp_ref.loss_grad(p, X, T)
    idx = rng.choice(mlp_ref.NPARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp_ref.loss(p + e, X, T) - mlp_ref.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_mlp_oracle_gelu_values():
    # GELU(x) = x Phi(x): Phi(0) = 1/2, Phi(1) = 0.8413447460685429
    np.testing.assert_allclose(mlp_ref.gelu(np.array([0.0, 1.0, -1.0])),
                               [0.0, 0.8413447460685429, -0.15865525393145707], rtol=1e-15)
