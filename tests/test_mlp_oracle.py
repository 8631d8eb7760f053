"""The numpy MLP oracle's backward pass against central finite differences
of its own loss (no GPU)."""
import numpy as np

import mlp


def _setup(seed=0, n=37):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(mlp.N_PARAMS) * 0.2
    X = rng.standard_normal((n, 8)) * [0.3, 0.3, 2.0, 0.5, 0.5, 0.4, 1.0, 1.0] + [0.8, 0.8, 3.0, 0, 0, 0.6, 4.5, 4.5]
    y = rng.normal(3.0, 0.4, n)
    norm = np.concatenate([X.mean(0), X.std(0), [y.mean(), y.std()]])
    return p, norm, X, y


def test_gradient_matches_finite_differences():
    p, norm, X, y = _setup()
    L, g = mlp.loss_grad(p, norm, X, y)
    assert abs(L - mlp.loss(p, norm, X, y)) <= 1e-12
    rng = np.random.default_rng(1)
    idx = np.concatenate([[mlp.O_B3], rng.choice(mlp.N_PARAMS, 60, replace=False)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(p + e, norm, X, y) - mlp.loss(p - e, norm, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_gelu_derivative():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp.gelu_d(z), fd, atol=1e-8)
    # GELU(0) = 0, GELU(x) -> x for large x, -> 0 for very negative x
    assert mlp.gelu(np.array([0.0]))[0] == 0 and abs(mlp.gelu(np.array([8.0]))[0] - 8) < 1e-12


def test_heldout_sampling_invariants(oracle):
    for m, h, r in [(113, 8, 30), (20, 8, 12), (50, 0, 10), (9, 3, 6)]:
        hp, rows = oracle.sample_heldout(42, 7, 0, m, h, r)
        assert hp.size == h and rows.size == r
        assert np.all(np.diff(hp) > 0) and np.all(np.diff(rows) > 0)
        assert not set(hp) & set(rows)
        assert (hp.size == 0 or hp.max() < m) and rows.max() < m
