"""Pins of the FP64 rank-predictor restatement: the analytic gradient against
central finite differences, and the GELU against its closed form at a few
points."""
import numpy as np

import mlp_ref as M


def test_param_count_matches_layout():
    assert M.NPARAM == 18 * 64 + 64 + 2 * (64 * 64 + 64) + 64 + 1


def test_gelu_known_values():
    assert M.gelu(np.array(0.0)) == 0.0
    # large |x|: gelu(x) -> max(x, 0)
    assert abs(M.gelu(np.array(10.0)) - 10.0) < 1e-12
    assert abs(M.gelu(np.array(-10.0))) < 1e-12
    # gelu'(0) = 1/2
    assert abs(M.gelu_grad(np.array(0.0)) - 0.5) < 1e-15
    xs = np.linspace(-3, 3, 13)
    fd = (M.gelu(xs + 1e-6) - M.gelu(xs - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - M.gelu_grad(xs))) < 1e-8


def test_gradient_finite_differences():
    rng = np.random.default_rng(3)
    p = M.init(1)
    X = rng.standard_normal((7, M.NIN))
    Y = rng.standard_normal(7)
    _, g = M.loss_grad(p, X, Y)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [M.NPARAM - 1, 0, 18 * 64 + 3]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        lp, _ = M.loss_grad(p + e, X, Y)
        lm, _ = M.loss_grad(p - e, X, Y)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_adam_decreases_loss():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((64, M.NIN))
    Y = np.sin(X[:, 0]) + 0.5 * X[:, 1]
    p = M.init(2)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0, _ = M.loss_grad(p, X, Y)
    for t in range(1, 201):
        _, g = M.loss_grad(p, X, Y)
        M.adam(p, g, m, v, 1e-3, t)
    l1, _ = M.loss_grad(p, X, Y)
    assert l1 < 0.2 * l0
