"""Pins of the oracle by mathematics (no reference run needed): brute-force
full-tensor checks on small grids, invariants and closed forms."""
import numpy as np
import pytest

import oracle_lib as O


def per_mode_stencil_full(T, sten, mode):
    """Periodic 3-point stencil along one axis of a full tensor."""
    return sum(s * np.roll(T, 1 - j, axis=mode) for j, s in enumerate(sten))


def test_lorth_orthonormal_and_invariant():
    x = O.gauss_tt([6, 5, 7, 4], [1, 4, 6, 3, 1], 5)
    y = O.lorth(x)
    assert np.allclose(y.full(), x.full(), rtol=0, atol=1e-13 * np.abs(x.full()).max())
    for k in range(x.d - 1):
        A = y.core(k).reshape(y.r[k] * y.n[k], y.r[k + 1], order="F")
        assert np.allclose(A.T @ A, np.eye(y.r[k + 1]), atol=1e-14)


def test_truncation_equals_full_tt_svd_small():
    # brute force: TT-SVD of the expanded tensor with the same per-bond rule
    x = O.gauss_tt([3, 4, 3, 4], [1, 2, 3, 2, 1], 8)
    y = O.gauss_tt([3, 4, 3, 4], [1, 2, 2, 2, 1], 9)
    z = O.axpby(1.0, x, 1e-3, y)
    tt = O.truncate(z, 1e-6 * O.norm(z))
    F = z.full()
    # with tol << spectrum gaps the rounded TT is the exact tensor
    assert np.allclose(tt.full(), F, atol=1e-12 * np.abs(F).max())
    t2 = O.truncate(z, 0.05 * O.norm(z))
    err = np.linalg.norm(t2.full() - F) / np.linalg.norm(F)
    assert err <= 0.05 + 1e-12


def test_axpby_hadamard_entries():
    x = O.gauss_tt([3, 4, 5], [1, 2, 3, 1], 1)
    y = O.gauss_tt([3, 4, 5], [1, 3, 2, 1], 2)
    np.testing.assert_allclose(O.axpby(0.3, x, -1.7, y).full(), 0.3 * x.full() - 1.7 * y.full(),
                               atol=1e-13)
    np.testing.assert_allclose(O.hadamard(x, y).full(), x.full() * y.full(), atol=1e-13)


def test_stencil_eigen_relation_on_sine():
    # periodic second difference of sin(2 pi m j / n) = -4 sin^2(pi m/n) * same
    n, m = 16, 3
    v = np.sin(2 * np.pi * m * np.arange(n) / n)
    x = O.HostTT([n, 5], [1, 1, 1], [v, np.ones(5)])
    y = O.periodic_stencil(x, 0, [1.0, -2.0, 1.0])
    lam = -4 * np.sin(np.pi * m / n) ** 2
    np.testing.assert_allclose(y.cores[0], lam * v, atol=1e-14)


def test_mode_sum_equals_sum_of_mode_stencils():
    x = O.gauss_tt([5, 6, 4, 5], [1, 2, 3, 2, 1], 3)
    st = [0.7, -1.3, 0.4]
    s = O.mode_sum(x, st)
    assert s.r == [1, 4, 6, 4, 1]
    F = x.full()
    ref = sum(per_mode_stencil_full(F, st, m) for m in range(x.d))
    np.testing.assert_allclose(s.full(), ref, atol=1e-13 * np.abs(ref).max())


def dense_rhs(U, nu, h):
    d = U.ndim
    d1 = [-0.5 / h, 0.0, 0.5 / h]
    d2 = [1 / h ** 2, -2 / h ** 2, 1 / h ** 2]
    grad = sum(per_mode_stencil_full(U, d1, m) for m in range(d))
    lap = sum(per_mode_stencil_full(U, d2, m) for m in range(d))
    return -U * grad + nu * lap + U * (1 - U)


def test_rxn_burgers_rhs_matches_dense_grid():
    d, n = 3, 12
    h = 2 * np.pi / n
    u = O.gaussian_ic(d, n, centre=[np.pi, 2.5, 3.6])
    f = O.rxn_burgers_rhs(u, 0.01, h, 1e-12, 64)
    ref = dense_rhs(u.full(), 0.01, h)
    assert np.linalg.norm(f.full() - ref) <= 1e-10 * np.linalg.norm(ref)


def test_rxn_burgers_rk4_matches_dense_rk4():
    d, n, dt, nu = 3, 10, 1e-3, 0.01
    h = 2 * np.pi / n
    u = O.gaussian_ic(d, n, centre=[np.pi, 2.9, 3.3])
    U = u.full()
    for _ in range(3):
        k1 = dense_rhs(U, nu, h)
        k2 = dense_rhs(U + 0.5 * dt * k1, nu, h)
        k3 = dense_rhs(U + 0.5 * dt * k2, nu, h)
        k4 = dense_rhs(U + dt * k3, nu, h)
        U = U + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    tt = O.rxn_burgers_rk4(u, nu, h, dt, 1e-12, 64, 3)
    assert np.linalg.norm(tt.full() - U) <= 1e-10 * np.linalg.norm(U)


def test_rk4_linear_amplification_closed_form():
    # u_t = u(1-u) restricted to small amplitude ~ linear growth; check the
    # RK4 polynomial on the pure reaction part via a spatially constant field:
    # constant u: gradients and Laplacian vanish, u_t = u - u^2 (logistic).
    d, n, dt = 3, 8, 0.05
    h = 2 * np.pi / n
    c0 = 0.1
    u = O.HostTT([n] * d, [1] * (d + 1), [np.full(n, c0 ** (1.0 / d))] * d)
    tt = O.rxn_burgers_rk4(u, 0.01, h, dt, 1e-13, 8, 1)
    f = lambda v: v - v * v
    k1 = f(c0); k2 = f(c0 + dt / 2 * k1); k3 = f(c0 + dt / 2 * k2); k4 = f(c0 + dt * k3)
    expect = c0 + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    np.testing.assert_allclose(tt.full(), expect, rtol=1e-12)
