"""Pins the CPU oracle of the quadrature-along-a-mode operators and of entry
evaluation (oracle/ptt_oracle_quad.c): against golden vectors written by the
compiled reference (pattac_marginal_trap / _cusum / pattac_scale_weight_sum /
pattac_apply_marginal / paratt_entry, paratt.c:471-499, 1252-1618), and
against brute-force contractions of the expanded tensor, which a dropped
term, a wrong sign or a shifted index would fail."""
import numpy as np
import pytest

import oracle_lib as O

G = O.golden()
D = 4


def golden_x():
    return O.HostTT.from_json(G["quad_x"])


@pytest.mark.parametrize("mode", range(D))
def test_scans_bit_exact_vs_reference(mode):
    x = golden_x()
    w = np.array(G["quad_w_%d" % mode])
    s = np.array(G["quad_s_%d" % mode])
    for kind, key, sc in ((1, "quad_trap_%d", None), (0, "quad_cusum_%d", None),
                          (2, "quad_sws_%d", s)):
        y = O.mode_scan(x, mode, kind, w, sc)
        got = y.cores[mode][: x.core_size(mode)]
        assert np.array_equal(got, np.array(G[key % mode])), (kind, mode)


@pytest.mark.parametrize("mode", range(D))
def test_marginal_bit_exact_vs_reference(mode):
    x = golden_x()
    w = np.array(G["quad_w_%d" % mode])
    m = O.apply_marginal(x, mode, w)
    ref = O.HostTT.from_json(G["quad_marginal_%d" % mode])
    assert m.n == ref.n and m.r == ref.r
    for k in range(m.d):
        assert np.array_equal(m.cores[k][: m.core_size(k)], ref.cores[k][: ref.core_size(k)]), k


def test_entry_vs_reference():
    # the reference's chain runs through BLAS dgemv, whose summation order
    # is the library's own: agreement to a few ulps is the bar
    x = golden_x()
    e = G["quad_entry"]
    for idx, val in zip(e["idx"], e["val"]):
        v = O.entry(x, idx)
        assert abs(v - val) <= 4e-16 * max(1.0, abs(val)) * 8, (idx, v, val)


def _trap_full(T, axis, w):
    T = np.moveaxis(T, axis, 0)
    out = np.zeros_like(T)
    for j in range(1, T.shape[0]):
        out[j] = out[j - 1] + 0.5 * w[j] * (T[j - 1] + T[j])
    return np.moveaxis(out, 0, axis)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_scans_vs_full_tensor(mode):
    rng = np.random.default_rng(5 + mode)
    n, r = [4, 5, 3], [1, 2, 3, 1]
    x = O.gauss_tt(n, r, 90 + mode)
    T = x.full()
    w = rng.standard_normal(n[mode])
    s = rng.standard_normal(n[mode])
    shape = [1] * 3
    shape[mode] = n[mode]
    cus = np.cumsum(T * w.reshape(shape), axis=mode)
    trap = _trap_full(T, mode, w)
    sws = trap + T * s.reshape(shape)
    for kind, want, sc in ((0, cus, None), (1, trap, None), (2, sws, s)):
        got = O.mode_scan(x, mode, kind, w, sc).full()
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max()), kind


def test_trapezoid_exact_on_linear_function():
    # the trapezoid rule integrates a linear function exactly:
    # int_0^{x_j} (a + b x) dx = a x_j + b x_j^2 / 2
    n, h, a, b = 33, 0.125, 0.7, -1.9
    xs = h * np.arange(n)
    f = a + b * xs
    x = O.HostTT([n, 2], [1, 1, 1], [f.copy(), np.array([1.0, 2.0])])
    y = O.mode_scan(x, 0, 1, np.full(n, h))
    want = a * xs + 0.5 * b * xs ** 2
    assert np.allclose(y.cores[0][:n], want, rtol=0, atol=1e-13)
    assert y.cores[0][0] == 0.0
    assert np.array_equal(y.cores[1][:2], x.cores[1][:2])


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_marginal_vs_full_tensor(mode):
    rng = np.random.default_rng(11 + mode)
    n, r = [3, 4, 2, 5], [1, 2, 3, 2, 1]
    x = O.gauss_tt(n, r, 70 + mode)
    w = rng.standard_normal(n[mode])
    want = np.tensordot(x.full(), w, axes=([mode], [0]))
    m = O.apply_marginal(x, mode, w)
    assert m.n == [n[k] for k in range(4) if k != mode]
    got = m.full()
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def test_entry_vs_full_tensor():
    n, r = [3, 4, 2, 5, 3], [1, 2, 3, 4, 2, 1]
    x = O.gauss_tt(n, r, 123)
    T = x.full()
    rng = np.random.default_rng(3)
    for _ in range(20):
        idx = [int(rng.integers(0, nk)) for nk in n]
        assert np.isclose(O.entry(x, idx), T[tuple(idx)], rtol=1e-13, atol=1e-15)


def test_entry_single_mode():
    x = O.HostTT([6], [1, 1], [np.arange(6, dtype=float) * 0.5])
    assert O.entry(x, [4]) == 2.0
