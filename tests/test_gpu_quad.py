"""GPU parity of the quadrature operators, marginals, entries and the
plain-norm rounding/norm (paratt.c:283-366, 471-499, 1252-1618) against the
compiled reference's golden vectors and the CPU oracle.  The scans,
marginals and entries sum in the reference's order with FP contraction off:
bit-exact against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import paratt_amd as P

pytestmark = pytest.mark.gpu
G = O.golden()


def dev(host):
    return P.DevTT.from_host(host)


def cores_of(t):
    return t.download_cores()


@pytest.mark.parametrize("mode", range(4))
def test_scans_bit_exact_vs_reference(mode):
    x = O.HostTT.from_json(G["quad_x"])
    w = np.array(G["quad_w_%d" % mode])
    s = np.array(G["quad_s_%d" % mode])
    for fn, key in ((lambda t: P.marginal_trap(t, mode, w), "quad_trap_%d"),
                    (lambda t: P.marginal_cusum(t, mode, w), "quad_cusum_%d"),
                    (lambda t: P.scale_weight_sum(t, mode, s, w), "quad_sws_%d")):
        t = dev(x)
        fn(t)
        got = cores_of(t)
        assert np.array_equal(got[mode], np.array(G[key % mode])), (key, mode)
        for k in range(4):
            if k != mode:
                assert np.array_equal(got[k], x.cores[k][: x.core_size(k)])
        t.free()


@pytest.mark.parametrize("mode", range(4))
def test_marginal_bit_exact_vs_reference(mode):
    x = O.HostTT.from_json(G["quad_x"])
    w = np.array(G["quad_w_%d" % mode])
    t = dev(x)
    m = P.apply_marginal(t, mode, w)
    ref = O.HostTT.from_json(G["quad_marginal_%d" % mode])
    assert m.n == ref.n and m.r == ref.r
    got = cores_of(m)
    for k in range(m.d):
        assert np.array_equal(got[k], ref.cores[k][: ref.core_size(k)]), k
    # the input is unchanged (the reference restores its sizes)
    assert t.n == x.n and t.r == x.r
    for k, c in enumerate(cores_of(t)):
        assert np.array_equal(c, x.cores[k][: x.core_size(k)])
    m.free()
    t.free()


def test_entry_vs_reference_and_oracle():
    x = O.HostTT.from_json(G["quad_x"])
    t = dev(x)
    e = G["quad_entry"]
    for idx, val in zip(e["idx"], e["val"]):
        v = P.entry(t, idx)
        assert v == O.entry(x, idx), idx
        assert abs(v - val) <= 3.2e-15 * max(1.0, abs(val)), (idx, v, val)
    t.free()


SCAN_SHAPES = [
    ([64, 256, 33], [1, 48, 96, 1]),
    ([1, 7, 5], [1, 3, 2, 1]),        # a mode of length 1
    ([300, 2], [1, 1, 1]),            # single fiber per core
    ([16, 1024, 8], [1, 16, 64, 1]),  # long fibers, 1024 of them
]


@pytest.mark.parametrize("shape", range(len(SCAN_SHAPES)))
def test_scans_bit_exact_vs_oracle(shape):
    n, r = SCAN_SHAPES[shape]
    x = O.gauss_tt(n, r, 300 + shape)
    rng = np.random.default_rng(shape)
    t0 = dev(x)
    for mode in range(len(n)):
        w = rng.standard_normal(n[mode])
        s = rng.standard_normal(n[mode])
        for kind in (0, 1, 2):
            t = dev(x)
            if kind == 0:
                P.marginal_cusum(t, mode, w)
            elif kind == 1:
                P.marginal_trap(t, mode, w)
            else:
                P.scale_weight_sum(t, mode, s, w)
            want = O.mode_scan(x, mode, kind, w, s if kind == 2 else None)
            got = cores_of(t)[mode]
            assert np.array_equal(got, want.cores[mode][: x.core_size(mode)]), (mode, kind)
            t.free()
    t0.free()


@pytest.mark.parametrize("n,r", [
    ([8, 40, 12, 30, 5], [1, 16, 64, 33, 7, 1]),
    ([3, 4], [1, 5, 1]),
    ([128, 64, 96], [1, 64, 64, 1]),
])
def test_marginal_bit_exact_vs_oracle(n, r):
    x = O.gauss_tt(n, r, 17 + len(n))
    rng = np.random.default_rng(len(n))
    t = dev(x)
    for mode in range(len(n)):
        w = rng.standard_normal(n[mode])
        m = P.apply_marginal(t, mode, w)
        want = O.apply_marginal(x, mode, w)
        assert m.n == want.n and m.r == want.r
        for k, c in enumerate(cores_of(m)):
            assert np.array_equal(c, want.cores[k][: want.core_size(k)]), (mode, k)
        m.free()
    t.free()


def test_entry_long_chain_and_single_mode():
    # d = 100 crosses the 48-core launch chunks of the entry chain
    d = 100
    n = [3 + (k % 5) for k in range(d)]
    r = [1] + [2 + (k % 4) for k in range(d - 1)] + [1]
    x = O.gauss_tt(n, r, 2024, scale=0.5)
    t = dev(x)
    rng = np.random.default_rng(9)
    for _ in range(8):
        idx = [int(rng.integers(0, nk)) for nk in n]
        assert P.entry(t, idx) == O.entry(x, idx)
    t.free()
    x1 = O.HostTT([9], [1, 1], [np.linspace(-1, 1, 9)])
    t1 = dev(x1)
    assert P.entry(t1, [7]) == x1.cores[0][7]
    t1.free()


def test_entry_wide_ranks():
    n, r = [4, 6, 5], [1, 256, 200, 1]
    x = O.gauss_tt(n, r, 77, scale=0.1)
    t = dev(x)
    T = x.full()
    for idx in ([0, 0, 0], [3, 5, 4], [2, 1, 3]):
        v = P.entry(t, idx)
        assert v == O.entry(x, idx)
        assert np.isclose(v, T[tuple(idx)], rtol=1e-12, atol=1e-14)
    t.free()


def test_truncate_plain_vs_oracle():
    # attac_tt_truncate: absolute tolerance in the plain Frobenius norm
    n, r = [6, 7, 5, 6], [1, 4, 6, 5, 1]
    x = O.gauss_tt(n, r, 55)
    y = O.gauss_tt(n, r, 56)
    z = O.axpby(1.0, x, 1e-9, y)
    nz = O.fro(z)
    for tol, cap in ((1e-7 * nz, None), (1e-12 * nz, None), (0.0, [1, 3, 3, 3, 1])):
        want = O.truncate_plain(z, tol, cap)
        t = dev(z)
        P.truncate_plain(t, tol, cap if cap is not None else z.r)
        got = O.HostTT(t.n, t.r, cores_of(t))
        assert got.r == want.r, (tol, got.r, want.r)
        assert O.rel_diff(got, want) < 1e-12
        t.free()


def test_norm_plain():
    n, r = [6, 7, 5, 6], [1, 4, 6, 5, 1]
    x = O.gauss_tt(n, r, 58)
    t = dev(x)
    v = P.norm_plain(t)
    assert abs(v - O.fro(x)) <= 1e-13 * O.fro(x)
    # x is left-orthogonalised in place and still represents the same tensor
    got = O.HostTT(t.n, t.r, cores_of(t))
    assert O.rel_diff(got, x) < 1e-13
    t.free()


def test_send_recv_self_exchange():
    L = P.lib()
    src = (C.c_double * 5)(1.5, -2.0, 3.25, 0.0, 9.0)
    dst = (C.c_double * 5)()
    sreq, rreq = C.c_int(7), C.c_int(7)
    MPI_DOUBLE = 0x4C00080B
    L.send_recv_above(src, dst, 5, MPI_DOUBLE, C.byref(sreq), C.byref(rreq))
    assert list(dst) == list(src)
    assert sreq.value == rreq.value == 0x2C000000
    ints = (C.c_int * 3)(4, 5, 6)
    got = (C.c_int * 3)()
    L.send_recv_below(ints, got, 3, 0x4C000405, C.byref(sreq), C.byref(rreq))
    assert list(got) == [4, 5, 6]
