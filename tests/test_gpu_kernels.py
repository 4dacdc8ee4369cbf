"""GPU parity of the building blocks: FP64 MFMA GEMM against numpy (fp64),
and the streaming TT operators against the CPU oracle / reference golden
vectors (bit-exact: the GPU kernels round exactly as the reference's loops)."""
import numpy as np
import pytest

import oracle_lib as O
import paratt_amd as P
from gpu_util import DevBuf, dgemm, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("M,N,K", [(7, 5, 3), (64, 64, 64), (130, 70, 33), (128, 128, 4096),
                                   (5000, 128, 128), (128, 3000, 96), (33, 200, 2100)])
def test_dgemm_matches_numpy(ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta == "T" else (M, K))
    B = rng.standard_normal((N, K) if tb == "T" else (K, N))
    Cm = rng.standard_normal((M, N))
    a, b, c = (DevBuf(np.asfortranarray(A).ravel(order="F")),
               DevBuf(np.asfortranarray(B).ravel(order="F")),
               DevBuf(np.asfortranarray(Cm).ravel(order="F")))
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    dgemm(ta, tb, M, N, K, 1.5, a, A.shape[0], b, B.shape[0], -0.5, c, M)
    got = c.get().reshape((M, N), order="F")
    ref = 1.5 * opA @ opB - 0.5 * Cm
    scale = np.abs(opA) @ np.abs(opB) + np.abs(Cm)
    assert np.max(np.abs(got - ref) / scale) < 1e-14 * max(K, 1) ** 0.5 * 4
    for bb in (a, b, c):
        bb.free()


def test_dgemm_in_place_tall_times_small():
    rng = np.random.default_rng(1)
    M, Kc = 20000, 96
    A = rng.standard_normal((M, Kc))
    T = np.triu(rng.standard_normal((Kc, Kc)))
    a, t = DevBuf(A.ravel(order="F")), DevBuf(T.ravel(order="F"))
    dgemm("N", "N", M, Kc, Kc, 1.0, a, M, t, Kc, 0.0, a, M)
    got = a.get().reshape((M, Kc), order="F")
    assert np.max(np.abs(got - A @ T)) < 1e-12 * np.max(np.abs(A @ T))


def _golden_x():
    return O.HostTT.from_json(O.golden()["ops_x"])


def test_axpby_bit_exact():
    g = O.golden()
    x, y = O.HostTT.from_json(g["ops_x"]), O.HostTT.from_json(g["ops_y"])
    dx, dy = P.DevTT.from_host(x), P.DevTT.from_host(y)
    z = P.axpby(0.5, dx, -2.0, dy)
    ref = O.HostTT.from_json(g["ops_axpby"])
    assert z.r == ref.r
    for k, c in enumerate(z.download_cores()):
        np.testing.assert_array_equal(c, ref.cores[k])


@pytest.mark.parametrize("rk", [1, 7, 64])
def test_axpby_random_shapes_bit_exact(rk):
    n = [5, 9, 4, 6]
    x = O.gauss_tt(n, [1, rk, rk + 1, 3, 1], 3)
    y = O.gauss_tt(n, [1, 2, rk, 5, 1], 4)
    z = P.axpby(1.25, P.DevTT.from_host(x), -0.75, P.DevTT.from_host(y))
    ref = O.axpby(1.25, x, -0.75, y)
    assert z.r == ref.r
    for k, c in enumerate(z.download_cores()):
        np.testing.assert_array_equal(c, ref.cores[k])


def test_hadamard_bit_exact():
    g = O.golden()
    x, y = O.HostTT.from_json(g["ops_x"]), O.HostTT.from_json(g["ops_y"])
    dx, dy = P.DevTT.from_host(x), P.DevTT.from_host(y)
    rz = [x.r[k] * y.r[k] for k in range(x.d + 1)]
    dz = P.DevTT(x.n, rz)
    import ctypes as C
    L = P.lib()
    xt, yt, zt = dx.tt.contents, dy.tt.contents, dz.tt.contents
    L.pattacHadamard(x.d, xt.dim, xt.ttRank, xt.coreList, yt.ttRank, yt.coreList, zt.ttRank,
                     zt.coreList)
    ref = O.HostTT.from_json(g["ops_hadamard"])
    for k, c in enumerate(dz.download_cores()):
        np.testing.assert_array_equal(c, ref.cores[k])


STEN = {3: [0.3, -1.1, 0.7], 5: [-1.0 / 12, 4.0 / 3, -2.5, 4.0 / 3, -1.0 / 12]}


@pytest.mark.parametrize("kind", ["pstencil", "sstencil", "cdfstencil"])
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("slen", [3, 5])
def test_stencils_bit_exact_vs_reference(kind, mode, slen):
    import ctypes as C
    x = _golden_x()
    dx = P.DevTT.from_host(x)
    s, sp = np.array(STEN[slen]), None
    sp = s.ctypes.data_as(C.POINTER(C.c_double))
    L = P.lib()
    if kind == "pstencil":
        L.pattac_apply_periodic_stencil(dx.p, mode, sp, slen)
    elif kind == "sstencil":
        L.attac_apply_periodic_stencil(dx.tt, mode, sp, slen)
    else:
        L.pattac_apply_cdf_stencil(dx.p, mode, sp, slen)
    ref = O.HostTT.from_json(O.golden()["ops_%s_m%d_s%d" % (kind, mode, slen)])
    for k, c in enumerate(dx.download_cores()):
        np.testing.assert_array_equal(c, ref.cores[k])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_flux_diag_bit_exact_vs_reference(mode):
    import ctypes as C
    g = O.golden()
    x = _golden_x()
    nn = x.n[mode]
    ws = 0.2 + 0.1 * np.arange(nn)
    src = 0.05 * np.arange(nn) - 0.1
    bw = 1.0 - 0.03 * np.arange(nn + 1)
    gcw = np.array([0.3, -0.45])
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    L = P.lib()
    dx = P.DevTT.from_host(x)
    L.pattac_apply_cdf_lf_flux(dx.p, mode, dp(ws), dp(gcw), dp(src), 0.01, 0.125)
    np.testing.assert_array_equal(dx.download_cores()[mode],
                                  O.HostTT.from_json(g["ops_lfflux_m%d" % mode]).cores[mode])
    dx = P.DevTT.from_host(x)
    L.pattac_apply_cdf_lw_rhs(dx.p, mode, dp(ws), dp(gcw), dp(bw), dp(src), 0.01, 0.125)
    np.testing.assert_array_equal(dx.download_cores()[mode],
                                  O.HostTT.from_json(g["ops_lwrhs_m%d" % mode]).cores[mode])
    dx = P.DevTT.from_host(x)
    L.attac_mult_diag_mat(dx.tt, mode, dp(ws))
    np.testing.assert_array_equal(dx.download_cores()[mode],
                                  O.HostTT.from_json(g["ops_diag_m%d" % mode]).cores[mode])


def test_scale_dot_norm_vs_reference():
    g = O.golden()
    x, y = O.HostTT.from_json(g["ops_x"]), O.HostTT.from_json(g["ops_y"])
    dx = P.DevTT.from_host(x)
    P.lib().attac_tt_scale(-1.75, dx.tt)
    np.testing.assert_array_equal(dx.download_cores()[0],
                                  O.HostTT.from_json(g["ops_scale"]).cores[0])
    dx, dy = P.DevTT.from_host(x), P.DevTT.from_host(y)
    v = P.dotprod(dx, dy)
    assert abs(v - g["ops_dotprod_xy"]) <= 1e-13 * abs(g["ops_dotprod_xy"])
    nv = P.norm(dx)
    assert abs(nv - g["ops_norm_x"]) <= 1e-13 * g["ops_norm_x"]
    # paratt_norm leaves x left-orthogonalised (same tensor, reference gauge)
    after = to_host(dx)
    ref_after = O.HostTT.from_json(g["ops_norm_x_after"])
    assert after.r == ref_after.r
    assert O.rel_diff(after, ref_after) < 1e-13


def test_gauss_fill_matches_cpu_generator():
    n, r = [6, 7, 5], [1, 3, 4, 1]
    dt = P.DevTT.gaussian(n, r, 77, 0.5)
    ref = O.gauss_tt(n, r, 77, 0.5)
    for k, c in enumerate(dt.download_cores()):
        np.testing.assert_allclose(c, ref.cores[k], rtol=1e-13, atol=1e-15)
