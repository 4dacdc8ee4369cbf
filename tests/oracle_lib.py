"""ctypes binding of the CPU oracle (oracle/liboracle.so) plus host-side TT
helpers.  Test infrastructure only: the oracle is the checker, never the
thing measured or shipped.

A host TT is `HostTT(n, r, cores)`: `cores[k]` is a flat float64 array in the
reference's column-major layout (element (a, j, b) of core k at
a + r_k*(j + n_k*b)).
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "golden_ref.json")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = C.CDLL(ORACLE_SO)
        _lib.o_inner.restype = C.c_double
        _lib.o_dotprod.restype = C.c_double
        _lib.o_norm.restype = C.c_double
    return _lib


_golden = None


def golden():
    global _golden
    if _golden is None:
        with open(GOLDEN) as f:
            _golden = json.load(f)
    return _golden


def _ip(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptrs(cores):
    arr = (C.POINTER(C.c_double) * len(cores))()
    for i, c in enumerate(cores):
        arr[i] = c.ctypes.data_as(C.POINTER(C.c_double))
    return arr


class HostTT:
    def __init__(self, n, r, cores):
        self.n = [int(v) for v in n]
        self.r = [int(v) for v in r]
        self.cores = [np.ascontiguousarray(c, dtype=np.float64) for c in cores]
        assert len(self.r) == len(self.n) + 1
        for k in range(self.d):
            assert self.cores[k].size >= self.core_size(k), (k, self.cores[k].size)

    @property
    def d(self):
        return len(self.n)

    def core_size(self, k):
        return self.r[k] * self.n[k] * self.r[k + 1]

    def core(self, k):
        """Core k as an (r_k, n_k, r_{k+1}) array view (Fortran order)."""
        return self.cores[k][: self.core_size(k)].reshape(
            (self.r[k], self.n[k], self.r[k + 1]), order="F")

    def copy(self, capacity=None):
        cores = []
        for k in range(self.d):
            sz = self.core_size(k)
            cap = max(sz, capacity[k] if capacity else sz)
            c = np.zeros(cap)
            c[:sz] = self.cores[k][:sz]
            cores.append(c)
        return HostTT(self.n, self.r, cores)

    def trimmed(self):
        return HostTT(self.n, self.r, [self.cores[k][: self.core_size(k)].copy()
                                       for k in range(self.d)])

    def full(self):
        t = self.core(0).reshape(self.n[0], self.r[1])
        for k in range(1, self.d):
            g = self.core(k)
            t = np.tensordot(t, g, axes=([t.ndim - 1], [0]))
        return t.reshape(self.n, order="C")

    def entry(self, idx):
        v = np.ones(1)
        for k in range(self.d):
            v = v @ self.core(k)[:, idx[k], :]
        return float(v[0])

    @staticmethod
    def from_json(obj):
        return HostTT(obj["n"], obj["r"], [np.array(c, dtype=np.float64) for c in obj["cores"]])


def gauss_tt(n, r, seed, scale=1.0):
    L = lib()
    cores = []
    for k in range(len(n)):
        cnt = r[k] * n[k] * r[k + 1]
        c = np.empty(max(cnt, 1))
        L.o_gauss_fill(C.c_uint64(seed), C.c_int(k), C.c_double(scale), C.c_long(cnt),
                       c.ctypes.data_as(C.POINTER(C.c_double)))
        cores.append(c[:cnt] if cnt else c)
    return HostTT(n, r, cores)


def uniform_ranks(d, rank):
    return [1] + [rank] * (d - 1) + [1]


# ---------------------------------------------------------------- oracle ops
def axpby(a, x, b, y):
    d = x.d
    if d < 2:
        # the formal sum needs two boundary cores (attac.c:81-112 writes
        # [x | y] into a one-core train's single n-element core)
        raise ValueError("formal axpby needs at least two modes")
    rz = [1] + [x.r[k] + y.r[k] for k in range(1, d)] + [1]
    z = HostTT(x.n, rz, [np.zeros(rz[k] * x.n[k] * rz[k + 1]) for k in range(d)])
    rzb = _ip(rz)
    lib().o_axpby(C.c_int(d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)), C.c_double(a),
                  _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(x.cores), C.c_double(b),
                  _ip(y.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores),
                  rzb.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(z.cores))
    return z


def hadamard(x, y):
    d = x.d
    rz = [x.r[k] * y.r[k] for k in range(d + 1)]
    z = HostTT(x.n, rz, [np.zeros(rz[k] * x.n[k] * rz[k + 1]) for k in range(d)])
    lib().o_hadamard(C.c_int(d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
                     _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(x.cores),
                     _ip(y.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores), _ptrs(z.cores))
    return z


def lorth(x):
    """Left-orthogonalise a copy of x (pattac.c:244-294)."""
    y = x.copy()
    r = _ip(y.r)
    lib().o_lorth(C.c_int(y.d), _ip(y.n).ctypes.data_as(C.POINTER(C.c_int)),
                  r.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores))
    y.r = [int(v) for v in r]
    return y.trimmed()


def truncate(x, tol, maxr=None, log=False):
    """paratt_truncate on a copy of x; returns (tt, sig_log, numrank_log)."""
    y = x.copy()
    d = y.d
    r = _ip(y.r)
    mr = _ip(maxr if maxr is not None else y.r)
    ld = max(y.r) + 1
    sig = np.zeros((d + 1) * ld)
    nr = np.zeros(d + 1, dtype=np.int32)
    lib().o_truncate(C.c_int(d), _ip(y.n).ctypes.data_as(C.POINTER(C.c_int)),
                     r.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores),
                     mr.ctypes.data_as(C.POINTER(C.c_int)), C.c_double(tol),
                     sig.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(ld),
                     nr.ctypes.data_as(C.POINTER(C.c_int)))
    y.r = [int(v) for v in r]
    out = y.trimmed()
    if log:
        return out, sig.reshape(d + 1, ld), nr
    return out


def inner(x, y):
    return lib().o_inner(C.c_int(x.d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
                         _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(x.cores),
                         _ip(y.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores))


def dotprod(x, y):
    """paratt_dotprod on copies (the reference mutates its arguments)."""
    xc, yc = x.copy(), y.copy()
    return lib().o_dotprod(C.c_int(x.d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
                           _ip(xc.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(xc.cores),
                           _ip(yc.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(yc.cores),
                           C.c_int(0))


def norm(x):
    """paratt_norm of a copy: ||x||_F / sqrt(prod n)."""
    y = x.copy()
    r = _ip(y.r)
    return lib().o_norm(C.c_int(y.d), _ip(y.n).ctypes.data_as(C.POINTER(C.c_int)),
                        r.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores))


def fro(x):
    """Plain Frobenius norm ||x||_F = paratt_norm * sqrt(prod n)."""
    return norm(x) * float(np.sqrt(np.prod(np.array(x.n, dtype=np.float64))))


def rel_diff(a, b):
    """||a - b||_F / ||b||_F via the formal difference TT (no cancellation)."""
    if a.d == 1:
        dn = float(np.linalg.norm(a.core(0).ravel() - b.core(0).ravel()) / np.sqrt(a.n[0]))
        bn = norm(b)
        return dn / bn if bn > 0 else dn
    dn = norm(axpby(1.0, a, -1.0, b))
    bn = norm(b)
    return dn / bn if bn > 0 else dn


def _core_op(fn, x, mode, *args):
    y = x.copy()
    c = y.cores[mode]
    fn(C.c_int(y.r[mode]), C.c_int(y.n[mode]), C.c_int(y.r[mode + 1]), *args,
       c.ctypes.data_as(C.POINTER(C.c_double)))
    return y


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def periodic_stencil(x, mode, sten, serial=False):
    s, sp = _dp(sten)
    f = lib().o_periodic_stencil_serial if serial else lib().o_periodic_stencil
    return _core_op(f, x, mode, sp, C.c_int(len(sten)))


def cdf_stencil(x, mode, sten):
    s, sp = _dp(sten)
    return _core_op(lib().o_cdf_stencil, x, mode, sp, C.c_int(len(sten)))


def lf_flux(x, mode, c, gc, s, dt, dx):
    a, ap = _dp(c)
    b, bp = _dp(gc)
    e, ep = _dp(s)
    return _core_op(lib().o_lf_flux, x, mode, ap, bp, ep, C.c_double(dt), C.c_double(dx))


def lw_rhs(x, mode, c, gc, bw, s, dt, dx):
    a, ap = _dp(c)
    b, bp = _dp(gc)
    w, wp = _dp(bw)
    e, ep = _dp(s)
    return _core_op(lib().o_lw_rhs, x, mode, ap, bp, wp, ep, C.c_double(dt), C.c_double(dx))


def diag(x, mode, v):
    a, ap = _dp(v)
    return _core_op(lib().o_diag, x, mode, ap)


def scale(x, alpha):
    y = x.copy()
    lib().o_scale(C.c_int(y.n[0]), C.c_int(y.r[1]), C.c_double(alpha),
                  y.cores[0].ctypes.data_as(C.POINTER(C.c_double)))
    return y


# ---------------------------------------------------------------- ODE composites
def _out_tt(n, rmax, d):
    cores = [np.zeros(rmax * n[k] * rmax) for k in range(d)]
    return cores


def mode_sum(x, sten3):
    d = x.d
    rmax = 2 * max(x.r)
    cores = _out_tt(x.n, rmax, d)
    rout = np.zeros(d + 1, dtype=np.int32)
    s, sp = _dp(sten3)
    lib().o_mode_sum(C.c_int(d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
                     _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(x.cores), sp,
                     rout.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(cores))
    r = [int(v) for v in rout]
    return HostTT(x.n, r, [cores[k][: r[k] * x.n[k] * r[k + 1]] for k in range(d)])


def rxn_burgers_rhs(x, nu, h, eps, cap):
    d = x.d
    cores = _out_tt(x.n, cap, d)
    rout = np.zeros(d + 1, dtype=np.int32)
    lib().o_rxn_burgers_rhs(C.c_int(d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
                            _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(x.cores),
                            C.c_double(nu), C.c_double(h), C.c_double(eps), C.c_int(cap),
                            rout.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(cores))
    r = [int(v) for v in rout]
    return HostTT(x.n, r, [cores[k][: r[k] * x.n[k] * r[k + 1]] for k in range(d)])


def rxn_burgers_rk4(x, nu, h, dt, eps, cap, steps):
    d = x.d
    cores = _out_tt(x.n, cap, d)
    rout = np.zeros(d + 1, dtype=np.int32)
    lib().o_rxn_burgers_rk4(C.c_int(d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
                            _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(x.cores),
                            C.c_double(nu), C.c_double(h), C.c_double(dt), C.c_double(eps),
                            C.c_int(cap), C.c_int(steps),
                            rout.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(cores))
    r = [int(v) for v in rout]
    return HostTT(x.n, r, [cores[k][: r[k] * x.n[k] * r[k + 1]] for k in range(d)])


def gaussian_ic(d, n, centre=np.pi, width=0.5):
    """Rank-1 TT of exp(-|x - centre|^2 / width) on the periodic grid
    x_j = 2 pi j / n (config 4's initial condition; per-mode factors)."""
    xs = 2 * np.pi * np.arange(n) / n
    cores = []
    cen = np.broadcast_to(np.asarray(centre, dtype=np.float64), (d,))
    for k in range(d):
        cores.append(np.exp(-(xs - cen[k]) ** 2 / width))
    return HostTT([n] * d, [1] * (d + 1), cores)


# ---------------------------------------------------------------- quadrature and entries
def mode_scan(x, mode, kind, w, scale=None):
    """Oracle of pattac_marginal_cusum (kind 0), pattac_marginal_trap (kind 1)
    and pattac_scale_weight_sum (kind 2) on a copy of x."""
    a, ap = _dp(w)
    s, sp = _dp(scale if scale is not None else np.zeros(len(w)))
    y = x.copy()
    c = y.cores[mode]
    lib().o_mode_scan(C.c_int(kind), C.c_int(y.r[mode]), C.c_int(y.n[mode]),
                      C.c_int(y.r[mode + 1]), ap, sp, c.ctypes.data_as(C.POINTER(C.c_double)))
    return y


def apply_marginal(x, mode, w):
    """Oracle of pattac_apply_marginal: a HostTT with d-1 modes."""
    a, ap = _dp(w)
    d = x.d
    n_out = [x.n[k] for k in range(d) if k != mode]
    if mode < d - 1:
        r_out = x.r[: mode + 1] + x.r[mode + 2:]
    else:
        r_out = x.r[: d - 1] + [1]
    outs = [np.zeros(max(r_out[k] * n_out[k] * r_out[k + 1], 1)) for k in range(d - 1)]
    rr = np.zeros(d, dtype=np.int32)
    xs = [np.ascontiguousarray(c) for c in x.cores]
    lib().o_apply_marginal(C.c_int(d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
                           _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(xs),
                           C.c_int(mode), ap, _ptrs(outs), rr.ctypes.data_as(C.POINTER(C.c_int)))
    assert list(rr) == r_out
    return HostTT(n_out, r_out, [outs[k][: r_out[k] * n_out[k] * r_out[k + 1]]
                                 for k in range(d - 1)])


def entry(x, idx):
    """Oracle of paratt_entry (right-to-left chain)."""
    xs = [np.ascontiguousarray(c) for c in x.cores]
    f = lib().o_entry
    f.restype = C.c_double
    return f(C.c_int(x.d), _ip(x.n).ctypes.data_as(C.POINTER(C.c_int)),
             _ip(x.r).ctypes.data_as(C.POINTER(C.c_int)), _ptrs(xs),
             _ip(idx).ctypes.data_as(C.POINTER(C.c_int)))


def truncate_plain(x, tol, maxr=None):
    """attac_tt_truncate on a copy of x: LTR orthogonalisation then RTL
    truncation with absolute tolerance tol, no grid scaling
    (attac.c:909-925 intent; pattac.c:244-522 steps)."""
    y = x.copy()
    d = y.d
    r = _ip(y.r)
    mr = _ip(maxr if maxr is not None else y.r)
    lib().o_lorth(C.c_int(d), _ip(y.n).ctypes.data_as(C.POINTER(C.c_int)),
                  r.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores))
    lib().o_ltruncate(C.c_int(d), _ip(y.n).ctypes.data_as(C.POINTER(C.c_int)),
                      r.ctypes.data_as(C.POINTER(C.c_int)), _ptrs(y.cores),
                      mr.ctypes.data_as(C.POINTER(C.c_int)), C.c_double(tol), None, C.c_int(0),
                      None)
    y.r = [int(v) for v in r]
    return y.trimmed()
