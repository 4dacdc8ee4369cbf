"""Learned rank prediction and the sketched right-to-left step.

The sketched step must give exactly the plain method's ranks (it is only
kept when its rank is certified) and a TT within 1e-12 + eps of the plain
result; a prediction that is too small must fall back to the plain step and
reproduce its result bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
import paratt_amd as P
from gpu_util import to_host

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _lib():
    L = P.lib()
    PP = C.POINTER(P.Paratt)
    L.paratt_amd_truncate_rel_sketch.argtypes = [PP, C.c_double, C.POINTER(C.c_int),
                                                 C.POINTER(C.c_int)]
    L.paratt_amd_last_sketch_counts.argtypes = [C.POINTER(C.c_int)]
    return L


def _hadamard_input():
    """u3 (.) v from the reaction-Burgers field (d=6, n=32): formal ranks
    35..48, numerically 8..12, eps-bound bonds."""
    d, n = 6, 32
    h = 2 * np.pi / n
    nu, dt, eps, cap = 0.01, 1e-3, 1e-8, 64
    mr = [1] + [cap] * (d - 1) + [1]

    def tr(z):
        return O.truncate(z, eps * O.norm(z), mr)

    ones = O.HostTT([n] * d, [1] * (d + 1), [np.ones(n) for _ in range(d)])
    d1 = np.array([-0.5 / h, 0.0, 0.5 / h])
    d2 = np.array([1.0, -2.0, 1.0]) / h ** 2

    def rhs(u):
        s1, lap = O.mode_sum(u, d1), O.mode_sum(u, d2)
        v = tr(O.axpby(1.0, O.axpby(1.0, ones, -1.0, u), -1.0, s1))
        return tr(O.axpby(1.0, tr(O.hadamard(u, v)), nu, lap))

    u0 = O.rxn_burgers_rk4(O.gaussian_ic(d, n), nu, h, dt, eps, cap, 1)
    u3 = tr(O.axpby(1.0, u0, 0.5 * dt, rhs(tr(O.axpby(1.0, u0, 0.5 * dt, rhs(u0))))))
    v = tr(O.axpby(1.0, O.axpby(1.0, ones, -1.0, u3), -1.0, O.mode_sum(u3, d1)))
    return O.hadamard(u3, v), mr, eps


def _sketch(z, mr, eps, pred):
    L = _lib()
    dz = P.DevTT.from_host(z)
    mra = (C.c_int * len(mr))(*mr)
    pa = (C.c_int * (z.d + 1))(*pred)
    L.paratt_amd_truncate_rel_sketch(dz.p, eps, mra, pa)
    cnt = (C.c_int * 2)()
    L.paratt_amd_last_sketch_counts(cnt)
    return to_host(dz), (cnt[0], cnt[1])


def test_sketch_with_exact_prediction_matches_plain():
    z, mr, eps = _hadamard_input()
    ref = O.truncate(z, eps * O.norm(z), mr)
    out, (att, acc) = _sketch(z, mr, eps, [0] + ref.r[1:-1] + [0])
    assert att >= 3 and acc == att
    assert out.r == ref.r
    assert O.rel_diff(out, ref) < 1e-12 + eps


def test_sketch_too_small_prediction_falls_back_bit_exact():
    """Ranks 24 at every interior bond, prediction 1 (sketch of 9 columns):
    every sketched step is refused and the plain steps run on untouched
    cores, so the result is the plain result bit for bit."""
    n = [32] * 5
    x = O.gauss_tt(n, [1, 12, 12, 12, 12, 1], 41)
    y = O.gauss_tt(n, [1, 12, 12, 12, 12, 1], 42)
    z = O.axpby(1.0, x, 1e-3, y)
    mr = [1, 64, 64, 64, 64, 1]
    eps = 1e-10
    dz = P.DevTT.from_host(z)
    P.truncate_rel(dz, eps, mr)
    plain = to_host(dz)
    assert min(plain.r[1:-1]) == 24
    out, (att, acc) = _sketch(z, mr, eps, [0] + [1] * (z.d - 1) + [0])
    assert att >= 3 and acc == 0
    assert out.r == plain.r
    for a, b in zip(out.cores, plain.cores):
        np.testing.assert_array_equal(a, b)


def test_sketch_small_prediction_mixed_outcome_matches_plain():
    """Prediction 1 on the Hadamard input: bonds of rank <= 9 may be
    certified from 9 columns, the others fall back; ranks equal plain."""
    z, mr, eps = _hadamard_input()
    ref = O.truncate(z, eps * O.norm(z), mr)
    out, (att, acc) = _sketch(z, mr, eps, [0] + [1] * (z.d - 1) + [0])
    assert att >= 3 and acc < att
    assert out.r == ref.r
    assert O.rel_diff(out, ref) < 1e-12 + eps


def test_committed_predictor_on_heldout_spectra():
    import paratt_amd.learned as Lr
    Pm = Lr.load()
    f = np.load(os.path.join(HERE, "golden", "rank_predictor_heldout.npz"), allow_pickle=False)
    pr = np.rint(Lr.predict(Pm, f["X"]) * 64)
    err = np.abs(pr - f["Y"] * 64)
    assert np.mean(err <= 1) >= 0.98


def test_learned_rk4_steps_equal_plain():
    """Learned mode on a small train: every step's ranks equal the plain run's
    and the states agree to 1e-10 after four steps."""
    import paratt_amd.learned as Lr
    d, n = 8, 64
    L = P.lib()
    Pm = Lr.load()
    states, ranks, cnts = {}, {}, {}
    for mode in (0, 1):
        Lr.setup(n)
        L.paratt_amd_learned_reset_history()
        Lr.enable(Pm, mode)
        u = Lr.initial_condition(d, n, 7)
        c0 = Lr.counts()
        rs = []
        for _ in range(4):
            Lr.rk4(u, 1)
            rs.append(list(u.r))
        c1 = Lr.counts()
        states[mode], ranks[mode] = to_host(u), rs
        cnts[mode] = c1["sketch_accepted"] - c0["sketch_accepted"]
    Lr.disable()
    assert ranks[1] == ranks[0]
    assert O.rel_diff(states[1], states[0]) < 1e-10
