"""GPU TT rounding parity: against the compiled reference's golden vectors
and against the CPU oracle on seeded inputs.  Bar: identical ranks, and
TT-to-TT relative error (through the formal difference, no cancellation)
<= 1e-12 plus the truncation eps where the cut is eps-bound."""
import numpy as np
import pytest

import oracle_lib as O
import paratt_amd as P
from gpu_util import to_host

pytestmark = pytest.mark.gpu

SMALL = ["small_tight", "small_loose", "small_cap2", "small_x_plus_x", "small_x_minus_x",
         "small_wide", "mid_eps_bind", "mid_eps_keep"]


@pytest.mark.parametrize("name", SMALL)
def test_round_matches_reference_golden(name):
    c = O.golden()[name]
    z = O.HostTT.from_json(c["z"])
    dz = P.DevTT.from_host(z)
    P.truncate(dz, c["tol"], c["maxr"])
    out = to_host(dz)
    ref = O.HostTT.from_json(c["out"])
    assert out.r == ref.r, (out.r, ref.r)
    if name == "small_x_minus_x":
        assert O.norm(out) < 1e-12 * O.norm(O.HostTT.from_json(c["x"]))
        return
    assert O.rel_diff(out, ref) < 1e-12


@pytest.mark.parametrize("name", ["cfg1", "cfg1_epsbind"])
def test_round_cfg1_matches_reference(name):
    c = O.golden()[name]
    d, nn, rk = c["d"], c["n"], c["rank"]
    n = [nn] * d
    sc = 1.0 / np.sqrt(nn * rk)
    x = O.gauss_tt(n, O.uniform_ranks(d, rk), c["seed_x"], sc)
    y = O.gauss_tt(n, O.uniform_ranks(d, rk), c["seed_y"], sc)
    z = O.axpby(1.0, x, c["beta"], y)
    dz = P.DevTT.from_host(z)
    nz = P.norm(P.DevTT.from_host(z))
    assert abs(nz - c["norm_in"]) <= 1e-13 * c["norm_in"]
    P.truncate(dz, c["tol"], dz.r)
    out = to_host(dz)
    assert out.r == c["ranks"]
    ref = O.truncate(z, c["tol"], z.r)
    assert O.rel_diff(out, ref) < 1e-12
    vals = np.array([out.entry(ix) for ix in c["entries"]["idx"]])
    refv = np.array(c["entries"]["val"])
    assert np.max(np.abs(vals - refv)) <= 1e-11 * np.max(np.abs(refv))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_round_relative_tolerance_equals_norm_then_truncate(seed):
    n = [12, 10, 14, 9, 11]
    x = O.gauss_tt(n, [1, 6, 8, 7, 5, 1], 100 + seed)
    y = O.gauss_tt(n, [1, 5, 6, 6, 4, 1], 200 + seed)
    z = O.axpby(1.0, x, 1e-6, y)
    tol = 1e-7 * O.norm(z)
    ref = O.truncate(z, tol, z.r)
    dz = P.DevTT.from_host(z)
    P.truncate_rel(dz, 1e-7, dz.r)
    out = to_host(dz)
    assert out.r == ref.r
    assert O.rel_diff(out, ref) < 1e-12 + 1e-7
