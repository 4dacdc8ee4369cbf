"""End to end: the reference's CDF driver (mpi_pattac_rxn_burgers_rank_max_read_file.c),
compiled against this repository's headers and library by
tests/driver/build_driver.sh with no change beyond initialising its two
uninitialised variables, runs on the GPU and reproduces the compiled
reference's outputs (tests/golden/cdf_driver_d20_rk5.npz, written by
tests/driver/gen_driver_golden.py): gaussian_dim_20 initial condition, rank
cap 5, 100 steps of 1e-5, outputs at t = 0 and t = 0.001.

The driver divides core 0 of several TTs on the host, so it runs with
PARATT_AMD_HOST_VISIBLE=1.  It exercises pattac_load_tt_folder,
pattac_marginal_trap, pattac_apply_cdf_lw_rhs, attac_mult_diag_mat,
attac_tt_axpby, copy_rank_change_tt, paratt_truncate (tolerance 0 and
absolute tolerances down to 1e-19 with and without binding caps),
paratt_entry and pattac_write_tt_folder -- about 6,000 roundings."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "driver"))
import driver_outputs as D  # noqa: E402

EXE = os.path.join(ROOT, "tests", "driver", "_bin", "cdf_driver_amd.x")
GOLD = os.path.join(ROOT, "tests", "golden", "cdf_driver_d20_rk5.npz")


@pytest.mark.gpu
@pytest.mark.timeout(400)
def test_cdf_driver_matches_reference(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("driver not built: tests/driver/build_driver.sh needs the reference sources")
    g = np.load(GOLD)
    ic = os.path.join(ROOT, "tests", "golden", "gaussian_dim_%d")
    env = dict(os.environ, PARATT_AMD_HOST_VISIBLE="1")
    print("running the CDF driver, 100 steps", flush=True)
    p = subprocess.run([EXE, ic, "20", "5", "1e-5", "0", "0.001", "0"], cwd=str(tmp_path),
                       capture_output=True, text=True, env=env, timeout=360)
    assert p.returncode == 0, p.stderr[-3000:]
    assert "success proc 0" in p.stdout
    assert D.printed_ranks(p.stdout) == g["ranks"].tolist()
    out = D.collect(str(tmp_path), 20, 5, 2)
    for key in ("marg_1", "mid_1", "marg_2", "mid_2"):
        diff = np.max(np.abs(out[key] - g[key]))
        print("%s: max |gpu - reference| = %.3e" % (key, diff))
        assert diff <= 1e-12, key
    for k, lab in ((1, "0.000"), (2, "0.001")):
        n, r, cores = D.read_tt(str(tmp_path), 20, 5, lab)
        ref = O.HostTT([64] * 20, g["tt%d_r" % k], [g["tt%d_c%d" % (k, j)] for j in range(20)])
        assert list(r) == ref.r
        rel = O.rel_diff(O.HostTT(n, r, cores), ref)
        print("TT at t=%s: relative difference %.3e" % (lab, rel))
        assert rel <= 1e-11
