"""Writes tests/golden/cdf_driver_d20_rk5.npz from the compiled reference's
build of the CDF driver (tests/driver/_bin/cdf_driver_ref.x, made by
build_driver.sh): gaussian_dim_20 initial condition, maximum rank 5,
t in [0, 0.001] (100 steps of 1e-5), outputs at t = 0 and t = 0.001.

The reference is run on two MPI ranks: at one rank its TT folder writer
blocks in a send to itself as soon as a core has right rank > 1
(paratt.c:1684-1686), and the second output has rank 5.  Its one-rank and
two-rank marginals agree to 1e-14.  Run in the build container only."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import driver_outputs as D  # noqa: E402

MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")


def main():
    exe = os.path.join(HERE, "_bin", "cdf_driver_ref.x")
    ic = os.path.join(ROOT, "tests", "golden", "gaussian_dim_%d")
    with tempfile.TemporaryDirectory() as wd:
        env = dict(os.environ, OPENBLAS_NUM_THREADS="4")
        p = subprocess.run([MPIEXEC, "-n", "2", exe, ic, "20", "5", "1e-5", "0", "0.001", "0"],
                           cwd=wd, capture_output=True, text=True, env=env, timeout=3000)
        if p.returncode != 0:
            sys.exit("reference driver failed: %s" % p.stderr[-2000:])
        out = D.collect(wd, 20, 5, 2)
        out["ranks"] = np.array(D.printed_ranks(p.stdout), dtype=np.int32)
        for k, lab in ((1, "0.000"), (2, "0.001")):
            n, r, cores = D.read_tt(wd, 20, 5, lab)
            out["tt%d_r" % k] = np.array(r, dtype=np.int32)
            for j, c in enumerate(cores):
                out["tt%d_c%d" % (k, j)] = c
    dst = os.path.join(ROOT, "tests", "golden", "cdf_driver_d20_rk5.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, "ranks", out["ranks"].tolist())


if __name__ == "__main__":
    main()
