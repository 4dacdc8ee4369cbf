"""The reference's CDF driver compiles against this repository's headers and
links against libparatt_amd.so: every entry point it calls resolves (source
compatibility of the drop-in).  Needs the reference sources and MPICH, so it
only runs where they are."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("PARATT_REFERENCE", "/root/reference")
DRIVER = os.path.join(REF, "mpi_pattac_rxn_burgers_rank_max_read_file.c")


@pytest.mark.skipif(not (os.path.exists(DRIVER) and os.path.exists("/opt/conda/bin/mpicc")),
                    reason="reference driver or MPICH not present")
def test_driver_builds_against_dropin():
    subprocess.run([os.path.join(ROOT, "tests", "driver", "build_driver.sh")], check=True,
                   timeout=600, capture_output=True)
    exe = os.path.join(ROOT, "tests", "driver", "_bin", "cdf_driver_amd.x")
    assert os.path.exists(exe)
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True, check=True).stdout
    line = [l for l in ldd.splitlines() if "libparatt_amd.so" in l][0]
    assert os.path.realpath(os.path.join(ROOT, "paratt_amd", "libparatt_amd.so")) in \
        os.path.realpath(line.split("=>")[1].split("(")[0].strip())
    assert "not found" not in ldd
