"""Readers for the files the CDF driver (mpi_pattac_rxn_burgers_rank_max_read_file.c)
writes in its working directory: the MATLAB-style marginal CDF files
data/bg_rxn_d<d>_rk<r>_mrg_0/marg_0_<k>.m (64 values, mode 0) and
data/bg_rxn_d<d>_rk<r>_mrg_0_mid/mrg_0_mid_<k>.m (64 x 64, modes 0 and d/2),
the TT folders data/bg_rxn_d<d>_rk<r>_tt/cdf_t_<time>/ and the ranks the
driver prints at each output time.  Used by gen_driver_golden.py to store
the compiled reference's outputs and by tests/test_gpu_cdf_driver.py to read
the drop-in build's."""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import ttio  # noqa: E402

_NUM = re.compile(r"([+-][0-9]\.[0-9]+e[+-][0-9]+)")


def read_m(path):
    """Values of a marginal .m file, in file order (column-major)."""
    with open(path) as f:
        return np.array([float(x) for x in _NUM.findall(f.read())])


def printed_ranks(stdout_text):
    """Rank lists printed after each 'tt rank:' header, in order."""
    lines = stdout_text.splitlines()
    out = []
    for i, l in enumerate(lines):
        if l.strip() == "tt rank:" and i + 1 < len(lines):
            out.append([int(v) for v in lines[i + 1].split()])
    return out


def collect(workdir, d, rk, nout):
    """Outputs 1..nout of one run as a dict of arrays."""
    res = {}
    for k in range(1, nout + 1):
        res["marg_%d" % k] = read_m(os.path.join(
            workdir, "data", "bg_rxn_d%d_rk%d_mrg_0" % (d, rk), "marg_0_%d.m" % k))
        res["mid_%d" % k] = read_m(os.path.join(
            workdir, "data", "bg_rxn_d%d_rk%d_mrg_0_mid" % (d, rk), "mrg_0_mid_%d.m" % k))
    return res


def read_tt(workdir, d, rk, time_label):
    return ttio.read_folder(os.path.join(workdir, "data", "bg_rxn_d%d_rk%d_tt" % (d, rk),
                                         "cdf_t_%s" % time_label), d)
