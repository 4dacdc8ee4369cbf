"""TEST INFRASTRUCTURE ONLY: CPU restatement of the reference's TT folder
format (paratt.c:1621-1892) used as the checker for the library's
pattac_write_tt_folder / pattac_load_tt_folder.  Only tests/, smoke() and
bench.py's CPU baseline may import it.

Folder layout: ``<folder>/core_<k+1>.cr`` holds core k.  Line 1 is
``"r_k,n_k,r_{k+1}"`` (paratt.c:1646-1647); then one value per line printed
with C's ``"%+1.16e"`` (paratt.c:1674-1676) in the loop order right rank,
mode index, left rank (innermost) -- the core's column-major order
(paratt.c:1672-1678).  The loader parses the header with atoi on the comma
fields (paratt.c:1756-1767) and each value with strtod (paratt.c:1875).

Python's ``%`` formatting of floats is correctly rounded, as glibc's printf
is, so ``format_value`` produces the same bytes as the reference; this is
pinned by rewriting the reference's own data folders byte for byte
(tests/test_oracle_ttio.py).
"""
import os

import numpy as np


def format_value(v):
    return "%+1.16e" % float(v)


def core_text(core, r0, n, r1):
    """Text of one core file; `core` is the column-major r0*n*r1 array."""
    flat = np.asarray(core, dtype=np.float64).reshape(-1)
    assert flat.size == r0 * n * r1
    return "%d,%d,%d\n" % (r0, n, r1) + "".join(format_value(v) + "\n" for v in flat)


def write_folder(folder, n, r, cores):
    os.makedirs(folder, exist_ok=True)
    for k in range(len(n)):
        with open(os.path.join(folder, "core_%d.cr" % (k + 1)), "w") as f:
            f.write(core_text(cores[k], r[k], n[k], r[k + 1]))


def _atoi(s):
    """C atoi: optional whitespace and sign, then leading digits."""
    s = s.lstrip()
    i = 0
    if i < len(s) and s[i] in "+-":
        i += 1
    while i < len(s) and s[i].isdigit():
        i += 1
    try:
        return int(s[:i])
    except ValueError:
        return 0


def read_header(path):
    with open(path) as f:
        line = f.readline()
    fields = line.rstrip("\n").split(",")
    fields += [""] * (3 - len(fields))
    return tuple(_atoi(t) for t in fields[:3])


def read_core(path, n_beg=0, n_end=None):
    """(r0, n_slice, r1, values) of the mode slice [n_beg, n_end)."""
    with open(path) as f:
        lines = f.read().split("\n")
    r0, n, r1 = (_atoi(t) for t in (lines[0].split(",") + ["", "", ""])[:3])
    if n_end is None:
        n_end = n
    vals = np.array([float(t) for t in lines[1:1 + r0 * n * r1]], dtype=np.float64)
    full = vals.reshape(r1, n, r0)  # right rank outermost, left rank innermost
    sl = full[:, n_beg:n_end, :]
    return r0, n_end - n_beg, r1, np.ascontiguousarray(sl).reshape(-1)


def read_folder(folder, d):
    n, r, cores = [], [1], []
    for k in range(d):
        r0, nk, r1, v = read_core(os.path.join(folder, "core_%d.cr" % (k + 1)))
        n.append(nk)
        if k == 0:
            r[0] = r0
        r.append(r1)
        cores.append(v)
    return n, r, cores
