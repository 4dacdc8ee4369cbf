"""Pins the oracle of the TT folder format (oracle/ttio.py, paratt.c:1621-1892)
and checks the host-only parts of the library's loader (no device calls).

Pins: the reference's own data folders (gaussian_dim_20 / gaussian_dim_40,
copied as fixtures under tests/golden/) are rewritten byte for byte; a
rank-1 TT written by the compiled reference matches the oracle's text; a
rank > 1 folder in the oracle's format is read back by the compiled
reference's loader into the identical TT (golden flag), and its mode slice
[1, 4) matches the reference's load_core_slice."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ttio  # noqa: E402

G = O.golden()
FIX = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("name,d", [("gaussian_dim_20", 20), ("gaussian_dim_40", 40)])
def test_reference_folders_rewrite_byte_identical(name, d, tmp_path):
    src = os.path.join(FIX, name)
    n, r, cores = ttio.read_folder(src, d)
    assert n == [64] * d and r == [1] * (d + 1)
    ttio.write_folder(str(tmp_path), n, r, cores)
    for k in range(d):
        a = open(os.path.join(src, "core_%d.cr" % (k + 1)), "rb").read()
        b = open(os.path.join(tmp_path, "core_%d.cr" % (k + 1)), "rb").read()
        assert a == b, k


def test_rank1_text_matches_reference_writer():
    y = O.HostTT.from_json(G["io_y_rank1"])
    for k in range(y.d):
        assert ttio.core_text(y.cores[k], y.r[k], y.n[k], y.r[k + 1]) == G["io_files_rank1"][k]


def test_rank_gt1_roundtrip_and_slice(tmp_path):
    assert G["io_reload_identical"] == 1
    x = O.HostTT.from_json(G["quad_x"])
    ttio.write_folder(str(tmp_path), x.n, x.r, x.cores)
    n, r, cores = ttio.read_folder(str(tmp_path), x.d)
    assert n == x.n and r == x.r
    for k in range(x.d):
        assert np.array_equal(cores[k], x.cores[k][: x.core_size(k)])
    s = G["io_slice_core2_1_4"]
    r0, nn, r1, v = ttio.read_core(os.path.join(tmp_path, "core_2.cr"), 1, 4)
    assert (r0, nn, r1) == (s["r0"], s["n"], s["r1"])
    assert np.array_equal(v, np.array(s["v"]))


def test_format_special_values():
    # 17 significant digits round-trip every double, including subnormals
    vals = [0.0, -0.0, 5e-324, 1.7976931348623157e308, -2.2250738585072014e-308,
            1.0 / 3.0, -123456789.125]
    for v in vals:
        t = ttio.format_value(v)
        assert float(t) == v and (np.signbit(float(t)) == np.signbit(v))
    assert ttio.format_value(1.0) == "+1.0000000000000000e+00"
    assert ttio.format_value(-0.0) == "-0.0000000000000000e+00"


# ---- host-only paths of the library (no GPU is touched) ----
def test_library_read_core_size():
    import paratt_amd as P
    assert P.read_core_size(os.path.join(FIX, "gaussian_dim_20", "core_7.cr")) == (1, 64, 1)


def test_library_read_core_size_missing_file_leaves_outputs(tmp_path):
    code = ("import paratt_amd as P; print(P.read_core_size(%r))"
            % str(tmp_path / "nope.cr"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         cwd=ROOT, timeout=120)
    assert out.returncode == 0
    assert "(-1, -1, -1)" in out.stdout
    assert "read_core_size error: can't open" in out.stderr


def test_library_load_missing_folder_returns_null(tmp_path):
    code = ("import paratt_amd as P; print(P.load_folder(%r, 3))" % str(tmp_path / "none"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         cwd=ROOT, timeout=120)
    assert out.returncode == 0 and "None" in out.stdout
    assert "cannot read the size" in out.stderr


def _fake_paratt_code(body):
    return """
import ctypes as C
import paratt_amd as P
d = 3
dim = (C.c_int * 3)(4, 5, 6)
rk = (C.c_int * 4)(1, 2, 2, 1)
cores = (C.c_void_p * 3)()
mem = (C.c_int * 3)(8, 20, 12)
tt = P.AttacTT(numDim=d, dim=dim, ttRank=rk, coreList=cores, coreMem=mem)
pt = P.Paratt(x=C.pointer(tt), mpi_rk=0, mpi_sz=1, dim=dim, numDim=d)
L = P.lib()
""" + body


@pytest.mark.parametrize("body,msg", [
    ("print(L.paratt_entry(C.byref(pt), (C.c_int * 3)(1, 5, 0)))", "out of range"),
    ("L.pattac_marginal_trap(C.byref(pt), 3, (C.c_double * 8)())", "out of range"),
    ("L.pattac_marginal_cusum(C.byref(pt), -1, (C.c_double * 8)())", "out of range"),
    ("m = C.pointer(pt); L.pattac_apply_marginal(C.byref(pt), C.byref(m), 0, "
     "(C.c_double * 8)())", "was not NULL"),
])
def test_library_fatal_argument_checks(body, msg):
    # each check fires on the host before any device call, then exit(1)
    out = subprocess.run([sys.executable, "-c", _fake_paratt_code(body)], capture_output=True,
                         text=True, cwd=ROOT, timeout=120)
    assert out.returncode == 1, (out.stdout, out.stderr)
    assert msg in out.stderr
