"""GPU-side TT folders (paratt.c:1621-1892): device-resident TTs written
and loaded through the .cr text format, checked byte for byte against the
oracle's text, the reference's own data folders (fixtures under
tests/golden/) and a rank-1 folder written by the compiled reference."""
import ctypes as C
import os
import sys
import time

import numpy as np
import pytest

import oracle_lib as O
import paratt_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ttio  # noqa: E402

pytestmark = pytest.mark.gpu
G = O.golden()
FIX = os.path.join(ROOT, "tests", "golden")


def read_bytes(folder, k):
    with open(os.path.join(folder, "core_%d.cr" % (k + 1)), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name,d", [("gaussian_dim_20", 20), ("gaussian_dim_40", 40)])
def test_load_and_rewrite_reference_folders(name, d, tmp_path):
    src = os.path.join(FIX, name)
    t = P.load_folder(src, d)
    n, r, cores = ttio.read_folder(src, d)
    assert t.n == n and t.r == r
    for k, c in enumerate(t.download_cores()):
        assert np.array_equal(c, cores[k]), k
    P.write_folder(t, str(tmp_path / "out"))
    for k in range(d):
        assert read_bytes(tmp_path / "out", k) == read_bytes(src, k), k
    t.free()


def test_write_matches_reference_writer_rank1(tmp_path):
    y = O.HostTT.from_json(G["io_y_rank1"])
    t = P.DevTT.from_host(y)
    P.write_folder(t, str(tmp_path))
    for k in range(y.d):
        assert read_bytes(tmp_path, k).decode() == G["io_files_rank1"][k]
    t.free()


@pytest.mark.parametrize("n,r", [
    ([5, 6, 4, 3], [1, 3, 2, 4, 1]),
    ([50, 17, 33, 9, 2], [1, 17, 40, 9, 2, 1]),
])
def test_write_load_roundtrip(n, r, tmp_path):
    x = O.gauss_tt(n, r, 4242, scale=1e-3)
    x.cores[1][0] = 0.0
    x.cores[1][1] = -0.0
    x.cores[1][2] = 5e-324
    x.cores[1][3] = -1.7976931348623157e308
    t = P.DevTT.from_host(x)
    P.write_folder(t, str(tmp_path))
    for k in range(x.d):
        want = ttio.core_text(x.cores[k][: x.core_size(k)], r[k], n[k], r[k + 1])
        assert read_bytes(tmp_path, k).decode() == want, k
    u = P.load_folder(str(tmp_path), x.d)
    assert u.n == n and u.r == r
    for k, c in enumerate(u.download_cores()):
        assert np.array_equal(c.view(np.uint64), x.cores[k][: x.core_size(k)].view(np.uint64)), k
    t.free()
    u.free()


def test_load_core_slice_into_device_buffer(tmp_path):
    x = O.HostTT.from_json(G["quad_x"])
    t = P.DevTT.from_host(x)
    P.write_folder(t, str(tmp_path))
    L = P.lib()
    buf = C.c_void_p(None)
    cap, rl, nn, rr = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    path = os.fsencode(str(tmp_path / "core_2.cr"))
    L.load_core_slice(path, C.byref(buf), C.byref(cap), C.byref(rl), C.byref(nn), C.byref(rr),
                      1, 4)
    s = G["io_slice_core2_1_4"]
    assert (rl.value, nn.value, rr.value) == (s["r0"], s["n"], s["r1"])
    cnt = rl.value * nn.value * rr.value
    assert cap.value == cnt
    host = np.empty(cnt)
    L.ptt_hip_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.ptt_hip_d2h(host.ctypes.data, buf, cnt * 8)
    assert np.array_equal(host, np.array(s["v"]))
    # a slice beyond the mode is refused and leaves the buffer alone
    L.load_core_slice(path, C.byref(buf), C.byref(cap), C.byref(rl), C.byref(nn), C.byref(rr),
                      2, 99)
    assert cap.value == cnt and nn.value == 3
    L.ptt_hip_free.argtypes = [C.c_void_p]
    L.ptt_hip_free(buf)
    t.free()


def test_rank_mismatch_is_reported(tmp_path, capfd):
    with open(tmp_path / "core_1.cr", "w") as f:
        f.write("1,3,2\n" + "".join("%+1.16e\n" % v for v in range(6)))
    with open(tmp_path / "core_2.cr", "w") as f:
        f.write("3,4,1\n" + "".join("%+1.16e\n" % v for v in range(12)))
    t = P.load_folder(str(tmp_path), 2)
    err = capfd.readouterr().err
    assert "Bad left TT rank in file 2" in err
    # the files' own sizes win (paratt.c:1736-1744 overwrites the ranks)
    assert t.r == [1, 3, 1] and t.n == [3, 4]
    t.free()


def test_io_throughput(tmp_path):
    # d=4, n=256, rank 64: 1M values per interior core, about 26 MB of text
    n, r = [256] * 4, [1, 64, 64, 64, 1]
    t = P.DevTT.gaussian(n, r, 99, 0.01)
    total = sum(r[k] * n[k] * r[k + 1] for k in range(4))
    t0 = time.perf_counter()
    P.write_folder(t, str(tmp_path))
    t1 = time.perf_counter()
    u = P.load_folder(str(tmp_path), 4)
    t2 = time.perf_counter()
    a, b = t.download_cores(), u.download_cores()
    for k in range(4):
        assert np.array_equal(a[k], b[k])
    print("tt folder: %d values, write %.2f s (%.1f M values/s), load %.2f s (%.1f M values/s)"
          % (total, t1 - t0, total / (t1 - t0) / 1e6, t2 - t1, total / (t2 - t1) / 1e6))
    t.free()
    u.free()
