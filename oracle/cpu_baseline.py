"""CPU baseline for bench.py (checker infrastructure, never the measured
product): times the compiled REFERENCE (oracle/_ref/libparatt_ref.so, built
from the reference's own sources) on a bounded sample of the TT-rounding
workload, at one MPI rank with threaded OpenBLAS on the host cores.

Run as a child process before the parent touches the GPU:
    python oracle/cpu_baseline.py --d 8 --n 1024 --rank 64 --eps 1e-8 --cap 64
prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--d", type=int, default=8)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--rank", type=int, default=64)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--cap", type=int, default=64)
    ap.add_argument("--reps", type=int, default=1)
    a = ap.parse_args()
    mpi = C.CDLL("/opt/conda/lib/libmpi.so", mode=C.RTLD_GLOBAL)
    mpi.MPI_Init(None, None)
    ref = C.CDLL(os.path.join(HERE, "_ref", "libparatt_ref.so"))
    orc = C.CDLL(os.path.join(HERE, "liboracle.so"))

    class AttacTT(C.Structure):
        _fields_ = [("numDim", C.c_int), ("dim", C.POINTER(C.c_int)),
                    ("ttRank", C.POINTER(C.c_int)), ("coreList", C.POINTER(C.POINTER(C.c_double))),
                    ("coreMem", C.POINTER(C.c_int)), ("wk", C.POINTER(C.c_double)),
                    ("wksz", C.c_int)]

    class Paratt(C.Structure):
        _fields_ = [("x", C.POINTER(AttacTT)), ("mpi_rk", C.c_int), ("mpi_sz", C.c_int),
                    ("distrib_dim_mat", C.POINTER(C.c_int)), ("dim", C.POINTER(C.c_int)),
                    ("numDim", C.c_int), ("iwk", C.POINTER(C.c_int)), ("iwksz", C.c_int),
                    ("orthog_wk", C.c_void_p)]

    P = C.POINTER(Paratt)
    ref.alloc_ptt.restype = P
    ref.paratt_norm.restype = C.c_double
    ref.paratt_norm.argtypes = [P]
    ref.paratt_truncate.argtypes = [P, C.c_double, C.POINTER(C.c_int)]
    ref.attac_tt_axpby.argtypes = [C.c_double, C.POINTER(AttacTT), C.c_double,
                                   C.POINTER(AttacTT), C.POINTER(AttacTT)]
    d, n, r = a.d, a.n, a.rank
    dims = (C.c_int * d)(*([n] * d))
    rr = (C.c_int * (d + 1))(*([1] + [r] * (d - 1) + [1]))
    rz = (C.c_int * (d + 1))(*([1] + [2 * r] * (d - 1) + [1]))
    sc = 1.0 / (n * r) ** 0.5

    def rand(seed):
        p = ref.alloc_ptt(d, dims, rr)
        for k in range(d):
            cnt = rr[k] * n * rr[k + 1]
            orc.o_gauss_fill(C.c_uint64(seed), C.c_int(k), C.c_double(sc), C.c_long(cnt),
                             p.contents.x.contents.coreList[k])
        return p

    x, y = rand(1), rand(2)
    times = []
    ranks = None
    for rep in range(a.reps + 1):
        z = ref.alloc_ptt(d, dims, rz)
        ref.attac_tt_axpby(1.0, x.contents.x, 1.0, y.contents.x, z.contents.x)
        zc = ref.alloc_ptt(d, dims, rz)
        ref.attac_tt_axpby(1.0, x.contents.x, 1.0, y.contents.x, zc.contents.x)
        tol = a.eps * ref.paratt_norm(zc)
        cap = (C.c_int * (d + 1))(*([1] + [a.cap] * (d - 1) + [1]))
        t0 = time.perf_counter()
        ref.paratt_truncate(z, tol, cap)
        t1 = time.perf_counter()
        if rep > 0 or a.reps == 0:
            times.append(t1 - t0)
        ranks = [z.contents.x.contents.ttRank[k] for k in range(d + 1)]
    times.sort()
    med = times[len(times) // 2]
    threads = int(os.environ.get("OPENBLAS_NUM_THREADS", os.cpu_count() or 1))
    print(json.dumps({"seconds": med, "d": d, "n": n, "rank_in": 2 * r, "ranks_out": ranks,
                      "threads": threads}))
    sys.stdout.flush()
    mpi.MPI_Finalize()


if __name__ == "__main__":
    main()
