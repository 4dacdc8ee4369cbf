"""Helpers for GPU tests: raw device buffers through the library's C-ABI."""
import ctypes as C

import numpy as np

import paratt_amd as P


class DevBuf:
    def __init__(self, arr, dtype=np.float64):
        arr = np.ascontiguousarray(arr, dtype=dtype)
        self.dtype = arr.dtype
        L = P.lib()
        L.ptt_hip_malloc.restype = C.c_void_p
        L.ptt_hip_malloc.argtypes = [C.c_size_t]
        L.ptt_hip_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.ptt_hip_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.ptt_hip_free.argtypes = [C.c_void_p]
        self.n = arr.size
        self.ptr = L.ptt_hip_malloc(max(arr.nbytes, 8))
        if arr.size:
            L.ptt_hip_h2d(self.ptr, arr.ctypes.data, arr.nbytes)

    def get(self):
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            P.lib().ptt_hip_d2h(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr:
            P.lib().ptt_hip_free(self.ptr)
            self.ptr = None


def dgemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, Cb, ldc):
    L = P.lib()
    L.ptt_hip_dgemm.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                C.c_void_p, C.c_long, C.c_void_p, C.c_long, C.c_double,
                                C.c_void_p, C.c_long]
    L.ptt_hip_dgemm(ta.encode(), tb.encode(), M, N, K, alpha, A.ptr, lda, B.ptr, ldb, beta,
                    Cb.ptr, ldc)
    L.ptt_hip_sync()


def to_host(dev):
    """DevTT -> oracle HostTT."""
    import oracle_lib as O
    return O.HostTT(dev.n, dev.r, dev.download_cores())
