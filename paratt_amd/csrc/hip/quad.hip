// quad.hip -- quadrature along one mode of a TT and entry evaluation, gfx950.
//
// Core k is column-major r0 x n x r1; element (a, j, b) at a + r0*(j + n*b).
// A "fiber" is the n values of one (a, b) pair.  The reference's cumulative
// sums (paratt.c:1386-1618) run sequentially along each fiber, so here one
// lane owns one fiber and walks it in the same order: lanes of a wave take
// consecutive a, so every step of the walk is one coalesced load per wave.
// All kernels compile with FP contraction off and add in the reference's
// order, which makes their results bit-identical to its scalar loops.
#include "common.hpp"

using namespace ptt;

namespace {

#pragma clang fp contract(off)

// kind 0: weighted cumulative sum (paratt.c:1386-1440):
//   g[j] = sum_{j' <= j} w[j'] f[j'].
// kind 1: trapezoid rule (paratt.c:1444-1516):
//   g[0] = 0, g[j] = g[j-1] + (0.5 w[j]) (f[j-1] + f[j]).
// kind 2: trapezoid rule on sum_w plus the scaled input
//   (paratt.c:1518-1601): g[j] = trap[j] + f[j] scale[j].
__global__ void k_mode_scan(int kind, int r0, int n, int r1, const double *__restrict__ w,
                            const double *__restrict__ scale, double *__restrict__ f) {
    const long nfib = (long)r0 * r1;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nfib;
         t += (long)gridDim.x * blockDim.x) {
        const long b = t / r0, a = t - b * r0;
        double *fb = f + a + (long)r0 * n * b;  // fiber base, stride r0
        if (kind == 0) {
            double acc = 0.0;
            for (int j = 0; j < n; j++) {
                double v = fb[(long)r0 * j] * w[j];
                acc = (j == 0) ? v : acc + v;
                fb[(long)r0 * j] = acc;
            }
        } else {
            double cum = 0.0, prev = fb[0];
            for (int j = 1; j < n; j++) {
                double cur = fb[(long)r0 * j];
                double delta = (0.5 * w[j]) * (prev + cur);
                fb[(long)r0 * (j - 1)] = (kind == 2) ? cum + prev * scale[j - 1] : cum;
                cum = cum + delta;
                prev = cur;
            }
            fb[(long)r0 * (n - 1)] = (kind == 2) ? cum + prev * scale[n - 1] : cum;
        }
    }
}

// Batched product with a sequential inner sum started from 0.0:
//   C[i + ldc*j + sC*q] = sum_{k=0}^{K-1} A[i + lda*k + sA*q] * B[k + ldb*j + sB*q]
// for i < M, j < N, q < batch.  One lane per output, i fastest.  These are
// the small contractions of the marginal (paratt.c:1311-1377), summed in the
// reference's loop order.
__global__ void k_seq_gemm(int M, int N, int K, int batch, const double *__restrict__ A, long lda,
                           long sA, const double *__restrict__ B, long ldb, long sB,
                           double *__restrict__ C, long ldc, long sC) {
    const long mn = (long)M * N, total = mn * batch;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const long q = e / mn, rem = e - q * mn;
        const long j = rem / M, i = rem - j * M;
        const double *a = A + i + sA * q;
        const double *bcol = B + ldb * j + sB * q;
        double acc = 0.0;
        for (int k = 0; k < K; k++) acc = acc + a[lda * k] * bcol[k];
        C[i + ldc * j + sC * q] = acc;
    }
}

// Entry of a TT by the right-to-left chain of matrix-vector products of
// pattacValue (pattac.c:82-121): v <- X_i[:, idx_i, :] v for i = d-2 .. 0,
// starting from v = X_{d-1}[:, idx_{d-1}, 0].  One workgroup; up to
// ENTRY_CHUNK cores per launch, v carried through global memory between
// launches.  Row p sums over q in order, as the reference's dgemv does.
constexpr int ENTRY_CHUNK = 48;
constexpr int ENTRY_NT = 256;
struct EntryChunk {
    const double *core[ENTRY_CHUNK];
    int r0[ENTRY_CHUNK], r1[ENTRY_CHUNK], n[ENTRY_CHUNK], idx[ENTRY_CHUNK];
    int count;  // cores in this chunk, processed from count-1 down to 0
    int first;  // 1 if the chunk holds the last core of the TT
};

__global__ __launch_bounds__(ENTRY_NT) void k_tt_entry(EntryChunk ch, double *__restrict__ vec,
                                                       int rmax) {
    extern __shared__ double sh[];
    double *va = sh, *vb = sh + rmax;
    int top = ch.count - 1;
    if (ch.first) {
        // v = X_{d-1}[:, idx, 0]
        const double *x = ch.core[top] + (long)ch.idx[top] * ch.r0[top];
        for (int p = threadIdx.x; p < ch.r0[top]; p += blockDim.x) va[p] = x[p];
        top--;
    } else {
        for (int p = threadIdx.x; p < ch.r1[top]; p += blockDim.x) va[p] = vec[p];
    }
    __syncthreads();
    for (int i = top; i >= 0; i--) {
        const int r0 = ch.r0[i], r1 = ch.r1[i];
        const long ld = (long)r0 * ch.n[i];
        const double *x = ch.core[i] + (long)ch.idx[i] * r0;
        for (int p = threadIdx.x; p < r0; p += blockDim.x) {
            double acc = 0.0;
            for (int q = 0; q < r1; q++) acc = acc + va[q] * x[p + ld * q];
            vb[p] = acc;
        }
        __syncthreads();
        double *t = va;
        va = vb;
        vb = t;
    }
    const int rout = ch.r0[0];
    for (int p = threadIdx.x; p < rout; p += blockDim.x) vec[p] = va[p];
}

}  // namespace

extern "C" {

void ptt_hip_mode_scan(int kind, int r0, int n, int r1, const double *w, const double *scale,
                       double *f) {
    ptt_hip_init();
    if (kind < 0 || kind > 2) {
        fprintf(stderr, "ptt_hip_mode_scan: unknown kind %d\n", kind);
        exit(1);
    }
    long nfib = (long)r0 * r1;
    if (nfib <= 0 || n <= 0) return;
    // 64-lane workgroups: small fiber counts still spread over many CUs
    PTT_LAUNCH_FB("mode_scan", 3.0 * nfib * n, 16.0 * nfib * n, k_mode_scan,
                  dim3(cdiv(nfib, 64) < 65535u ? cdiv(nfib, 64) : 65535u), dim3(64), 0, kind, r0,
                  n, r1, w, scale, f);
}

void ptt_hip_seq_gemm(int M, int N, int K, int batch, const double *A, long lda, long sA,
                      const double *B, long ldb, long sB, double *C, long ldc, long sC) {
    ptt_hip_init();
    long total = (long)M * N * batch;
    if (total <= 0) return;
    PTT_LAUNCH_FB("seq_gemm", 2.0 * total * K, 8.0 * total, k_seq_gemm,
                  dim3(stream_blocks(total, 256)), dim3(256), 0, M, N, K, batch, A, lda, sA, B,
                  ldb, sB, C, ldc, sC);
}

void ptt_hip_tt_entry(int d, const int *n, const int *r, double *const *cores, const int *idx,
                      double *vec, double *out) {
    ptt_hip_init();
    int rmax = 1;
    for (int k = 0; k <= d; k++) rmax = r[k] > rmax ? r[k] : rmax;
    if ((size_t)rmax * 2 * sizeof(double) > 64 * 1024) {
        fprintf(stderr, "ptt_hip_tt_entry: rank %d too large\n", rmax);
        exit(1);
    }
    // chunks of cores from the right end
    int hi = d;
    while (hi > 0) {
        int lo = hi - ENTRY_CHUNK > 0 ? hi - ENTRY_CHUNK : 0;
        EntryChunk ch;
        ch.count = hi - lo;
        ch.first = (hi == d);
        for (int i = 0; i < ch.count; i++) {
            int k = lo + i;
            ch.core[i] = cores[k];
            ch.r0[i] = r[k];
            ch.r1[i] = r[k + 1];
            ch.n[i] = n[k];
            ch.idx[i] = idx[k];
        }
        PTT_LAUNCH("tt_entry", k_tt_entry, dim3(1), dim3(ENTRY_NT),
                   sizeof(double) * 2 * (size_t)rmax, ch, vec, rmax);
        hi = lo;
    }
    ptt_hip_d2h(out, vec, sizeof(double));
}

}  // extern "C"
