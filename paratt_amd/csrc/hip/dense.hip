// dense.hip -- small dense FP64 factorizations on one workgroup (gfx950).
//
// ptt_hip_chol_qr_factor: the r x r step of Cholesky-QR.  The Gram matrix is
// Jacobi-scaled (unit diagonal) so column scaling never costs accuracy, the
// factorization guards tiny pivots (the index is marked deficient: its row
// of R and column of R^-1 become zero, so A*T has a zero column there and
// A = (A*T)*R still holds to rounding), and the inverse is formed in place.
//
// ptt_hip_svd_jacobi: one-sided (Hestenes) Jacobi SVD with round-robin
// parallel pairing; each pair is owned by a lane group that reduces its
// three inner products with cross-lane shuffles.  Columns live in LDS when
// they fit (160 KB per CU on gfx950), else in the caller's global buffers.
// The reference's rank rule (pattac.c:468-481) is applied on the device.
#include "common.hpp"

#include <cmath>

using namespace ptt;

namespace {

constexpr int NT = 1024;
constexpr int LDS_DOUBLES = 17408;  // 136 KB dynamic + static arrays <= 160 KB LDS

__global__ __launch_bounds__(NT) void k_chol_qr_factor(int c, const double *__restrict__ G,
                                                       double shift, double tiny,
                                                       double *__restrict__ R,
                                                       double *__restrict__ T, int *info,
                                                       double *dinfo, double *gwork) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    // Work matrix S (c x c): LDS if it fits, else global scratch.
    double *S = (c * c + 3 * c <= LDS_DOUBLES) ? lds : gwork;
    double *D = (c * c + 3 * c <= LDS_DOUBLES) ? lds + c * c : gwork + (long)c * c;
    double *col = D + c;   // temp column
    double *dd = col + c;  // pivots (as double)
    __shared__ int s_def[1];
    __shared__ double s_red[NT / 64];
    __shared__ double s_rjj;
    __shared__ int s_isdef;

    for (int j = tid; j < c; j += NT) {
        double g = G[j + (long)j * c];
        D[j] = g > 0.0 ? sqrt(g) : 0.0;
    }
    if (tid == 0) s_def[0] = 0;
    __syncthreads();
    for (long e = tid; e < (long)c * c; e += NT) {
        int i = (int)(e % c), k = (int)(e / c);
        double di = D[i], dk = D[k];
        S[e] = (di > 0.0 && dk > 0.0) ? G[e] / (di * dk) : 0.0;
    }
    __syncthreads();
    // Right-looking Cholesky on the upper triangle: S[i + c*k], i <= k.
    for (int j = 0; j < c; j++) {
        if (tid == 0) {
            double p = S[j + (long)j * c] + shift;
            int dead = !(D[j] > 0.0) || !(p > tiny);
            s_isdef = dead;
            s_rjj = dead ? 0.0 : sqrt(p);
            S[j + (long)j * c] = s_rjj;
            dd[j] = dead ? 1.0 : 0.0;
            if (dead) s_def[0] += 1;
        }
        __syncthreads();
        const double rjj = s_rjj;
        const int dead = s_isdef;
        for (int k = j + 1 + tid; k < c; k += NT)
            S[j + (long)k * c] = dead ? 0.0 : S[j + (long)k * c] / rjj;
        __syncthreads();
        if (!dead) {
            const int w = c - j - 1;
            for (long t = tid; t < (long)w * w; t += NT) {
                int i = j + 1 + (int)(t % w), k = j + 1 + (int)(t / w);
                if (i <= k) S[i + (long)k * c] -= S[j + (long)i * c] * S[j + (long)k * c];
            }
        }
        __syncthreads();
    }
    // ||Rs||_F and R = Rs * D (upper, zeros below).
    double fr = 0.0;
    for (long e = tid; e < (long)c * c; e += NT) {
        int i = (int)(e % c), k = (int)(e / c);
        double v = (i <= k) ? S[e] : 0.0;
        fr += v * v;
        R[e] = v * D[k];
    }
    for (int o = 32; o > 0; o >>= 1) fr += __shfl_xor(fr, o);
    if ((tid & 63) == 0) s_red[tid >> 6] = fr;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int i = 0; i < NT / 64; i++) t += s_red[i];
        dinfo[0] = sqrt(t);
    }
    __syncthreads();
    // In-place inverse of the upper triangle, column by column; 8 lanes per
    // row share each dot product.
    for (int j = 0; j < c; j++) {
        const bool dead = dd[j] != 0.0;
        for (int i = tid; i < j; i += NT) col[i] = S[i + (long)j * c];
        __syncthreads();
        const double tjj = dead ? 0.0 : 1.0 / S[j + (long)j * c];
        const int row = tid >> 3, sub = tid & 7;
        for (int i0 = 0; i0 < j; i0 += NT / 8) {
            int i = i0 + row;
            double s = 0.0;
            if (i < j && !dead)
                for (int l = i + sub; l < j; l += 8) s += S[i + (long)l * c] * col[l];
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            s += __shfl_xor(s, 4);
            if (i < j && sub == 0) S[i + (long)j * c] = dead ? 0.0 : -s * tjj;
        }
        __syncthreads();
        if (tid == 0) S[j + (long)j * c] = tjj;
        __syncthreads();
    }
    double ft = 0.0;
    for (long e = tid; e < (long)c * c; e += NT) {
        int i = (int)(e % c), k = (int)(e / c);
        double v = (i <= k) ? S[e] : 0.0;
        ft += v * v;
        T[e] = (D[i] > 0.0) ? v / D[i] : 0.0;
    }
    for (int o = 32; o > 0; o >>= 1) ft += __shfl_xor(ft, o);
    __syncthreads();
    if ((tid & 63) == 0) s_red[tid >> 6] = ft;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int i = 0; i < NT / 64; i++) t += s_red[i];
        dinfo[0] *= sqrt(t);
        info[0] = s_def[0];
    }
}

__device__ __forceinline__ void rr_pair(int kp, int t, int p, int &a, int &b) {
    // Round-robin (circle) schedule over kp players, round t in [0, kp-1).
    const int m = kp - 1;
    if (p == 0) {
        a = m;
        b = t;
    } else {
        a = (t + p) % m;
        b = (t - p + m) % m;
    }
}

__global__ __launch_bounds__(NT) void k_svd_jacobi(int m, int k, const double *__restrict__ A,
                                                   long lda, double *__restrict__ sig,
                                                   double *__restrict__ U, long ldu,
                                                   double *__restrict__ V, long ldv,
                                                   double log_scale, double delta, int cap,
                                                   int *iout, double *gwork, int use_lds) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    double *W = use_lds ? lds : gwork;            // m x k, ld m
    double *Vw = V ? (use_lds ? lds + (long)m * k : gwork + (long)m * k) : nullptr;  // k x k
    __shared__ int s_rot;
    __shared__ double s_sig[1024];
    __shared__ double s_sorted[1024];
    __shared__ int s_pos[1024];

    for (long e = tid; e < (long)m * k; e += NT) {
        long j = e / m, i = e - j * m;
        W[e] = A[i + j * lda];
    }
    if (Vw)
        for (long e = tid; e < (long)k * k; e += NT) {
            long j = e / k, i = e - j * k;
            Vw[e] = (i == j) ? 1.0 : 0.0;
        }
    __syncthreads();

    const int kp = k + (k & 1);
    const int P = kp / 2;
    int G = 64;
    while (G > 1 && G * P > NT) G >>= 1;
    const int ngroups = NT / G;
    const int g = tid / G, gl = tid - g * G;
    const double tol = 0x1.0p-52 * sqrt((double)m);
    int sweep = 0;
    const int max_sweeps = 40;
    for (; sweep < max_sweeps && kp >= 2; sweep++) {
        if (tid == 0) s_rot = 0;
        __syncthreads();
        int rot_local = 0;
        for (int t = 0; t < kp - 1; t++) {
            for (int p0 = 0; p0 < P; p0 += ngroups) {
                int p = p0 + g;
                int a = -1, b = -1;
                if (p < P) rr_pair(kp, t, p, a, b);
                bool valid = (p < P) && a < k && b < k;
                double al = 0.0, be = 0.0, ga = 0.0;
                if (valid) {
                    const double *ca = W + (long)a * m, *cb = W + (long)b * m;
                    for (int i = gl; i < m; i += G) {
                        double x = ca[i], y = cb[i];
                        al += x * x;
                        be += y * y;
                        ga += x * y;
                    }
                }
                for (int o = 1; o < G; o <<= 1) {
                    al += __shfl_xor(al, o);
                    be += __shfl_xor(be, o);
                    ga += __shfl_xor(ga, o);
                }
                if (valid && fabs(ga) > tol * sqrt(al * be) && al > 0.0 && be > 0.0) {
                    double zeta = (be - al) / (2.0 * ga);
                    double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
                    double *ca = W + (long)a * m, *cb = W + (long)b * m;
                    for (int i = gl; i < m; i += G) {
                        double x = ca[i], y = cb[i];
                        ca[i] = cs * x - sn * y;
                        cb[i] = sn * x + cs * y;
                    }
                    if (Vw) {
                        double *va = Vw + (long)a * k, *vb = Vw + (long)b * k;
                        for (int i = gl; i < k; i += G) {
                            double x = va[i], y = vb[i];
                            va[i] = cs * x - sn * y;
                            vb[i] = sn * x + cs * y;
                        }
                    }
                    rot_local = 1;
                }
            }
            __syncthreads();
        }
        if (rot_local) s_rot = 1;
        __syncthreads();
        if (!s_rot) break;
        __syncthreads();
    }
    // column norms
    for (int j = g; j < k; j += ngroups) {
        const double *cj = W + (long)j * m;
        double s = 0.0;
        for (int i = gl; i < m; i += G) s += cj[i] * cj[i];
        for (int o = 1; o < G; o <<= 1) s += __shfl_xor(s, o);
        if (gl == 0) s_sig[j] = sqrt(s);
    }
    __syncthreads();
    // descending order by rank counting (ties by index)
    for (int j = tid; j < k; j += NT) {
        double sj = s_sig[j];
        int pos = 0;
        for (int i = 0; i < k; i++) {
            double si = s_sig[i];
            pos += (si > sj) || (si == sj && i < j);
        }
        s_pos[j] = pos;
    }
    __syncthreads();
    for (int j = tid; j < k; j += NT) {
        sig[s_pos[j]] = s_sig[j];
        s_sorted[s_pos[j]] = s_sig[j];
    }
    for (long e = tid; e < (long)m * k; e += NT) {
        long j = e / m, i = e - j * m;
        double sj = s_sig[j];
        U[i + (long)s_pos[j] * ldu] = sj > 0.0 ? W[e] / sj : 0.0;
    }
    if (Vw)
        for (long e = tid; e < (long)k * k; e += NT) {
            long j = e / k, i = e - j * k;
            V[i + (long)s_pos[j] * ldv] = Vw[e];
        }
    __syncthreads();
    if (tid == 0) {
        const double s0 = s_sorted[0];
        const double thr = sqrt((double)k) * 0x1.0p-53 * s0;
        int numrank = 0;
        for (int j = 0; j < k; j++) numrank += (s_sorted[j] > thr) && (s0 > 0.0);
        int rho = numrank;
        double tail = 0.0;
        const double ld = delta > 0.0 ? log(delta) : -INFINITY;
        for (int q = numrank - 1; q >= 0; q--) {
            double v = s_sorted[q];
            tail += v * v;
            double lhs = (tail > 0.0) ? 0.5 * log(tail) + log_scale : -INFINITY;
            if (delta > 0.0 && lhs < ld) rho = q;
            else break;
        }
        if (rho > k) rho = k;
        if (rho > cap) rho = cap;
        if (rho < 1) rho = 1;
        iout[0] = numrank;
        iout[1] = rho;
        iout[2] = sweep + 1;
    }
}

__global__ void k_svd_factors(int m, int rhomax, const int *__restrict__ drho,
                              const double *__restrict__ U, long ldu,
                              const double *__restrict__ sig, double wscale, double *W,
                              double *US) {
    const int rho = *drho;
    long total = (long)m * rhomax;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        int j = (int)(e / m), i = (int)(e - (long)j * m);
        if (j >= rho) continue;
        double u = U[i + (long)j * ldu], s = sig[j];
        if (US) US[i + (long)j * m] = u * s;
        if (W) W[j + (long)i * rho] = (s > 0.0) ? wscale * u / s : 0.0;
    }
}

}  // namespace

void allow_big_lds() {
    static bool done = false;
    if (done) return;
    const int bytes = LDS_DOUBLES * (int)sizeof(double);
    HIPCHECK(hipFuncSetAttribute((const void *)k_chol_qr_factor,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    HIPCHECK(hipFuncSetAttribute((const void *)k_svd_jacobi,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done = true;
}

extern "C" {

void ptt_hip_chol_qr_factor(int c, const double *G, double shift, double tiny, double *R,
                            double *T, int *info, double *dinfo) {
    ptt_hip_init();
    allow_big_lds();
    size_t need = (size_t)c * c + 3 * (size_t)c;
    double *gwork = nullptr;
    size_t shmem = 0;
    if (need <= (size_t)LDS_DOUBLES) {
        shmem = need * sizeof(double);
    } else {
        gwork = (double *)ptt_hip_scratch(PTT_WS_SKETCH, need * sizeof(double));
    }
    PTT_LAUNCH("chol_qr_factor", k_chol_qr_factor, dim3(1), dim3(NT), shmem, c, G, shift, tiny, R,
               T, info, dinfo, gwork);
}

void ptt_hip_svd_jacobi(int m, int k, double *A, long lda, double *sig, double *U, long ldu,
                        double *V, long ldv, double log_scale, double delta, int cap,
                        int *iout) {
    ptt_hip_init();
    allow_big_lds();
    if (k > 1024) {
        fprintf(stderr, "paratt_amd: Jacobi SVD limited to k <= 1024 (k = %d)\n", k);
        exit(1);
    }
    size_t need = (size_t)m * k + (V ? (size_t)k * k : 0);
    int use_lds = need <= (size_t)LDS_DOUBLES ? 1 : 0;
    double *gwork = nullptr;
    size_t shmem = 0;
    if (use_lds)
        shmem = need * sizeof(double);
    else
        gwork = (double *)ptt_hip_scratch(PTT_WS_SKETCH, need * sizeof(double));
    PTT_LAUNCH("svd_jacobi", k_svd_jacobi, dim3(1), dim3(NT), shmem, m, k, A, lda, sig, U, ldu, V,
               ldv, log_scale, delta, cap, iout, gwork, use_lds);
}

void ptt_hip_svd_factors(int m, int rhomax, const int *drho, const double *U, long ldu,
                         const double *sig, double wscale, double *W, double *US) {
    ptt_hip_init();
    long total = (long)m * rhomax;
    if (total <= 0) return;
    PTT_LAUNCH("svd_factors", k_svd_factors, dim3(stream_blocks(total, 256)), dim3(256), 0, m,
               rhomax, drho, U, ldu, sig, wscale, W, US);
}

void ptt_hip_trmm_small(int c, const double *A, const double *B, double *C) {
    ptt_hip_dgemm('N', 'N', c, c, c, 1.0, A, c, B, c, 0.0, C, c);
}

}  // extern "C"
