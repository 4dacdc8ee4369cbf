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
// The Jacobi kernel holds 24 KB of static LDS (sorting and packing tables),
// so its dynamic budget is smaller: 17280 doubles + 24.1 KB < 160 KB.
constexpr int SVD_LDS_DOUBLES = 17280;

// Work matrix S is stored ROW-major in LDS (S[i*c + k] = entry (i,k)) so the
// pivot row read by every thread in the trailing update is contiguous and
// each thread's own entries are consecutive across lanes: no bank conflicts.
template <bool INLDS>
__global__ __launch_bounds__(NT) void k_chol_qr_factor(int c, const double *__restrict__ G,
                                                       double shift, double tiny,
                                                       double *__restrict__ R,
                                                       double *__restrict__ T, int *info,
                                                       double *dinfo, double *gwork,
                                                       int want_kappa) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    // INLDS: the work matrix is LDS-resident (the compiler sees the shared
    // address space and issues ds_* instructions); else global scratch.
    double *S = INLDS ? lds : gwork;
    double *D = S + (long)c * c;
    double *col = D + c;   // temp vector
    double *dd = col + c;  // deficient flags
    double *pv = dd + c;   // power-iteration vector (2c)
    __shared__ int s_def;
    __shared__ double s_red[NT / 64];
    __shared__ double s_rjj, s_scal;
    __shared__ int s_isdef;

    for (int j = tid; j < c; j += NT) {
        double g = G[j + (long)j * c];
        D[j] = g > 0.0 ? sqrt(g) : 0.0;
    }
    if (tid == 0) s_def = 0;
    __syncthreads();
    // G is symmetric: S(i,k) = G[i + k*c] / (D_i D_k), stored row-major.
    for (long e = tid; e < (long)c * c; e += NT) {
        int k = (int)(e % c), i = (int)(e / c);
        double di = D[i], dk = D[k];
        S[e] = (di > 0.0 && dk > 0.0) ? G[k + (long)i * c] / (di * dk) : 0.0;
    }
    __syncthreads();
    // lambda_max of the scaled Gram by power iteration (before it is overwritten)
    auto block_sum = [&](double v) -> double {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        __syncthreads();
        if ((tid & 63) == 0) s_red[tid >> 6] = v;
        __syncthreads();
        double t = 0.0;
        for (int i = 0; i < NT / 64; i++) t += s_red[i];
        return t;
    };
    const int PIT = want_kappa ? 10 : 0;
    for (int i = tid; i < c; i += NT) pv[i] = 1.0 + 0.01 * (i % 7);
    __syncthreads();
    double lmax = 0.0;
    for (int it = 0; it < PIT; it++) {
        // y = S x (rows in parallel, 8 lanes per row)
        const int row = tid >> 3, sub = tid & 7;
        for (int i0 = 0; i0 < c; i0 += NT / 8) {
            int i = i0 + row;
            double acc = 0.0;
            if (i < c)
                for (int l = sub; l < c; l += 8) acc += S[(long)i * c + l] * pv[l];
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            acc += __shfl_xor(acc, 4);
            if (i < c && sub == 0) pv[c + i] = acc;
        }
        __syncthreads();
        double nn = 0.0;
        for (int i = tid; i < c; i += NT) nn += pv[c + i] * pv[c + i];
        nn = block_sum(nn);
        double xn = 0.0;
        for (int i = tid; i < c; i += NT) xn += pv[i] * pv[i];
        xn = block_sum(xn);
        lmax = (xn > 0.0) ? sqrt(nn / xn) : 0.0;
        double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
        for (int i = tid; i < c; i += NT) pv[i] = pv[c + i] * inv;
        __syncthreads();
    }
    // Right-looking Cholesky, upper triangle, row-major.
    for (int j = 0; j < c; j++) {
        if (tid == 0) {
            double p = S[(long)j * c + j] + shift;
            int dead = !(D[j] > 0.0) || !(p > tiny);
            s_isdef = dead;
            s_rjj = dead ? 0.0 : sqrt(p);
            S[(long)j * c + j] = s_rjj;
            dd[j] = dead ? 1.0 : 0.0;
            if (dead) s_def += 1;
        }
        __syncthreads();
        const double rjj = s_rjj;
        const int dead = s_isdef;
        for (int k = j + 1 + tid; k < c; k += NT)
            S[(long)j * c + k] = dead ? 0.0 : S[(long)j * c + k] / rjj;
        __syncthreads();
        if (!dead) {
            const int w = c - j - 1;
            // entries (i,k), j < i <= k: thread -> consecutive k for fixed i
            for (long t = tid; t < (long)w * w; t += NT) {
                int i = j + 1 + (int)(t / w), k = j + 1 + (int)(t % w);
                if (i <= k) S[(long)i * c + k] -= S[(long)j * c + i] * S[(long)j * c + k];
            }
        }
        __syncthreads();
    }
    // ||Rs||_F and R = Rs * D (column-major output, zeros below).
    double fr = 0.0;
    for (long e = tid; e < (long)c * c; e += NT) {
        int i = (int)(e % c), k = (int)(e / c);
        double v = (i <= k) ? S[(long)i * c + k] : 0.0;
        fr += v * v;
        R[e] = v * D[k];
    }
    fr = block_sum(fr);
    // In-place inverse of the upper triangle, column by column (row-major S:
    // row i of the inverse is contiguous); 8 lanes per row share each dot.
    for (int j = 0; j < c; j++) {
        const bool dead = dd[j] != 0.0;
        for (int i = tid; i < j; i += NT) col[i] = S[(long)i * c + j];
        __syncthreads();
        const double tjj = dead ? 0.0 : 1.0 / S[(long)j * c + j];
        const int row = tid >> 3, sub = tid & 7;
        for (int i0 = 0; i0 < j; i0 += NT / 8) {
            int i = i0 + row;
            double sacc = 0.0;
            if (i < j && !dead)
                for (int l = i + sub; l < j; l += 8) sacc += S[(long)i * c + l] * col[l];
            sacc += __shfl_xor(sacc, 1);
            sacc += __shfl_xor(sacc, 2);
            sacc += __shfl_xor(sacc, 4);
            if (i < j && sub == 0) S[(long)i * c + j] = dead ? 0.0 : -sacc * tjj;
        }
        __syncthreads();
        if (tid == 0) S[(long)j * c + j] = tjj;
        __syncthreads();
    }
    double ft = 0.0;
    for (long e = tid; e < (long)c * c; e += NT) {
        int i = (int)(e % c), k = (int)(e / c);
        double v = (i <= k) ? S[(long)i * c + k] : 0.0;
        ft += v * v;
        T[e] = (D[i] > 0.0) ? v / D[i] : 0.0;
    }
    ft = block_sum(ft);
    // lambda_max(G^-1) = ||Rs^-1||_2^2 by power iteration on Ts Ts^T
    for (int i = tid; i < c; i += NT) pv[i] = 1.0 + 0.01 * (i % 5);
    __syncthreads();
    double lminv = 0.0;
    for (int it = 0; it < PIT; it++) {
        const int row = tid >> 3, sub = tid & 7;
        // u = Ts^T x  (u_k = sum_{i<=k} Ts(i,k) x_i)
        for (int k0 = 0; k0 < c; k0 += NT / 8) {
            int k = k0 + row;
            double acc = 0.0;
            if (k < c)
                for (int i = sub; i <= k; i += 8) acc += S[(long)i * c + k] * pv[i];
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            acc += __shfl_xor(acc, 4);
            if (k < c && sub == 0) col[k] = acc;
        }
        __syncthreads();
        // y = Ts u  (y_i = sum_{k>=i} Ts(i,k) u_k)
        for (int i0 = 0; i0 < c; i0 += NT / 8) {
            int i = i0 + row;
            double acc = 0.0;
            if (i < c)
                for (int k = i + sub; k < c; k += 8) acc += S[(long)i * c + k] * col[k];
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            acc += __shfl_xor(acc, 4);
            if (i < c && sub == 0) pv[c + i] = acc;
        }
        __syncthreads();
        double nn = 0.0;
        for (int i = tid; i < c; i += NT) nn += pv[c + i] * pv[c + i];
        nn = block_sum(nn);
        double xn = 0.0;
        for (int i = tid; i < c; i += NT) xn += pv[i] * pv[i];
        xn = block_sum(xn);
        lminv = (xn > 0.0) ? sqrt(nn / xn) : 0.0;
        double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
        for (int i = tid; i < c; i += NT) pv[i] = pv[c + i] * inv;
        __syncthreads();
    }
    if (tid == 0) {
        dinfo[0] = sqrt(fr) * sqrt(ft);      // Frobenius condition (upper bound, >= c)
        dinfo[1] = want_kappa ? sqrt(lmax * lminv) : 0.0;  // 2-norm condition of Rs
        dinfo[2] = (double)s_def;
        info[0] = s_def;
    }
    (void)s_scal;
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

// INLDS: columns of A (and V when VLDS) live in LDS.  LOGROT: V is not
// kept here; every pair's rotation (c, s) is logged and replayed afterwards
// on the rows of V by k_svd_replay_v (rows of V evolve independently).
template <bool INLDS, bool VLDS, bool LOGROT>
__global__ __launch_bounds__(NT) void k_svd_jacobi(int m, int k, const double *__restrict__ A,
                                                   long lda, double *__restrict__ sig,
                                                   double *__restrict__ U, long ldu,
                                                   double *__restrict__ V, long ldv,
                                                   double log_scale, double delta, int cap,
                                                   int *iout, double *gwork,
                                                   double2 *__restrict__ rotlog,
                                                   int *__restrict__ posout) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    double *W = INLDS ? lds : gwork;  // m x k, ld m
    double *Vw = (V && !LOGROT) ? (VLDS ? lds + (long)m * k : gwork + (long)m * k) : nullptr;
    __shared__ int s_rot;
    __shared__ double s_sig[1024];
    __shared__ double s_sorted[1024];
    __shared__ int s_pos[1024];

    // Exactly-zero columns (deficient directions) take no part in the
    // rotations: nonzero columns are packed first, in their original order.
    __shared__ int s_perm[1024];
    __shared__ int s_knz;
    for (int j = tid; j < k; j += NT) {
        double s = 0.0;
        for (int i = 0; i < m; i++) s = fmax(s, fabs(A[i + (long)j * lda]));
        s_sig[j] = s;
    }
    __syncthreads();
    if (tid == 0) {
        int q = 0;
        for (int j = 0; j < k; j++)
            if (s_sig[j] > 0.0) s_perm[q++] = j;
        s_knz = q;
        for (int j = 0; j < k; j++)
            if (!(s_sig[j] > 0.0)) s_perm[q++] = j;
    }
    __syncthreads();
    for (long e = tid; e < (long)m * k; e += NT) {
        long j = e / m, i = e - j * m;
        W[e] = A[i + (long)s_perm[j] * lda];
    }
    if (Vw)
        for (long e = tid; e < (long)k * k; e += NT) {
            long j = e / k, i = e - j * k;
            Vw[e] = (i == s_perm[j]) ? 1.0 : 0.0;
        }
    if (LOGROT)
        for (int j = tid; j < k; j += NT) posout[1024 + j] = s_perm[j];
    __syncthreads();
    const int knz = s_knz;

    const int kp = knz + (knz & 1);
    const int P = kp / 2;
    int G = 64;
    while (G > 1 && G * P > NT) G >>= 1;
    const int ngroups = NT / G;
    const int g = tid / G, gl = tid - g * G;
    // Rotate only when the pair's cosine exceeds the rounding noise of an
    // m-term dot product; below that the columns are orthogonal to working
    // accuracy (stricter thresholds never settle).
    const double tol = 0x1.0p-53 * (double)m;
    int sweep = 0;
    const int max_sweeps = 40;
    for (; sweep < max_sweeps && knz >= 2; sweep++) {
        if (tid == 0) s_rot = 0;
        __syncthreads();
        int rot_local = 0;
        for (int t = 0; t < kp - 1; t++) {
            for (int p0 = 0; p0 < P; p0 += ngroups) {
                int p = p0 + g;
                int a = -1, b = -1;
                if (p < P) rr_pair(kp, t, p, a, b);
                bool valid = (p < P) && a < knz && b < knz;
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
                bool rot = valid && fabs(ga) > tol * sqrt(al * be) && al > 0.0 && be > 0.0;
                if (LOGROT && p < P && gl == 0) {
                    double2 cslog = make_double2(1.0, 0.0);
                    if (rot) {
                        double zeta = (be - al) / (2.0 * ga);
                        double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        double cs = 1.0 / sqrt(1.0 + tt * tt);
                        cslog = make_double2(cs, cs * tt);
                    }
                    rotlog[((long)sweep * (kp - 1) + t) * P + p] = cslog;
                }
                if (rot) {
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
        if (LOGROT) posout[j] = s_pos[j];
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
        iout[2] = (sweep < max_sweeps) ? sweep + 1 : max_sweeps;
        // sweeps written to the log: the loop breaks after a rotation-free
        // sweep (logged as identities) or stops at max_sweeps
        if (LOGROT) iout[3] = (sweep < max_sweeps) ? sweep + 1 : max_sweeps;
        if (LOGROT) iout[4] = knz;
    }
}

// Replays the logged Jacobi rotations on the rows of V (V starts as I):
// one wave per row, the row kept in LDS; each round applies the P disjoint
// pair rotations in parallel.  Row i of V ends at V[i, pos[j]] = row[j].
__global__ __launch_bounds__(64) void k_svd_replay_v(int k, const double2 *__restrict__ rotlog,
                                                    const int *__restrict__ nsw,
                                                    const int *__restrict__ pos,
                                                    double *__restrict__ V, long ldv) {
    __shared__ double row[1024];
    const int i = blockIdx.x, lane = threadIdx.x;
    const int knz = nsw[1];
    const int kp = knz + (knz & 1), P = kp / 2;
    const int *perm = pos + 1024;
    for (int j = lane; j < 1024; j += 64) row[j] = (j < k && perm[j] == i) ? 1.0 : 0.0;
    __syncthreads();
    const int sweeps = *nsw;
    for (int sw = 0; sw < sweeps; sw++)
        for (int t = 0; t < kp - 1; t++) {
            for (int p = lane; p < P; p += 64) {
                int a, b;
                rr_pair(kp, t, p, a, b);
                double2 cs = rotlog[((long)sw * (kp - 1) + t) * P + p];
                if (a < knz && b < knz && cs.y != 0.0) {
                    double x = row[a], y = row[b];
                    row[a] = cs.x * x - cs.y * y;
                    row[b] = cs.y * x + cs.x * y;
                }
            }
            __syncthreads();
        }
    for (int j = lane; j < k; j += 64) V[i + (long)pos[j] * ldv] = row[j];
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

template <typename KF>
void set_lds_attr(KF kern, int ndoubles) {
    HIPCHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 ndoubles * (int)sizeof(double)));
}

void allow_big_lds() {
    static bool done = false;
    if (done) return;
    set_lds_attr(k_chol_qr_factor<true>, LDS_DOUBLES);
    set_lds_attr(k_svd_jacobi<true, true, false>, SVD_LDS_DOUBLES);
    set_lds_attr(k_svd_jacobi<true, false, true>, SVD_LDS_DOUBLES);
    done = true;
}

}  // namespace

extern "C" {

void ptt_hip_chol_qr_factor(int c, const double *G, double shift, double tiny, double *R,
                            double *T, int *info, double *dinfo, int want_kappa) {
    ptt_hip_init();
    allow_big_lds();
    size_t need = (size_t)c * c + 5 * (size_t)c;
    if (need <= (size_t)LDS_DOUBLES) {
        PTT_LAUNCH("chol_qr_factor", k_chol_qr_factor<true>, dim3(1), dim3(NT),
                   need * sizeof(double), c, G, shift, tiny, R, T, info, dinfo, (double *)nullptr,
                   want_kappa);
    } else {
        double *gwork = (double *)ptt_hip_scratch(PTT_WS_SKETCH, need * sizeof(double));
        PTT_LAUNCH("chol_qr_factor", k_chol_qr_factor<false>, dim3(1), dim3(NT), 0, c, G, shift,
                   tiny, R, T, info, dinfo, gwork, want_kappa);
    }
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
    const size_t mk = (size_t)m * k, kk = (size_t)k * k;
    if (mk + (V ? kk : 0) <= (size_t)SVD_LDS_DOUBLES) {
        PTT_LAUNCH("svd_jacobi", (k_svd_jacobi<true, true, false>), dim3(1), dim3(NT),
                   (mk + (V ? kk : 0)) * sizeof(double), m, k, A, lda, sig, U, ldu, V, ldv,
                   log_scale, delta, cap, iout, (double *)nullptr, (double2 *)nullptr,
                   (int *)nullptr);
    } else if (mk <= (size_t)SVD_LDS_DOUBLES) {
        // columns in LDS; V (if wanted) from the rotation log
        const int kp = k + (k & 1), P = kp / 2;
        const size_t logn = (size_t)40 * (kp - 1) * P;
        char *ws = (char *)ptt_hip_scratch(PTT_WS_SKETCH, logn * sizeof(double2) + 2048 * sizeof(int));
        double2 *rotlog = (double2 *)ws;
        int *pos = (int *)(ws + logn * sizeof(double2));
        PTT_LAUNCH("svd_jacobi", (k_svd_jacobi<true, false, true>), dim3(1), dim3(NT),
                   mk * sizeof(double), m, k, A, lda, sig, U, ldu, V, ldv, log_scale, delta, cap,
                   iout, (double *)nullptr, rotlog, pos);
        if (V)
            PTT_LAUNCH("svd_replay_v", k_svd_replay_v, dim3(k), dim3(64), 0, k, rotlog, iout + 3,
                       pos, V, ldv);
    } else {
        double *gwork = (double *)ptt_hip_scratch(PTT_WS_SKETCH, (mk + (V ? kk : 0)) * sizeof(double));
        PTT_LAUNCH("svd_jacobi", (k_svd_jacobi<false, false, false>), dim3(1), dim3(NT), 0, m, k,
                   A, lda, sig, U, ldu, V, ldv, log_scale, delta, cap, iout, gwork,
                   (double2 *)nullptr, (int *)nullptr);
    }
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
