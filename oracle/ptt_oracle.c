/*
 * ptt_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room FP64 CPU restatement of the reference's TT hot path at one
 * MPI rank, used as the checker for the GPU library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; no
 * product code links or calls it.
 *
 * It calls the same LAPACK/BLAS routines with the same arguments as the
 * reference (LAPACK 3.9.1 API; netlib LAPACK from scipy's OpenBLAS), so its rounding
 * fingerprint is the reference's up to BLAS summation order.  Each function
 * cites the reference lines it restates.  Pinning (see tests/test_oracle_*):
 *   - against golden vectors written by the compiled reference
 *     (tests/golden/, produced by tests/golden/gen_golden.c);
 *   - against mathematics: orthogonality, tensor invariance, brute-force
 *     TT-SVD of small full tensors, closed-form stencil eigenvalues,
 *     exact entries of formal sums and Hadamard products.
 *
 * Layout: core k is column-major r_k x n_k x r_{k+1}, element (a,j,b) at
 * a + r_k*(j + n_k*b) (source/paratt.h:38-42).  Cores are host arrays owned
 * by the caller, sized for the largest shape they will hold.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* LAPACK/BLAS: the netlib LAPACK inside scipy's bundled OpenBLAS (LP64),
 * whose Fortran symbols carry a "scipy_" prefix -- the same library the
 * compiled reference is linked to (oracle/build_ref.sh). */
#define dgeqrf_ scipy_dgeqrf_
#define dorgqr_ scipy_dorgqr_
#define dgelqf_ scipy_dgelqf_
#define dorglq_ scipy_dorglq_
#define dgesvdq_ scipy_dgesvdq_
#define dgemm_ scipy_dgemm_
#define dtrmm_ scipy_dtrmm_
#define dlange_ scipy_dlange_
#define dsyev_ scipy_dsyev_

extern void dgeqrf_(const int *, const int *, double *, const int *, double *,
                    double *, const int *, int *);
extern void dorgqr_(const int *, const int *, const int *, double *,
                    const int *, const double *, double *, const int *, int *);
extern void dgelqf_(const int *, const int *, double *, const int *, double *,
                    double *, const int *, int *);
extern void dorglq_(const int *, const int *, const int *, double *,
                    const int *, const double *, double *, const int *, int *);
extern void dgesvdq_(const char *, const char *, const char *, const char *,
                     const char *, const int *, const int *, double *,
                     const int *, double *, double *, const int *, double *,
                     const int *, int *, int *, const int *, double *,
                     int *, double *, const int *, int *);
extern void dgemm_(const char *, const char *, const int *, const int *,
                   const int *, const double *, const double *, const int *,
                   const double *, const int *, const double *, double *,
                   const int *);
extern void dtrmm_(const char *, const char *, const char *, const char *,
                   const int *, const int *, const double *, const double *,
                   const int *, double *, const int *);
extern double dlange_(const char *, const int *, const int *, const double *,
                      const int *, double *);

static void *xmalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 8);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
    return p;
}
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* Seeded input generator (project-defined, not in the reference).    */
/* counter-based: splitmix64 finaliser -> two uniforms -> Box-Muller.  */
/* ------------------------------------------------------------------ */
static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
/* Entry e of core `core` of the TT with seed `seed`: N(0,1) * scale. */
void o_gauss_fill(uint64_t seed, int core, double scale, long count,
                  double *out) {
    uint64_t h = mix64(mix64(seed) + (uint64_t)core);
    for (long e = 0; e < count; e += 2) {
        uint64_t p = (uint64_t)(e >> 1);
        uint64_t a = mix64(h + 2 * p), b = mix64(h + 2 * p + 1);
        double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
        double u2 = (double)(b >> 11) * 0x1.0p-53;
        double rad = sqrt(-2.0 * log(u1));
        double th = 6.283185307179586476925286766559 * u2;
        out[e] = rad * cos(th) * scale;
        if (e + 1 < count) out[e + 1] = rad * sin(th) * scale;
    }
}

/* ------------------------------------------------------------------ */
/* Formal sum / axpby / Hadamard                                      */
/* ------------------------------------------------------------------ */
/* attac.c:81-112 (attacFormalAXPBY). alpha, beta act on core 0 only;
 * core d-1 stacks x over y; interior cores are block diagonal per slice. */
void o_axpby(int d, const int *n, double alpha, const int *rx, double **x,
             double beta, const int *ry, double **y, int *rz, double **z) {
    rz[0] = 1;
    rz[d] = 1;
    for (int i = 1; i < d; i++) rz[i] = rx[i] + ry[i];
    for (long e = 0; e < (long)n[0] * rx[1]; e++) z[0][e] = x[0][e] * alpha;
    for (long e = 0; e < (long)n[0] * ry[1]; e++)
        z[0][(long)rx[1] * n[0] + e] = y[0][e] * beta;
    {
        int k = d - 1;
        for (int j = 0; j < n[k]; j++) {
            for (int a = 0; a < rx[k]; a++)
                z[k][a + (long)rz[k] * j] = x[k][a + (long)rx[k] * j];
            for (int a = 0; a < ry[k]; a++)
                z[k][rx[k] + a + (long)rz[k] * j] = y[k][a + (long)ry[k] * j];
        }
    }
    for (int i = 1; i < d - 1; i++) {
        long zn = (long)rz[i] * n[i];
        memset(z[i], 0, sizeof(double) * zn * rz[i + 1]);
        for (int j = 0; j < n[i]; j++) {
            for (int b = 0; b < rx[i + 1]; b++)
                for (int a = 0; a < rx[i]; a++)
                    z[i][a + (long)rz[i] * j + zn * b] =
                        x[i][a + (long)rx[i] * j + (long)rx[i] * n[i] * b];
            for (int b = 0; b < ry[i + 1]; b++)
                for (int a = 0; a < ry[i]; a++)
                    z[i][rx[i] + a + (long)rz[i] * j + zn * (rx[i + 1] + b)] =
                        y[i][a + (long)ry[i] * j + (long)ry[i] * n[i] * b];
        }
    }
}

/* attac.c:128-136 with utilities.c:112-129: per mode slice j,
 * z[ry*ax+ay, j, ry'*bx+by] = x[ax,j,bx] * y[ay,j,by]. */
void o_hadamard(int d, const int *n, const int *rx, double **x, const int *ry,
                double **y, double **z) {
    for (int i = 0; i < d; i++) {
        int zr = rx[i] * ry[i], zr1 = rx[i + 1] * ry[i + 1];
        for (int j = 0; j < n[i]; j++)
            for (int bx = 0; bx < rx[i + 1]; bx++)
                for (int ax = 0; ax < rx[i]; ax++)
                    for (int by = 0; by < ry[i + 1]; by++)
                        for (int ay = 0; ay < ry[i]; ay++) {
                            long zi = (long)(ry[i] * ax + ay) +
                                      (long)zr * (j + (long)n[i] * (ry[i + 1] * bx + by));
                            z[i][zi] =
                                x[i][ax + (long)rx[i] * (j + (long)n[i] * bx)] *
                                y[i][ay + (long)ry[i] * (j + (long)n[i] * by)];
                        }
        (void)zr1;
    }
}

/* ------------------------------------------------------------------ */
/* Orthogonalisation and truncation (P = 1)                            */
/* ------------------------------------------------------------------ */
/* pattac.c:244-294 with tsqr.c:541-714/716-1045 at one rank (the tree is a
 * single node: dgeqrf + dorgqr) and utilities.c:233-308 (trapezoid_mult 'l'). */
void o_lorth(int d, const int *n, int *r, double **x) {
    for (int i = 0; i < d - 1; i++) {
        int m = r[i] * n[i], c = r[i + 1], k = imin(m, c), info = 0, lw = -1;
        double *Q = xmalloc(sizeof(double) * (size_t)m * c);
        double *tau = xmalloc(sizeof(double) * (size_t)imax(k, 1));
        double q = 0;
        memcpy(Q, x[i], sizeof(double) * (size_t)m * c);
        dgeqrf_(&m, &c, Q, &m, tau, &q, &lw, &info);
        lw = (int)q;
        double *wk = xmalloc(sizeof(double) * (size_t)imax(lw, 1));
        dgeqrf_(&m, &c, Q, &m, tau, wk, &lw, &info);
        free(wk);
        /* R: k x c upper trapezoid (tsqr.c:666-670 pack, pattac.c:271-272 unpack) */
        double *R = xmalloc(sizeof(double) * (size_t)k * c);
        for (int col = 0; col < c; col++)
            for (int row = 0; row < k; row++)
                R[row + (long)k * col] = row <= col ? Q[row + (long)m * col] : 0.0;
        /* trapezoid_mult('l'): x[i+1] (c x w, ld c) <- R * x[i+1] */
        int w = n[i + 1] * r[i + 2];
        double one = 1.0;
        dtrmm_("L", "U", "N", "N", &k, &w, &one, R, &k, x[i + 1], &c);
        if (c > k) {
            int del = c - k;
            dgemm_("N", "N", &k, &w, &del, &one, R + (long)k * k, &k,
                   x[i + 1] + k, &c, &one, x[i + 1], &c);
            /* unpad_rows: ld c -> ld k */
            for (int col = 1; col < w; col++)
                for (int row = 0; row < k; row++)
                    x[i + 1][row + (long)k * col] = x[i + 1][row + (long)c * col];
        }
        r[i + 1] = k;
        /* dorgqr on the factored copy, then copy out m x k (tsqr.c:745-777, 1034-1038) */
        lw = -1;
        dorgqr_(&m, &k, &k, Q, &m, tau, &q, &lw, &info);
        lw = (int)q;
        wk = xmalloc(sizeof(double) * (size_t)imax(lw, 1));
        dorgqr_(&m, &k, &k, Q, &m, tau, wk, &lw, &info);
        free(wk);
        memcpy(x[i], Q, sizeof(double) * (size_t)m * k);
        free(Q); free(tau); free(R);
    }
}

/* pattac.c:297-522 at one rank: right-to-left, LQ of each core
 * (tsqr.c:1048-1222, 1225-1554 = dgelqf + dorglq), SVD of L by
 * dgesvdq('A','P','N','R','R'), tail rule with delta = abs_tol/sqrt(d-1)
 * (strict <), caps k and maxr[i], floor 1.
 * If sig_log != NULL, bond i's singular values go to sig_log[i*ldsig + j]
 * (j < k), numrank_log[i] gets NUMRANK. */
void o_ltruncate(int d, const int *n, int *r, double **x, const int *maxr,
                 double abs_tol, double *sig_log, int ldsig, int *numrank_log) {
    const double delta = abs_tol / sqrt(d - 1.0);
    for (int i = d - 1; i > 0; i--) {
        int rows = r[i], w = n[i] * r[i + 1], k = imin(rows, w), info = 0;
        int lw = -1;
        double q = 0;
        double *Q = xmalloc(sizeof(double) * (size_t)rows * w);
        double *tau = xmalloc(sizeof(double) * (size_t)imax(k, 1));
        memcpy(Q, x[i], sizeof(double) * (size_t)rows * w);
        dgelqf_(&rows, &w, Q, &rows, tau, &q, &lw, &info);
        lw = (int)q;
        double *wk = xmalloc(sizeof(double) * (size_t)imax(lw, 1));
        dgelqf_(&rows, &w, Q, &rows, tau, wk, &lw, &info);
        free(wk);
        /* L: rows x k lower trapezoid, ld rows (lower_tri_unpack, pattac.c:417-421) */
        double *L = xmalloc(sizeof(double) * (size_t)rows * k);
        for (int col = 0; col < k; col++)
            for (int row = 0; row < rows; row++)
                L[row + (long)rows * col] = row >= col ? Q[row + (long)rows * col] : 0.0;
        /* dgesvdq workspace query then run (pattac.c:370-452) */
        int numrank = k, liw = -1, lrw = -1, lwk = -1, ldv = k;
        int iwq[4];
        double wq[4], rwq[4];
        double *S = xmalloc(sizeof(double) * (size_t)k);
        double *U = xmalloc(sizeof(double) * (size_t)rows * k);
        double *VT = xmalloc(sizeof(double) * (size_t)k * k);
        dgesvdq_("A", "P", "N", "R", "R", &rows, &k, L, &rows, S, U, &rows, VT,
                 &ldv, &numrank, iwq, &liw, wq, &lwk, rwq, &lrw, &info);
        lwk = (int)wq[0] + 1;
        lrw = (int)rwq[0] + 1;
        liw = iwq[0];
        int *iw = xmalloc(sizeof(int) * (size_t)imax(liw, 1));
        double *sw = xmalloc(sizeof(double) * (size_t)lwk);
        double *rw = xmalloc(sizeof(double) * (size_t)lrw);
        dgesvdq_("A", "P", "N", "R", "R", &rows, &k, L, &rows, S, U, &rows, VT,
                 &ldv, &numrank, iw, &liw, sw, &lwk, rw, &lrw, &info);
        if (info != 0) {
            fprintf(stderr, "oracle: dgesvdq info=%d at core %d\n", info, i);
            exit(1);
        }
        if (sig_log) {
            for (int j = 0; j < k && j < ldsig; j++)
                sig_log[(long)i * ldsig + j] = j < numrank ? S[j] : 0.0;
        }
        if (numrank_log) numrank_log[i] = numrank;
        /* tail rule (pattac.c:468-481) */
        double tail = 0.0;
        int rho = numrank;
        for (int j = numrank - 1; j >= 0; j--) {
            tail += S[j] * S[j];
            if (sqrt(tail) < delta) rho = j; else break;
        }
        rho = imin(rho, k);
        rho = imin(rho, maxr[i]);
        rho = imax(rho, 1);
        /* U_rho * diag(S) (pattac.c:483-488) */
        double *US = xmalloc(sizeof(double) * (size_t)rows * rho);
        for (int j = 0; j < rho; j++)
            for (int a = 0; a < rows; a++)
                US[a + (long)rows * j] = S[j] * U[a + (long)rows * j];
        /* explicit Q (k x w) from dorglq, rows copied to ld k (tsqr.c:1253-1286, 1541-1547) */
        lw = -1;
        dorglq_(&k, &w, &k, Q, &rows, tau, &q, &lw, &info);
        lw = (int)q;
        wk = xmalloc(sizeof(double) * (size_t)imax(lw, 1));
        dorglq_(&k, &w, &k, Q, &rows, tau, wk, &lw, &info);
        free(wk);
        double *Qk = xmalloc(sizeof(double) * (size_t)k * w);
        for (int t = 0; t < w; t++)
            memcpy(Qk + (long)t * k, Q + (long)t * rows, sizeof(double) * k);
        /* x[i] <- VT[:rho,:] * Q (pattac.c:490-502) */
        double one = 1.0, zero = 0.0;
        dgemm_("N", "N", &rho, &w, &k, &one, VT, &ldv, Qk, &k, &zero, x[i], &rho);
        /* x[i-1] <- x[i-1] * U S (pattac.c:504-519) */
        int m1 = r[i - 1] * n[i - 1];
        double *tmp = xmalloc(sizeof(double) * (size_t)m1 * rho);
        dgemm_("N", "N", &m1, &rho, &rows, &one, x[i - 1], &m1, US, &rows, &zero,
               tmp, &m1);
        memcpy(x[i - 1], tmp, sizeof(double) * (size_t)m1 * rho);
        r[i] = rho;
        free(tmp); free(Qk); free(US); free(iw); free(sw); free(rw);
        free(S); free(U); free(VT); free(L); free(Q); free(tau);
    }
}

/* pattac.c:590-675 (ortho 'R', tol = abs_tol) with paratt.c:316-342. */
void o_truncate(int d, const int *n, int *r, double **x, const int *maxr,
                double abs_tol, double *sig_log, int ldsig, int *numrank_log) {
    int *mr = xmalloc(sizeof(int) * (size_t)(d + 1));
    memcpy(mr, maxr, sizeof(int) * (size_t)(d + 1)); /* maxRank may alias r */
    for (int i = 0; i < d; i++) {
        double s = sqrt((double)n[i]);
        long cnt = (long)r[i] * n[i] * r[i + 1];
        for (long e = 0; e < cnt; e++) x[i][e] /= s;
    }
    o_lorth(d, n, r, x);
    o_ltruncate(d, n, r, x, mr, abs_tol, sig_log, ldsig, numrank_log);
    for (int i = 0; i < d; i++) {
        double s = sqrt((double)n[i]);
        long cnt = (long)r[i] * n[i] * r[i + 1];
        for (long e = 0; e < cnt; e++) x[i][e] *= s;
    }
    free(mr);
}

/* pattac.c:693-741: a = x0^T y0; for i>=1: b = a*y_i, a = x_i^T b. */
double o_inner(int d, const int *n, const int *rx, double **x, const int *ry,
               double **y) {
    int rmax = 1;
    for (int i = 0; i <= d; i++) rmax = imax(rmax, imax(rx[i], ry[i]));
    int nmax = 1;
    for (int i = 0; i < d; i++) nmax = imax(nmax, n[i]);
    double *a = xmalloc(sizeof(double) * (size_t)rmax * rmax);
    double *b = xmalloc(sizeof(double) * (size_t)rmax * rmax * nmax);
    double one = 1.0, zero = 0.0;
    int rxn = rx[0] * n[0], ryn = ry[0] * n[0];
    dgemm_("T", "N", &rx[1], &ry[1], &n[0], &one, x[0], &rxn, y[0], &ryn, &zero,
           a, &rx[1]);
    for (int i = 1; i < d; i++) {
        rxn = rx[i] * n[i];
        int nry = n[i] * ry[i + 1];
        dgemm_("N", "N", &rx[i], &nry, &ry[i], &one, a, &rx[i], y[i], &ry[i],
               &zero, b, &rx[i]);
        dgemm_("T", "N", &rx[i + 1], &ry[i + 1], &rxn, &one, x[i], &rxn, b,
               &rxn, &zero, a, &rx[i + 1]);
    }
    double v = a[0];
    free(a); free(b);
    return v;
}

static void scale_cores(int d, const int *n, const int *r, double **x, int up) {
    for (int i = 0; i < d; i++) {
        double s = sqrt((double)n[i]);
        long cnt = (long)r[i] * n[i] * r[i + 1];
        if (up) for (long e = 0; e < cnt; e++) x[i][e] *= s;
        else    for (long e = 0; e < cnt; e++) x[i][e] /= s;
    }
}

/* paratt.c:412-469: scale x (and y if distinct) by 1/sqrt(n), inner
 * product, unscale.  Mutates the cores by the rounding of that round trip,
 * as the reference does. */
double o_dotprod(int d, const int *n, const int *rx, double **x, const int *ry,
                 double **y, int same) {
    scale_cores(d, n, rx, x, 0);
    if (!same) scale_cores(d, n, ry, y, 0);
    double v = o_inner(d, n, rx, x, ry, y);
    scale_cores(d, n, rx, x, 1);
    if (!same) scale_cores(d, n, ry, y, 1);
    return v;
}

/* paratt.c:369-409: scale, left-orthogonalise (x keeps that form), Frobenius
 * norm of the last core, unscale. */
double o_norm(int d, const int *n, int *r, double **x) {
    scale_cores(d, n, r, x, 0);
    o_lorth(d, n, r, x);
    int m = r[d - 1];
    double v = dlange_("F", &m, &n[d - 1], x[d - 1], &m, NULL);
    scale_cores(d, n, r, x, 1);
    return v;
}

/* ------------------------------------------------------------------ */
/* Core-wise operators at one rank (paratt.c:271-281, 571-1250)         */
/* f is core `mode`: r x nn x r1.                                       */
/* ------------------------------------------------------------------ */
#define F(a, j, b) f[(a) + (long)r * ((j) + (long)nn * (b))]

/* paratt.c:271-281: scales core 0 only. */
void o_scale(int n0, int r1, double alpha, double *f) {
    for (long e = 0; e < (long)n0 * r1; e++) f[e] *= alpha;
}

/* paratt.c:571-581 */
void o_diag(int r, int nn, int r1, const double *v, double *f) {
    for (int a = 0; a < r; a++)
        for (int j = 0; j < nn; j++)
            for (int b = 0; b < r1; b++) F(a, j, b) *= v[j];
}

/* paratt.c:584-661 (serial periodic stencil: correct wrap indices). */
void o_periodic_stencil_serial(int r, int nn, int r1, const double *sten,
                               int slen, double *f) {
    slen = (slen % 2) ? slen : slen - 1;
    int h = (slen - 1) / 2;
    double *out = xmalloc(sizeof(double) * (size_t)nn);
    for (int b = 0; b < r1; b++)
        for (int a = 0; a < r; a++) {
            for (int i = 0; i < nn; i++) {
                double acc = 0.0;
                if (i < h) {
                    for (int j = 0; j < h - i; j++) acc += sten[j] * F(a, nn - h + j + i, b);
                    for (int j = 0; j <= i + h; j++) acc += sten[j + h - i] * F(a, j, b);
                } else if (i < nn - h) {
                    for (int j = 0; j < slen; j++) acc += sten[j] * F(a, i + j - h, b);
                } else {
                    for (int j = i - h; j < nn; j++) acc += sten[j + h - i] * F(a, j, b);
                    int nb = h - (nn - i) + 1;
                    for (int j = 0; j < nb; j++) acc += sten[slen - nb + j] * F(a, j, b);
                }
                out[i] = acc;
            }
            for (int i = 0; i < nn; i++) F(a, i, b) = out[i];
        }
    free(out);
}

/* paratt.c:665-770 at one rank.  Ghost rows above are f[n-h .. n-1]
 * (:689-704, :722-723) and are consumed with a counter that restarts at 0
 * for each output row (:738-744); so for output row i < h the j-th
 * negative offset reads f[n - h + (j - (i - h))] = f[n + j - i]
 * (equal to the periodic f[n + j] only for i = 0). */
void o_periodic_stencil(int r, int nn, int r1, const double *sten, int slen,
                        double *f) {
    slen = (slen % 2) ? slen : slen - 1;
    int h = (slen - 1) / 2;
    double *out = xmalloc(sizeof(double) * (size_t)nn);
    for (int b = 0; b < r1; b++)
        for (int a = 0; a < r; a++) {
            for (int i = 0; i < nn; i++) {
                double acc = 0.0;
                int sc = 0, gc = 0;
                for (int j = i - h; j < i; j++) {
                    double v;
                    if (j < 0) { v = F(a, nn - h + gc, b); gc++; }
                    else v = F(a, j, b);
                    acc += sten[sc++] * v;
                }
                acc += sten[sc++] * F(a, i, b);
                gc = 0;
                for (int j = i + 1; j <= i + h; j++) {
                    double v;
                    if (j < nn) v = F(a, j, b);
                    else { v = F(a, gc, b); gc++; }
                    acc += sten[sc++] * v;
                }
                out[i] = acc;
            }
            for (int i = 0; i < nn; i++) F(a, i, b) = out[i];
        }
    free(out);
}

/* paratt.c:772-923 at one rank: mirror ghosts.  Above: ghost g holds
 * f[h-1-g] (:826-832), read with a per-row counter; below: ghost g holds
 * f[n-1-g] (:853-859). */
void o_cdf_stencil(int r, int nn, int r1, const double *sten, int slen,
                   double *f) {
    slen = (slen % 2) ? slen : slen - 1;
    int h = (slen - 1) / 2;
    double *out = xmalloc(sizeof(double) * (size_t)nn);
    for (int b = 0; b < r1; b++)
        for (int a = 0; a < r; a++) {
            for (int i = 0; i < nn; i++) {
                double acc = 0.0;
                int sc = 0, gc = 0;
                for (int j = i - h; j < i; j++) {
                    double v;
                    if (j < 0) { v = F(a, h - 1 - gc, b); gc++; }
                    else v = F(a, j, b);
                    acc += sten[sc++] * v;
                }
                acc += sten[sc++] * F(a, i, b);
                gc = 0;
                for (int j = i + 1; j <= i + h; j++) {
                    double v;
                    if (j < nn) v = F(a, j, b);
                    else { v = F(a, nn - 1 - gc, b); gc++; }
                    acc += sten[sc++] * v;
                }
                out[i] = acc;
            }
            for (int i = 0; i < nn; i++) F(a, i, b) = out[i];
        }
    free(out);
}

/* paratt.c:926-1093 at one rank (h = 1, mirror ghosts):
 * out[i] = sten0*(-c[i-1] f[i-1]) + sten2*(-c[i+1] f[i+1]), with
 * sten = [-0.5/dx, 0, 0.5/dx], ghost speeds gc[0] (left) and gc[1] (right);
 * then f[i] <- out[i] + s[i]*f[i]. */
void o_lf_flux(int r, int nn, int r1, const double *c, const double *gc,
               const double *s, double dt, double dx, double *f) {
    (void)dt;
    double sten0 = -0.5 / dx, sten2 = 0.5 / dx;
    double *out = xmalloc(sizeof(double) * (size_t)nn);
    for (int b = 0; b < r1; b++)
        for (int a = 0; a < r; a++) {
            for (int i = 0; i < nn; i++) {
                double acc = 0.0, v;
                if (i - 1 < 0) v = -gc[0] * F(a, 0, b);
                else v = -c[i - 1] * F(a, i - 1, b);
                acc += sten0 * v;
                if (i + 1 < nn) v = -c[i + 1] * F(a, i + 1, b);
                else v = -gc[1] * F(a, nn - 1, b);
                acc += sten2 * v;
                out[i] = acc;
            }
            for (int i = 0; i < nn; i++) F(a, i, b) = out[i] + s[i] * F(a, i, b);
        }
    free(out);
}

/* paratt.c:1096-1250 at one rank (Lax-Wendroff, mirror ghosts g0 = f[0],
 * g1 = f[n-1]); bw has n+1 entries. */
void o_lw_rhs(int r, int nn, int r1, const double *c, const double *gc,
              const double *bw, const double *s, double dt, double dx,
              double *f) {
    double *out = xmalloc(sizeof(double) * (size_t)nn);
    for (int b = 0; b < r1; b++)
        for (int a = 0; a < r; a++) {
            double g0 = F(a, 0, b), g1 = F(a, nn - 1, b);
            double uh0 = 0.5 * (g0 + F(a, 0, b)) -
                         ((c[0] * F(a, 0, b) - g0 * gc[0]) * dt / (2.0 * dx));
            for (int i = 0; i < nn; i++) {
                double uh1;
                if (i + 1 < nn)
                    uh1 = 0.5 * (F(a, i + 1, b) + F(a, i, b)) -
                          ((c[i + 1] * F(a, i + 1, b) - c[i] * F(a, i, b)) * dt / (2.0 * dx));
                else
                    uh1 = 0.5 * (g1 + F(a, i, b)) -
                          ((g1 * gc[1] - c[i] * F(a, i, b)) * dt / (2.0 * dx));
                out[i] = -(bw[i + 1] * uh1 - bw[i] * uh0) / dx + s[i] * F(a, i, b);
                uh0 = uh1;
            }
            for (int i = 0; i < nn; i++) F(a, i, b) = out[i];
        }
    free(out);
}
#undef F
