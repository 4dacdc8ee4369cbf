/*
 * ops.c -- TT algebra and core-wise operators on device-resident cores:
 * formal axpby/sum (attac.c:37-112, paratt.c:237-268), Hadamard
 * (attac.c:128-136), scaling (paratt.c:271-281), the mode operators of
 * paratt.c:571-1250 at one rank, and the inner-product sweep
 * (pattac.c:693-741, paratt.c:412-469).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "paratt.h"
#include "paratt_amd.h"
#include "ptt_hip.h"
#include "ptt_host.h"

static long core_count3(int r0, int n, int r1) { return (long)r0 * n * r1; }

/* ---------------- formal sums ---------------- */

void attacFormalAXPBY(const int d, const int *n, const double alpha, const int *rx,
                      double *const *x, const double beta, const int *ry, double *const *y,
                      int *rz, double **z) {
    if (d < 2) ptt_host_fatal("attacFormalAXPBY", "needs at least two modes");
    rz[0] = 1;
    rz[d] = 1;
    for (int i = 1; i < d; i++) rz[i] = rx[i] + ry[i];
    ptt_hip_axpby_core(0, n[0], rx[0], rx[1], x[0], ry[0], ry[1], y[0], alpha, beta, z[0]);
    for (int i = 1; i < d - 1; i++)
        ptt_hip_axpby_core(1, n[i], rx[i], rx[i + 1], x[i], ry[i], ry[i + 1], y[i], 1.0, 1.0, z[i]);
    ptt_hip_axpby_core(2, n[d - 1], rx[d - 1], rx[d], x[d - 1], ry[d - 1], ry[d], y[d - 1], 1.0,
                       1.0, z[d - 1]);
    ptt_hip_sync();
}

void pattacFormalAXPBY(const int d, const int *n, const double alpha, const int *rx,
                       double *const *x, const double beta, const int *ry, double *const *y,
                       int *rz, double **z) {
    attacFormalAXPBY(d, n, alpha, rx, x, beta, ry, y, rz, z);
}

void attacFormalSum(const int d, const int *n, const int *rx, double *const *x, const int *ry,
                    double *const *y, int *rz, double **z) {
    attacFormalAXPBY(d, n, 1.0, rx, x, 1.0, ry, y, rz, z);
}

void pattacFormalSum(const int d, const int *n, const int *rx, double *const *x, const int *ry,
                     double *const *y, int *rz, double **z) {
    attacFormalAXPBY(d, n, 1.0, rx, x, 1.0, ry, y, rz, z);
}

void attac_tt_axpby(double a, attac_tt *x, double b, attac_tt *y, attac_tt *z) {
    int d = z->numDim;
    if (z == x || z == y) ptt_host_fatal("attac_tt_axpby", "z must not alias x or y");
    for (int k = 0; k < d; k++) {
        int r0 = (k == 0) ? 1 : x->ttRank[k] + y->ttRank[k];
        int r1 = (k == d - 1) ? 1 : x->ttRank[k + 1] + y->ttRank[k + 1];
        ptt_host_reserve_core(z, k, core_count3(r0, z->dim[k], r1));
        z->ttRank[k] = r0;
    }
    z->ttRank[d] = 1;
    attacFormalAXPBY(x->numDim, x->dim, a, x->ttRank, x->coreList, b, y->ttRank, y->coreList,
                     z->ttRank, z->coreList);
}

/* ---------------- Hadamard ---------------- */

void attacHadamard(const int d, const int *n, const int *rx, double *const *x, const int *ry,
                   double *const *y, int *rz, double **z) {
    for (int i = 0; i <= d; i++)
        if (rz[i] != rx[i] * ry[i])
            ptt_host_fatal("attacHadamard", "rz must equal rx*ry (caller-sized, as in attac.c)");
    for (int i = 0; i < d; i++)
        ptt_hip_hadamard_core(n[i], rx[i], rx[i + 1], x[i], ry[i], ry[i + 1], y[i], z[i]);
    ptt_hip_sync();
}

void pattacHadamard(const int d, const int *n, const int *rx, double *const *x, const int *ry,
                    double *const *y, int *rz, double **z) {
    attacHadamard(d, n, rx, x, ry, y, rz, z);
}

/* ---------------- scaling and mode operators ---------------- */

void attac_tt_scale(double a, attac_tt *x) {
    ptt_hip_scale(x->coreList[0], core_count3(x->ttRank[0], x->dim[0], x->ttRank[1]), a);
    ptt_hip_sync();
}

void attac_mult_diag_mat(attac_tt *f, int mode, double *vec) {
    if (mode < 0 || mode >= f->numDim) ptt_host_fatal("attac_mult_diag_mat", "bad mode");
    double *dv = ptt_host_coef(vec, f->dim[mode]);
    ptt_hip_diag_core(f->ttRank[mode], f->dim[mode], f->ttRank[mode + 1], dv, f->coreList[mode]);
    ptt_hip_sync();
}

/* Out-of-place operator into TMP, then the result becomes the core. */
static void stencil_apply(attac_tt *f, int mode, int kind, const double *sten, int slen) {
    if (mode < 0 || mode >= f->numDim) ptt_host_fatal("stencil", "bad mode");
    slen = (slen % 2) ? slen : slen - 1;
    if (slen < 1) ptt_host_fatal("stencil", "stencil length must be >= 1");
    int r0 = f->ttRank[mode], n = f->dim[mode], r1 = f->ttRank[mode + 1];
    if ((slen - 1) / 2 > n) ptt_host_fatal("stencil", "stencil wider than the mode");
    long cnt = core_count3(r0, n, r1);
    double *ds = ptt_host_coef(sten, slen);
    double *tmp = (double *)ptt_hip_scratch(PTT_WS_TMP, sizeof(double) * (size_t)cnt);
    ptt_hip_stencil_core(kind, r0, n, r1, ds, slen, f->coreList[mode], tmp);
    ptt_host_swap_core_with_tmp(f, mode, cnt);
    ptt_hip_sync();
}

void attac_apply_periodic_stencil(attac_tt *f, int mode, double *sten, int slen) {
    stencil_apply(f, mode, 1, sten, slen);
}

void pattac_apply_periodic_stencil(paratt *f, int mode, double *sten, int slen) {
    stencil_apply(f->x, mode, 0, sten, slen);
}

void pattac_apply_cdf_stencil(paratt *f, int mode, double *sten, int slen) {
    stencil_apply(f->x, mode, 2, sten, slen);
}

void pattac_apply_cdf_lf_flux(paratt *fp, int mode, double *wave_speed, double *gc_wave_speed,
                              double *source_coef, double dt, double dx) {
    attac_tt *f = fp->x;
    if (mode < 0 || mode >= f->numDim) ptt_host_fatal("pattac_apply_cdf_lf_flux", "bad mode");
    int r0 = f->ttRank[mode], n = f->dim[mode], r1 = f->ttRank[mode + 1];
    double *h = (double *)malloc(sizeof(double) * (size_t)(2 * n + 2));
    memcpy(h, wave_speed, sizeof(double) * n);
    h[n] = gc_wave_speed[0];
    h[n + 1] = gc_wave_speed[1];
    memcpy(h + n + 2, source_coef, sizeof(double) * n);
    double *dc = ptt_host_coef(h, 2 * n + 2);
    free(h);
    long cnt = core_count3(r0, n, r1);
    double *tmp = (double *)ptt_hip_scratch(PTT_WS_TMP, sizeof(double) * (size_t)cnt);
    ptt_hip_lf_flux_core(r0, n, r1, dc, dt, dx, f->coreList[mode], tmp);
    ptt_host_swap_core_with_tmp(f, mode, cnt);
    ptt_hip_sync();
}

void pattac_apply_cdf_lw_rhs(paratt *fp, int mode, double *wave_speed, double *gc_wave_speed,
                             double *bdry_wave_speed, double *source_coef, double dt, double dx) {
    attac_tt *f = fp->x;
    if (mode < 0 || mode >= f->numDim) ptt_host_fatal("pattac_apply_cdf_lw_rhs", "bad mode");
    int r0 = f->ttRank[mode], n = f->dim[mode], r1 = f->ttRank[mode + 1];
    double *h = (double *)malloc(sizeof(double) * (size_t)(3 * n + 3));
    memcpy(h, wave_speed, sizeof(double) * n);
    h[n] = gc_wave_speed[0];
    h[n + 1] = gc_wave_speed[1];
    memcpy(h + n + 2, bdry_wave_speed, sizeof(double) * (n + 1));
    memcpy(h + 2 * n + 3, source_coef, sizeof(double) * n);
    double *dc = ptt_host_coef(h, 3 * n + 3);
    free(h);
    long cnt = core_count3(r0, n, r1);
    double *tmp = (double *)ptt_hip_scratch(PTT_WS_TMP, sizeof(double) * (size_t)cnt);
    ptt_hip_lw_rhs_core(r0, n, r1, dc, dt, dx, f->coreList[mode], tmp);
    ptt_host_swap_core_with_tmp(f, mode, cnt);
    ptt_hip_sync();
}

/* ---------------- inner products ---------------- */

/* Left-to-right contraction a <- x_i^T (a y_i), scaled by s[i] after core i
 * (s = 1 reproduces pattacInnerProduct; s = 1/n_i the grid-normalised
 * product of paratt_dotprod without touching the cores). */
static double inner_sweep(int d, const int *n, const int *rx, double *const *x, const int *ry,
                          double *const *y, const double *s) {
    int rmax = 1, nmax = 1;
    for (int i = 0; i <= d; i++) {
        rmax = rx[i] > rmax ? rx[i] : rmax;
        rmax = ry[i] > rmax ? ry[i] : rmax;
    }
    for (int i = 0; i < d; i++) nmax = n[i] > nmax ? n[i] : nmax;
    double *a = (double *)ptt_hip_scratch(PTT_WS_SKETCH, sizeof(double) * (size_t)rmax * rmax);
    double *b = (double *)ptt_hip_scratch(PTT_WS_TMP2, sizeof(double) * (size_t)rmax * rmax * nmax);
    /* a = x0^T y0 (rx1 x ry1), K = rx0*n0 */
    ptt_hip_dgemm('T', 'N', rx[1], ry[1], rx[0] * n[0], s[0], x[0], (long)rx[0] * n[0], y[0],
                  (long)ry[0] * n[0], 0.0, a, rx[1]);
    for (int i = 1; i < d; i++) {
        long nry = (long)n[i] * ry[i + 1];
        /* b = a * y_i : (rx_i x ry_i) * (ry_i x n*ry_{i+1}) */
        ptt_hip_dgemm('N', 'N', rx[i], (int)nry, ry[i], 1.0, a, rx[i], y[i], ry[i], 0.0, b, rx[i]);
        /* a = x_i^T b : (rx_i*n x rx_{i+1})^T (rx_i*n x ry_{i+1}) */
        long rxn = (long)rx[i] * n[i];
        ptt_hip_dgemm('T', 'N', rx[i + 1], ry[i + 1], (int)rxn, s[i], x[i], rxn, b, rxn, 0.0, a,
                      rx[i + 1]);
    }
    double v = 0.0;
    ptt_hip_d2h(&v, a, sizeof(double));
    return v;
}

double pattacInnerProduct(const MPI_Comm comm, const int *isRedundant, const int d, const int *n,
                          const int *rx, double *const *x, const int *ry, double *const *y,
                          double *work) {
    (void)comm; (void)isRedundant; (void)work;
    double *s = (double *)malloc(sizeof(double) * d);
    for (int i = 0; i < d; i++) s[i] = 1.0;
    double v = inner_sweep(d, n, rx, x, ry, y, s);
    free(s);
    return v;
}

double paratt_dotprod(paratt *xp, paratt *yp) {
    int d = xp->numDim;
    double *s = (double *)malloc(sizeof(double) * d);
    /* the reference scales x and y cores by 1/sqrt(n_i) each (paratt.c:431-446) */
    for (int i = 0; i < d; i++) s[i] = 1.0 / ((double)xp->dim[i]);
    double v = inner_sweep(d, xp->x->dim, xp->x->ttRank, xp->x->coreList, yp->x->ttRank,
                           yp->x->coreList, s);
    free(s);
    return v;
}
