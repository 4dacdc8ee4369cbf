/*
 * ptt_oracle_quad.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement, at one MPI rank, of the reference's quadrature along one
 * mode and entry evaluation:
 *   o_mode_scan      paratt.c:1386-1440 (cusum), 1444-1516 (trapezoid),
 *                    1518-1601 (trapezoid + scaled input);
 *   o_apply_marginal paratt.c:1252-1383;
 *   o_entry          paratt.c:471-499 with pattac.c:82-121.
 * Sums run in the reference's loop order with separate multiply and add, so
 * results are bit-identical to the reference compiled without FMA.  The
 * entry chain sums each row in column order, as reference (netlib) dgemv.
 * Pinned by golden vectors of the compiled reference and by brute-force
 * contractions of small full tensors (tests/test_oracle_quad.py).
 * Only tests/, smoke() and bench.py's CPU baseline may load this code.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* In place on core f (r0 x n x r1).  kind 0: cusum; 1: trapezoid;
 * 2: trapezoid on w plus f*scale. */
void o_mode_scan(int kind, int r0, int n, int r1, const double *w, const double *scale,
                 double *f) {
    for (int b = 0; b < r1; b++)
        for (int a = 0; a < r0; a++) {
            double *fb = f + a + (long)r0 * n * b;
#define F(j) fb[(long)r0 * (j)]
            if (kind == 0) {
                for (int j = 0; j < n; j++) F(j) *= w[j];
                for (int j = 1; j < n; j++) F(j) += F(j - 1);
            } else {
                double *orig = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
                for (int j = 0; j < n; j++) orig[j] = F(j);
                double cum = 0.0;
                for (int j = 1; j < n; j++) {
                    double delta = 0.5 * w[j] * (orig[j - 1] + orig[j]);
                    F(j - 1) = cum;
                    cum += delta;
                }
                if (n > 0) F(n - 1) = cum;
                if (kind == 2)
                    for (int j = 0; j < n; j++) F(j) += orig[j] * scale[j];
                free(orig);
            }
#undef F
        }
}

/* Marginal of mode `mode` against w.  out[] must hold d-1 host buffers of
 * the marginal's core sizes; out_r (d entries) receives its ranks. */
void o_apply_marginal(int d, const int *n, const int *r, double **x, int mode, const double *w,
                      double **out, int *out_r) {
    const int r0 = r[mode], r1 = r[mode + 1], nm = n[mode];
    double *wk = (double *)calloc((size_t)r0 * r1, sizeof(double));
    for (int k = 0; k < r1; k++)
        for (int i = 0; i < r0; i++) {
            wk[i + r0 * k] = 0.0;
            for (int j = 0; j < nm; j++) wk[i + r0 * k] += w[j] * x[mode][i + (long)r0 * (j + (long)nm * k)];
        }
    if (mode < d - 1) {
        for (int k = 0; k <= mode; k++) out_r[k] = r[k];
        for (int k = mode + 1; k < d; k++) out_r[k] = r[k + 1];
        long cols = (long)n[mode + 1] * r[mode + 2];
        for (long j = 0; j < cols; j++)
            for (int i = 0; i < r0; i++) {
                double s = 0.0;
                for (int k = 0; k < r1; k++) s += wk[i + r0 * k] * x[mode + 1][k + (long)r1 * j];
                out[mode][i + (long)r0 * j] = s;
            }
        for (int k = 0; k < mode; k++)
            memcpy(out[k], x[k], sizeof(double) * (size_t)r[k] * n[k] * r[k + 1]);
        for (int k = mode + 1; k < d - 1; k++)
            memcpy(out[k], x[k + 1], sizeof(double) * (size_t)r[k + 1] * n[k + 1] * r[k + 2]);
    } else {
        for (int k = 0; k < d - 1; k++) out_r[k] = r[k];
        out_r[d - 1] = 1;
        long rows = (long)r[mode - 1] * n[mode - 1];
        for (long i = 0; i < rows; i++) {
            double s = 0.0;
            for (int j = 0; j < r0; j++) s += wk[j] * x[mode - 1][i + rows * j];
            out[mode - 1][i] = s;
        }
        for (int k = 0; k < mode - 1; k++)
            memcpy(out[k], x[k], sizeof(double) * (size_t)r[k] * n[k] * r[k + 1]);
    }
    free(wk);
}

/* Entry by the right-to-left chain: v = X_{d-1}[:, i_{d-1}, 0], then
 * v <- X_k[:, i_k, :] v for k = d-2 .. 0. */
double o_entry(int d, const int *n, const int *r, double **x, const int *idx) {
    int rmax = 1;
    for (int k = 0; k <= d; k++) rmax = r[k] > rmax ? r[k] : rmax;
    double *a = (double *)calloc((size_t)rmax, sizeof(double));
    double *b = (double *)calloc((size_t)rmax, sizeof(double));
    memcpy(a, x[d - 1] + (long)idx[d - 1] * r[d - 1], sizeof(double) * (size_t)r[d - 1]);
    for (int k = d - 2; k >= 0; k--) {
        long ld = (long)r[k] * n[k];
        const double *s = x[k] + (long)idx[k] * r[k];
        for (int p = 0; p < r[k]; p++) b[p] = 0.0;
        for (int q = 0; q < r[k + 1]; q++)
            for (int p = 0; p < r[k]; p++) b[p] += a[q] * s[p + ld * q];
        double *t = a;
        a = b;
        b = t;
    }
    double v = a[0];
    free(a);
    free(b);
    return v;
}
