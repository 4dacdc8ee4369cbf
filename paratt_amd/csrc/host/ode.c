/*
 * ode.c -- step-truncation explicit Euler (paratt_ode.c:13-37) over the
 * GPU-resident primitives.  The vector field is the caller's host C code.
 */
#include "paratt.h"
#include "paratt_ode.h"
#include "ptt_host.h"

/* paratt_ode.c:26-37: f_new = T2(f_old + dt*T1(F(f_old))), caps = current ranks. */
void pattac_exp_euler_st(paratt *f_old, paratt *f_new, paratt *work,
                         void (*vecField)(paratt *, paratt *, int, ...), double dt, double eps_1,
                         double eps_2) {
    vecField(f_old, work, 0);
    paratt_truncate(work, eps_1, work->x->ttRank);
    attac_tt_axpby(1.0, f_old->x, dt, work->x, f_new->x);
    paratt_truncate(f_new, eps_2, f_new->x->ttRank);
}

/* paratt_ode.c:13-24, with the working GPU rounding in place of the
 * reference's serial truncation (whose SVD calls are commented out). */
void attac_exp_euler_st(attac_tt *f_old, attac_tt *f_new, attac_tt *work,
                        void (*vecField)(attac_tt *, attac_tt *, int, ...), double dt,
                        double eps_1, double eps_2) {
    vecField(f_old, work, 0);
    ptt_host_round(work, eps_1, work->ttRank);
    attac_tt_axpby(1.0, f_old, dt, work, f_new);
    ptt_host_round(f_new, eps_2, f_new->ttRank);
}

/* ------------------------------------------------------------------ */
/* Reaction-Burgers field and step-truncation RK4 (project composite;  */
/* the reference ships only explicit Euler).                           */
/* ------------------------------------------------------------------ */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "paratt_amd.h"
#include "ptt_hip.h"

static paratt *ptt_clone(paratt *s) {
    paratt *p = alloc_ptt(s->numDim, s->dim, s->x->ttRank);
    copy_rank_change_tt(p->x, s->x);
    return p;
}

static paratt *ptt_axpby_new(double a, paratt *x, double b, paratt *y) {
    int d = x->numDim;
    int *rz = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++)
        rz[k] = (k == 0 || k == d) ? 1 : x->x->ttRank[k] + y->x->ttRank[k];
    paratt *z = alloc_ptt(d, x->dim, rz);
    attac_tt_axpby(a, x->x, b, y->x, z->x);
    free(rz);
    return z;
}

static void round_rel_cap(paratt *x, double eps, int cap) {
    int d = x->numDim;
    int *mr = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) mr[k] = (k == 0 || k == d) ? 1 : cap;
    paratt_amd_truncate_rel(x, eps, mr);
    free(mr);
}

void paratt_amd_mode_sum(paratt *u, const double *sten3, paratt *s) {
    int d = u->numDim;
    int *rs = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) rs[k] = (k == 0 || k == d) ? 1 : 2 * u->x->ttRank[k];
    rank_change_tt(s->x, rs);
    double *dst = ptt_host_coef(sten3, 3);
    for (int k = 0; k < d; k++) {
        int r0 = u->x->ttRank[k], r1 = u->x->ttRank[k + 1], n = u->dim[k];
        int s0 = rs[k], s1 = rs[k + 1];
        long cnt = (long)r0 * n * r1;
        double *du = (double *)ptt_hip_scratch(PTT_WS_TMP2, sizeof(double) * (size_t)cnt);
        ptt_hip_stencil_core(0, r0, n, r1, dst, 3, u->x->coreList[k], du);
        double *sc = s->x->coreList[k];
        const double *uk = u->x->coreList[k];
        if (d == 1) {
            ptt_hip_d2d(sc, du, sizeof(double) * (size_t)cnt);
        } else if (k == 0) {
            ptt_hip_block_place(sc, s0, n, s1, 0, 0, du, r0, r1, 1.0);
            ptt_hip_block_place(sc, s0, n, s1, 0, r1, uk, r0, r1, 1.0);
        } else if (k == d - 1) {
            ptt_hip_block_place(sc, s0, n, s1, 0, 0, uk, r0, r1, 1.0);
            ptt_hip_block_place(sc, s0, n, s1, r0, 0, du, r0, r1, 1.0);
        } else {
            ptt_hip_zero(sc, sizeof(double) * (size_t)s0 * n * s1);
            ptt_hip_block_place(sc, s0, n, s1, 0, 0, uk, r0, r1, 1.0);
            ptt_hip_block_place(sc, s0, n, s1, r0, 0, du, r0, r1, 1.0);
            ptt_hip_block_place(sc, s0, n, s1, r0, r1, uk, r0, r1, 1.0);
        }
    }
    ptt_hip_sync();
    free(rs);
}

static struct {
    double nu, h, eps;
    int cap;
} g_rb = {0.01, 1.0, 1e-8, 64};

void paratt_amd_rxn_burgers_setup(double nu, double h, double eps, int cap) {
    g_rb.nu = nu;
    g_rb.h = h;
    g_rb.eps = eps;
    g_rb.cap = cap;
}

/* out = u .* T(1 - u - sum_i D1_i u) + nu * sum_i D2_i u  (vecField-compatible) */
void paratt_amd_rxn_burgers_field(paratt *u, paratt *out, int nargs, ...) {
    (void)nargs;
    int d = u->numDim;
    double h = g_rb.h;
    double d1[3] = {-0.5 / h, 0.0, 0.5 / h};
    double d2[3] = {1.0 / (h * h), -2.0 / (h * h), 1.0 / (h * h)};
    int *one = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) one[k] = 1;
    paratt *ones = alloc_ptt(d, u->dim, one);
    for (int k = 0; k < d; k++) ptt_hip_fill(ones->x->coreList[k], u->dim[k], 1.0);
    paratt *s1 = alloc_ptt(d, u->dim, one);
    paratt *lap = alloc_ptt(d, u->dim, one);
    paratt_amd_mode_sum(u, d1, s1);
    paratt_amd_mode_sum(u, d2, lap);
    paratt *v0 = ptt_axpby_new(1.0, ones, -1.0, u);
    paratt *v = ptt_axpby_new(1.0, v0, -1.0, s1);
    round_rel_cap(v, g_rb.eps, g_rb.cap);
    int *rh = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) rh[k] = u->x->ttRank[k] * v->x->ttRank[k];
    paratt *hd = alloc_ptt(d, u->dim, rh);
    pattacHadamard(d, u->dim, u->x->ttRank, u->x->coreList, v->x->ttRank, v->x->coreList,
                   hd->x->ttRank, hd->x->coreList);
    round_rel_cap(hd, g_rb.eps, g_rb.cap);
    paratt *f = ptt_axpby_new(1.0, hd, g_rb.nu, lap);
    round_rel_cap(f, g_rb.eps, g_rb.cap);
    copy_rank_change_tt(out->x, f->x);
    free_ptt(ones); free_ptt(s1); free_ptt(lap); free_ptt(v0); free_ptt(v); free_ptt(hd);
    free_ptt(f);
    free(one); free(rh);
}

/* Step-truncation RK4: stages and the final combination are rounded to eps
 * (relative) with the rank cap; vecField rounds its own output. */
void paratt_amd_rk4_st(paratt *f_old, paratt *f_new,
                       void (*vecField)(paratt *, paratt *, int, ...), double dt, double eps,
                       int cap) {
    int d = f_old->numDim;
    int *one = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) one[k] = 1;
    paratt *k1 = alloc_ptt(d, f_old->dim, one), *k2 = alloc_ptt(d, f_old->dim, one);
    paratt *k3 = alloc_ptt(d, f_old->dim, one), *k4 = alloc_ptt(d, f_old->dim, one);
    vecField(f_old, k1, 0);
    paratt *u2 = ptt_axpby_new(1.0, f_old, 0.5 * dt, k1);
    round_rel_cap(u2, eps, cap);
    vecField(u2, k2, 0);
    paratt *u3 = ptt_axpby_new(1.0, f_old, 0.5 * dt, k2);
    round_rel_cap(u3, eps, cap);
    vecField(u3, k3, 0);
    paratt *u4 = ptt_axpby_new(1.0, f_old, dt, k3);
    round_rel_cap(u4, eps, cap);
    vecField(u4, k4, 0);
    paratt *s12 = ptt_axpby_new(1.0, k1, 2.0, k2);
    paratt *s123 = ptt_axpby_new(1.0, s12, 2.0, k3);
    paratt *s = ptt_axpby_new(1.0, s123, 1.0, k4);
    round_rel_cap(s, eps, cap);
    paratt *un = ptt_axpby_new(1.0, f_old, dt / 6.0, s);
    round_rel_cap(un, eps, cap);
    copy_rank_change_tt(f_new->x, un->x);
    free_ptt(k1); free_ptt(k2); free_ptt(k3); free_ptt(k4); free_ptt(u2); free_ptt(u3);
    free_ptt(u4); free_ptt(s12); free_ptt(s123); free_ptt(s); free_ptt(un);
    free(one);
}
