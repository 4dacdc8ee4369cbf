/*
 * ptt_oracle_ode.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the step-truncation ODE composites, built from the
 * oracle's restatements of the reference primitives (ptt_oracle.c):
 *  - o_mode_sum: sum_i D_i u as an exact rank-2r TT (equal to the formal sum
 *    of d stenciled copies, paratt.c:665-770 + attac.c:81-112);
 *  - o_rxn_burgers_rhs / o_rxn_burgers_rk4: the reaction-Burgers composite
 *    u_t = -u (sum_i d_i u) + nu lap(u) + u(1-u) on a periodic grid.  The
 *    reference has no RK4 and no such vector field; this composition is the
 *    project's definition and the GPU library (paratt_amd/csrc/host/ode.c)
 *    composes the same steps in the same order.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* primitives from ptt_oracle.c */
void o_axpby(int d, const int *n, double alpha, const int *rx, double **x, double beta,
             const int *ry, double **y, int *rz, double **z);
void o_hadamard(int d, const int *n, const int *rx, double **x, const int *ry, double **y,
                double **z);
void o_truncate(int d, const int *n, int *r, double **x, const int *maxr, double abs_tol,
                double *sig_log, int ldsig, int *numrank_log);
double o_norm(int d, const int *n, int *r, double **x);
void o_periodic_stencil(int r, int nn, int r1, const double *sten, int slen, double *f);

typedef struct {
    int d;
    int *n;
    int *r;
    double **c;
} ott;

static long csz(const ott *t, int k) { return (long)t->r[k] * t->n[k] * t->r[k + 1]; }

static ott *ott_alloc(int d, const int *n, const int *r) {
    ott *t = (ott *)malloc(sizeof(ott));
    t->d = d;
    t->n = (int *)malloc(sizeof(int) * d);
    t->r = (int *)malloc(sizeof(int) * (d + 1));
    t->c = (double **)malloc(sizeof(double *) * d);
    memcpy(t->n, n, sizeof(int) * d);
    memcpy(t->r, r, sizeof(int) * (d + 1));
    for (int k = 0; k < d; k++) t->c[k] = (double *)calloc((size_t)(csz(t, k) > 0 ? csz(t, k) : 1), sizeof(double));
    return t;
}

static void ott_free(ott *t) {
    if (!t) return;
    for (int k = 0; k < t->d; k++) free(t->c[k]);
    free(t->c); free(t->n); free(t->r); free(t);
}

static ott *ott_copy(const ott *s) {
    ott *t = ott_alloc(s->d, s->n, s->r);
    for (int k = 0; k < s->d; k++) memcpy(t->c[k], s->c[k], sizeof(double) * (size_t)csz(s, k));
    return t;
}

static ott *ott_axpby(double a, const ott *x, double b, const ott *y) {
    int d = x->d;
    int *rz = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) rz[k] = (k == 0 || k == d) ? 1 : x->r[k] + y->r[k];
    ott *z = ott_alloc(d, x->n, rz);
    o_axpby(d, x->n, a, x->r, x->c, b, y->r, y->c, z->r, z->c);
    free(rz);
    return z;
}

/* paratt_truncate with a relative tolerance eps * ||t||_n and a uniform cap */
static void ott_round_rel(ott *t, double eps, int cap) {
    int d = t->d;
    ott *c = ott_copy(t);
    double nrm = o_norm(d, c->n, c->r, c->c);
    ott_free(c);
    int *mr = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) mr[k] = (k == 0 || k == d) ? 1 : cap;
    o_truncate(d, t->n, t->r, t->c, mr, eps * nrm, NULL, 0, NULL);
    free(mr);
}

/* S = sum_i D_i u as a rank-2r TT (formal, exact):
 *   core 0      : [D u_0 | u_0]
 *   core k      : [[u_k, 0], [D u_k, u_k]]  (rows: done / pending)
 *   core d-1    : [u_{d-1}; D u_{d-1}]
 * D is the one-rank periodic stencil of paratt.c:665-770 (slen 3). */
static ott *mode_sum(const ott *u, const double *sten) {
    int d = u->d;
    int *rs = (int *)calloc((size_t)d + 1, sizeof(int));
    for (int k = 0; k <= d; k++) rs[k] = (k == 0 || k == d) ? 1 : 2 * u->r[k];
    ott *s = ott_alloc(d, u->n, rs);
    for (int k = 0; k < d; k++) {
        int r0 = u->r[k], r1 = u->r[k + 1], nn = u->n[k];
        int s0 = rs[k], s1 = rs[k + 1];
        double *du = (double *)malloc(sizeof(double) * (size_t)csz(u, k));
        memcpy(du, u->c[k], sizeof(double) * (size_t)csz(u, k));
        o_periodic_stencil(r0, nn, r1, sten, 3, du);
#define PUT(ra, rb, src)                                                            \
    for (int b = 0; b < r1; b++)                                                    \
        for (int j = 0; j < nn; j++)                                                \
            for (int a = 0; a < r0; a++)                                            \
                s->c[k][(ra) + a + (long)s0 * (j + (long)nn * ((rb) + b))] =        \
                    (src)[a + (long)r0 * (j + (long)nn * b)];
        if (d == 1) {
            memcpy(s->c[0], du, sizeof(double) * (size_t)csz(u, 0));
        } else if (k == 0) {
            PUT(0, 0, du);
            PUT(0, r1, u->c[k]);
        } else if (k == d - 1) {
            PUT(0, 0, u->c[k]);
            PUT(r0, 0, du);
        } else {
            PUT(0, 0, u->c[k]);
            PUT(r0, 0, du);
            PUT(r0, r1, u->c[k]);
        }
#undef PUT
        (void)s1;
        free(du);
    }
    free(rs);
    return s;
}

typedef struct {
    double nu, h, eps;
    int cap;
} rb_params;

/* F(u) = u .* T(1 - u - sum_i D1_i u) + nu * sum_i D2_i u, every sum
 * truncated to eps (relative) with the rank cap. */
static ott *rb_rhs(const ott *u, const rb_params *p) {
    int d = u->d;
    double d1[3] = {-0.5 / p->h, 0.0, 0.5 / p->h};
    double d2[3] = {1.0 / (p->h * p->h), -2.0 / (p->h * p->h), 1.0 / (p->h * p->h)};
    int *one = (int *)calloc((size_t)d + 1, sizeof(int));
    for (int k = 0; k <= d; k++) one[k] = 1;
    ott *ones = ott_alloc(d, u->n, one);
    for (int k = 0; k < d; k++)
        for (int j = 0; j < u->n[k]; j++) ones->c[k][j] = 1.0;
    ott *s1 = mode_sum(u, d1);
    ott *lap = mode_sum(u, d2);
    ott *v0 = ott_axpby(1.0, ones, -1.0, u);
    ott *v = ott_axpby(1.0, v0, -1.0, s1);
    ott_round_rel(v, p->eps, p->cap);
    int *rh = (int *)malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) rh[k] = u->r[k] * v->r[k];
    ott *h = ott_alloc(d, u->n, rh);
    o_hadamard(d, u->n, u->r, u->c, v->r, v->c, h->c);
    ott_round_rel(h, p->eps, p->cap);
    ott *f = ott_axpby(1.0, h, p->nu, lap);
    ott_round_rel(f, p->eps, p->cap);
    ott_free(ones); ott_free(s1); ott_free(lap); ott_free(v0); ott_free(v); ott_free(h);
    free(one); free(rh);
    return f;
}

/* One step-truncation RK4 step of the reaction-Burgers field. */
static ott *rb_rk4(const ott *u, double dt, const rb_params *p) {
    ott *k1 = rb_rhs(u, p);
    ott *u2 = ott_axpby(1.0, u, 0.5 * dt, k1);
    ott_round_rel(u2, p->eps, p->cap);
    ott *k2 = rb_rhs(u2, p);
    ott *u3 = ott_axpby(1.0, u, 0.5 * dt, k2);
    ott_round_rel(u3, p->eps, p->cap);
    ott *k3 = rb_rhs(u3, p);
    ott *u4 = ott_axpby(1.0, u, dt, k3);
    ott_round_rel(u4, p->eps, p->cap);
    ott *k4 = rb_rhs(u4, p);
    ott *s12 = ott_axpby(1.0, k1, 2.0, k2);
    ott *s123 = ott_axpby(1.0, s12, 2.0, k3);
    ott *s = ott_axpby(1.0, s123, 1.0, k4);
    ott_round_rel(s, p->eps, p->cap);
    ott *un = ott_axpby(1.0, u, dt / 6.0, s);
    ott_round_rel(un, p->eps, p->cap);
    ott_free(k1); ott_free(u2); ott_free(k2); ott_free(u3); ott_free(k3); ott_free(u4);
    ott_free(k4); ott_free(s12); ott_free(s123); ott_free(s);
    return un;
}

/* ctypes entry: runs `steps` RK4 steps from u (d, n, r, cores); returns the
 * result by filling r_out (d+1) and cores_out (caller-sized to cap*n*cap). */
void o_rxn_burgers_rk4(int d, const int *n, const int *r, double **cores, double nu, double h,
                       double dt, double eps, int cap, int steps, int *r_out, double **cores_out) {
    ott *u = ott_alloc(d, n, r);
    for (int k = 0; k < d; k++) memcpy(u->c[k], cores[k], sizeof(double) * (size_t)csz(u, k));
    rb_params p = {nu, h, eps, cap};
    for (int s = 0; s < steps; s++) {
        ott *un = rb_rk4(u, dt, &p);
        ott_free(u);
        u = un;
    }
    memcpy(r_out, u->r, sizeof(int) * (d + 1));
    for (int k = 0; k < d; k++) memcpy(cores_out[k], u->c[k], sizeof(double) * (size_t)csz(u, k));
    ott_free(u);
}

/* ctypes entry: the reaction-Burgers right-hand side alone. */
void o_rxn_burgers_rhs(int d, const int *n, const int *r, double **cores, double nu, double h,
                       double eps, int cap, int *r_out, double **cores_out) {
    ott *u = ott_alloc(d, n, r);
    for (int k = 0; k < d; k++) memcpy(u->c[k], cores[k], sizeof(double) * (size_t)csz(u, k));
    rb_params p = {nu, h, eps, cap};
    ott *f = rb_rhs(u, &p);
    memcpy(r_out, f->r, sizeof(int) * (d + 1));
    for (int k = 0; k < d; k++) memcpy(cores_out[k], f->c[k], sizeof(double) * (size_t)csz(f, k));
    ott_free(u);
    ott_free(f);
}

/* ctypes entry: S = sum_i D_i u (rank 2r) with a 3-point stencil. */
void o_mode_sum(int d, const int *n, const int *r, double **cores, const double *sten,
                int *r_out, double **cores_out) {
    ott *u = ott_alloc(d, n, r);
    for (int k = 0; k < d; k++) memcpy(u->c[k], cores[k], sizeof(double) * (size_t)csz(u, k));
    ott *s = mode_sum(u, sten);
    memcpy(r_out, s->r, sizeof(int) * (d + 1));
    for (int k = 0; k < d; k++) memcpy(cores_out[k], s->c[k], sizeof(double) * (size_t)csz(s, k));
    ott_free(u);
    ott_free(s);
}
