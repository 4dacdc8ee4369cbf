/*
 * tt.c -- TT storage on the device: allocation, rank changes and copies
 * (paratt.c:73-189), rank-1 and sum-of-rank-1 builders (paratt.c:502-568),
 * host <-> HBM transfers.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "paratt.h"
#include "paratt_amd.h"
#include "ptt_hip.h"
#include "ptt_host.h"

void ptt_host_fatal(const char *fn, const char *msg) {
    fprintf(stderr, "%s: %s\n", fn, msg);
    exit(1);
}

static void *xmalloc(size_t b) {
    void *p = malloc(b ? b : 8);
    if (!p) ptt_host_fatal("paratt_amd", "host allocation failed");
    return p;
}

void *ptt_host_malloc(size_t bytes) { return xmalloc(bytes); }

void *ptt_host_calloc(size_t count, size_t size) {
    void *p = calloc(count ? count : 1, size ? size : 1);
    if (!p) ptt_host_fatal("calloc", "host allocation failed");
    return p;
}

void *ptt_host_realloc(void *q, size_t bytes) {
    void *p = realloc(q, bytes ? bytes : 1);
    if (!p) ptt_host_fatal("realloc", "host allocation failed");
    return p;
}

static long core_count(const int *dim, const int *r, int k) {
    return (long)r[k] * dim[k] * r[k + 1];
}

attac_tt *alloc_tt(int ndims, int *d, int *ttR) {
    if (ndims < 1) ptt_host_fatal("alloc_tt", "ndims must be >= 1");
    attac_tt *tt = (attac_tt *)xmalloc(sizeof(attac_tt));
    tt->numDim = ndims;
    tt->dim = (int *)xmalloc(sizeof(int) * ndims);
    tt->ttRank = (int *)xmalloc(sizeof(int) * (ndims + 1));
    tt->coreList = (double **)xmalloc(sizeof(double *) * ndims);
    tt->coreMem = (int *)xmalloc(sizeof(int) * ndims);
    for (int k = 0; k < ndims; k++) tt->dim[k] = d[k];
    for (int k = 0; k <= ndims; k++) tt->ttRank[k] = ttR[k];
    for (int k = 0; k < ndims; k++) {
        long cnt = core_count(d, ttR, k);
        tt->coreList[k] = (double *)ptt_hip_malloc(sizeof(double) * (size_t)(cnt > 0 ? cnt : 1));
        tt->coreMem[k] = (int)cnt;
    }
    tt->wk = NULL;
    tt->wksz = 0;
    return tt;
}

paratt *alloc_ptt(int ndims, int *d, int *ttR) {
    paratt *p = (paratt *)xmalloc(sizeof(paratt));
    p->mpi_rk = 0;
    p->mpi_sz = 1;
    p->numDim = ndims;
    p->dim = (int *)xmalloc(sizeof(int) * ndims);
    p->distrib_dim_mat = (int *)xmalloc(sizeof(int) * ndims);
    for (int k = 0; k < ndims; k++) p->dim[k] = p->distrib_dim_mat[k] = d[k];
    p->x = alloc_tt(ndims, d, ttR);
    p->iwksz = ndims + 1;
    p->iwk = (int *)xmalloc(sizeof(int) * p->iwksz);
    p->orthog_wk = NULL;
    return p;
}

attac_tt *rank_change_tt(attac_tt *x, int *ttR) {
    for (int k = 0; k < x->numDim; k++) {
        long need = (long)x->dim[k] * ttR[k] * ttR[k + 1];
        if (x->coreMem[k] < need) {
            /* realloc semantics: keep the old contents */
            double *nb = (double *)ptt_hip_malloc(sizeof(double) * (size_t)need);
            ptt_hip_d2d(nb, x->coreList[k], sizeof(double) * (size_t)x->coreMem[k]);
            ptt_hip_free(x->coreList[k]);
            x->coreList[k] = nb;
            x->coreMem[k] = (int)need;
        }
        x->ttRank[k] = ttR[k];
    }
    x->ttRank[x->numDim] = ttR[x->numDim];
    return x;
}

void copy_rank_change_tt(attac_tt *dest, attac_tt *src) {
    rank_change_tt(dest, src->ttRank);
    for (int k = 0; k < dest->numDim; k++)
        ptt_hip_d2d(dest->coreList[k], src->coreList[k],
                    sizeof(double) * (size_t)core_count(src->dim, src->ttRank, k));
    ptt_hip_sync();
}

void free_tt(attac_tt *x) {
    if (!x) return;
    for (int k = 0; k < x->numDim; k++) ptt_hip_free(x->coreList[k]);
    free(x->dim);
    free(x->ttRank);
    free(x->coreList);
    free(x->coreMem);
    if (x->wk) free(x->wk);
    free(x);
}

void free_ptt(paratt *p) {
    if (!p) return;
    free_tt(p->x);
    free(p->distrib_dim_mat);
    if (p->iwk) free(p->iwk);
    free(p->dim);
    free(p);
}

void ptt_host_reserve_core(attac_tt *x, int k, long count) {
    if (x->coreMem == NULL) {
        ptt_host_fatal("paratt_amd", "borrowed core too small");
    }
    if (x->coreMem[k] < count) {
        ptt_hip_free(x->coreList[k]);
        x->coreList[k] = (double *)ptt_hip_malloc(sizeof(double) * (size_t)count);
        x->coreMem[k] = (int)count;
    }
}

void ptt_host_swap_core_with_tmp(attac_tt *x, int k, long count) {
    if (x->coreMem == NULL) {
        /* borrowed view: copy the result back into the caller's buffer
         * (rounding never outgrows the input core) */
        double *tmp = (double *)ptt_hip_scratch(PTT_WS_TMP, 8);
        ptt_hip_d2d(x->coreList[k], tmp, sizeof(double) * (size_t)count);
        return;
    }
    size_t cap = (size_t)x->coreMem[k] * sizeof(double);
    ptt_hip_swap_tmp(PTT_WS_TMP, &x->coreList[k], &cap);
    size_t el = cap / sizeof(double);
    x->coreMem[k] = (int)(el > 2147483647UL ? 2147483647UL : el);
}

void ptt_host_set_diag(double *A, int m, double alpha) { ptt_hip_set_diag(A, m, m, alpha); }

void ptt_host_view(attac_tt *view, int d, int *n, int *r, double **x) {
    view->numDim = d;
    view->dim = n;
    view->ttRank = r;
    view->coreList = x;
    view->coreMem = NULL; /* borrowed */
    view->wk = NULL;
    view->wksz = 0;
}

void ptt_host_view_release(attac_tt *view, double **x) { (void)view; (void)x; }

double *ptt_host_coef(const double *host, long count) {
    double *d = (double *)ptt_hip_scratch(PTT_WS_COEF, sizeof(double) * (size_t)(count > 0 ? count : 1));
    ptt_hip_h2d(d, host, sizeof(double) * (size_t)count);
    return d;
}

void paratt_amd_core_upload(attac_tt *x, int k, const double *host) {
    ptt_hip_h2d(x->coreList[k], host, sizeof(double) * (size_t)core_count(x->dim, x->ttRank, k));
}

void paratt_amd_core_download(const attac_tt *x, int k, double *host) {
    ptt_hip_d2h(host, x->coreList[k], sizeof(double) * (size_t)core_count(x->dim, x->ttRank, k));
}

void paratt_amd_fill_gaussian(attac_tt *x, uint64_t seed, double scale) {
    for (int k = 0; k < x->numDim; k++)
        ptt_hip_gauss_fill(x->coreList[k], core_count(x->dim, x->ttRank, k), seed, k, scale);
    ptt_hip_sync();
}

void paratt_amd_sync(void) { ptt_hip_sync(); }

void attac_set_rank1(attac_tt **x_ptr, double **vec, int ndims, int *d) {
    int *one = (int *)xmalloc(sizeof(int) * (ndims + 1));
    for (int k = 0; k <= ndims; k++) one[k] = 1;
    *x_ptr = alloc_tt(ndims, d, one);
    for (int k = 0; k < ndims; k++) ptt_hip_h2d((*x_ptr)->coreList[k], vec[k], sizeof(double) * d[k]);
    free(one);
}

attac_tt *attac_sum_of_product(double ***vec, int ndims, int *d, int all_ttR) {
    /* Same construction as paratt.c:520-568: repeated formal sums of the
     * rank-1 terms, so the cores come out in the reference's block order. */
    attac_tt *acc = NULL;
    attac_set_rank1(&acc, vec[0], ndims, d);
    for (int t = 1; t < all_ttR; t++) {
        attac_tt *term = NULL;
        attac_set_rank1(&term, vec[t], ndims, d);
        int *rz = (int *)xmalloc(sizeof(int) * (ndims + 1));
        for (int k = 0; k <= ndims; k++) rz[k] = (k == 0 || k == ndims) ? 1 : acc->ttRank[k] + 1;
        attac_tt *z = alloc_tt(ndims, d, rz);
        attacFormalSum(ndims, d, acc->ttRank, acc->coreList, term->ttRank, term->coreList,
                       z->ttRank, z->coreList);
        free_tt(acc);
        free_tt(term);
        free(rz);
        acc = z;
    }
    return acc;
}
