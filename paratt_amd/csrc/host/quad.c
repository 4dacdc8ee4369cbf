/*
 * quad.c -- quadrature along one mode and entry evaluation of device-resident
 * TTs: pattac_apply_marginal, pattac_marginal_cusum, pattac_marginal_trap,
 * pattac_scale_weight_sum (paratt.c:1252-1618), paratt_entry
 * (paratt.c:471-499) with pattacValue (pattac.c:82-121), and the one-rank
 * ghost exchange helpers send_recv_above / send_recv_below
 * (paratt.c:192-234).
 *
 * At one rank the reference's pipelined exchange of the running sum between
 * neighbouring ranks (paratt.c:1420-1440, 1465-1516) has no partner, so each
 * operator is one sequential pass per fiber on the device (quad.hip), summed
 * in the reference's order.  Weight vectors are host arrays, as in the
 * reference; they are uploaded per call.  Indices and modes are checked on
 * the host before any launch: an out-of-range mode or index is fatal.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "paratt.h"
#include "ptt_hip.h"
#include "ptt_host.h"

static void check_mode(const char *fn, const paratt *f, int mode) {
    if (!f || !f->x) ptt_host_fatal(fn, "NULL tensor");
    if (mode < 0 || mode >= f->numDim) {
        fprintf(stderr, "%s: mode %d out of range [0, %d)\n", fn, mode, f->numDim);
        exit(1);
    }
}

static void mode_scan(const char *fn, paratt *f, int mode, int kind, const double *w,
                      const double *scale) {
    check_mode(fn, f, mode);
    attac_tt *x = f->x;
    int n = x->dim[mode];
    if (n <= 0) return;
    double *host = (double *)ptt_host_malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(host, w, sizeof(double) * (size_t)n);
    if (scale) memcpy(host + n, scale, sizeof(double) * (size_t)n);
    double *dc = ptt_host_coef(host, scale ? 2L * n : (long)n);
    free(host);
    ptt_hip_mode_scan(kind, x->ttRank[mode], n, x->ttRank[mode + 1], dc, scale ? dc + n : NULL,
                      x->coreList[mode]);
    ptt_hip_sync();
}

void pattac_marginal_cusum(paratt *f_ptt, int mode, double *row_vector) {
    mode_scan("pattac_marginal_cusum", f_ptt, mode, 0, row_vector, NULL);
}

void pattac_marginal_trap(paratt *f_ptt, int mode, double *row_vector) {
    mode_scan("pattac_marginal_trap", f_ptt, mode, 1, row_vector, NULL);
}

void pattac_scale_weight_sum(paratt *f_ptt, int mode, double *scale, double *sum_w) {
    mode_scan("pattac_scale_weight_sum", f_ptt, mode, 2, sum_w, scale);
}

void pattac_apply_marginal(paratt *f_ptt, paratt **marginal_ptt_ptr, int mode,
                           double *row_vector) {
    if (*marginal_ptt_ptr != NULL) {
        fprintf(stderr, "pattac_apply_marginal: Error.\nInput marginal_tt_ptr was not NULL.\n");
        exit(1);
    }
    check_mode("pattac_apply_marginal", f_ptt, mode);
    attac_tt *x = f_ptt->x;
    const int d = x->numDim;
    if (d < 2) ptt_host_fatal("pattac_apply_marginal", "needs at least two modes");
    const int *r = x->ttRank, *nn = x->dim;
    /* Sizes of the marginal: mode removed; below the last mode the bond
     * r[mode+1] disappears, at the last mode the new right bond is 1. */
    int *md = (int *)ptt_host_malloc(sizeof(int) * (size_t)(d - 1));
    int *mr = (int *)ptt_host_malloc(sizeof(int) * (size_t)d);
    for (int k = 0, t = 0; k < d; k++)
        if (k != mode) md[t++] = nn[k];
    if (mode < d - 1) {
        for (int k = 0; k <= mode; k++) mr[k] = r[k];
        for (int k = mode + 1; k < d; k++) mr[k] = r[k + 1];
    } else {
        for (int k = 0; k < d - 1; k++) mr[k] = r[k];
        mr[d - 1] = 1;
    }
    paratt *m = alloc_ptt(d - 1, md, mr);
    *marginal_ptt_ptr = m;
    const int n = nn[mode], r0 = r[mode], r1 = r[mode + 1];
    double *w = ptt_host_coef(row_vector, n);
    /* wk (r0 x r1) = sum_j w[j] X_mode[:, j, :] */
    double *wk = (double *)ptt_hip_scratch(PTT_WS_TMP2, sizeof(double) * (size_t)r0 * r1);
    ptt_hip_seq_gemm(r0, 1, n, r1, x->coreList[mode], r0, (long)r0 * n, w, n, 0, wk, r0, r0);
    if (mode < d - 1) {
        /* marginal core `mode` = wk * X_{mode+1}: r0 x (n_{mode+1} r_{mode+2}) */
        long cols = (long)nn[mode + 1] * r[mode + 2];
        ptt_hip_seq_gemm(r0, (int)cols, r1, 1, wk, r0, 0, x->coreList[mode + 1], r1, 0,
                         m->x->coreList[mode], r0, 0);
        for (int k = 0; k < mode; k++)
            ptt_hip_d2d(m->x->coreList[k], x->coreList[k],
                        sizeof(double) * (size_t)r[k] * nn[k] * r[k + 1]);
        for (int k = mode + 1; k < d - 1; k++)
            ptt_hip_d2d(m->x->coreList[k], x->coreList[k + 1],
                        sizeof(double) * (size_t)r[k + 1] * nn[k + 1] * r[k + 2]);
    } else {
        /* last mode: marginal core d-2 = X_{d-2} (r_{d-2} n_{d-2} x r_{d-1}) * wk */
        long rows = (long)r[mode - 1] * nn[mode - 1];
        ptt_hip_seq_gemm((int)rows, 1, r0, 1, x->coreList[mode - 1], rows, 0, wk, r0, 0,
                         m->x->coreList[mode - 1], rows, 0);
        for (int k = 0; k < mode - 1; k++)
            ptt_hip_d2d(m->x->coreList[k], x->coreList[k],
                        sizeof(double) * (size_t)r[k] * nn[k] * r[k + 1]);
    }
    ptt_hip_sync();
    free(md);
    free(mr);
}

static double tt_value(const char *fn, int d, const int *n, const int *r, double *const *cores,
                       const int *idx) {
    if (d < 1) ptt_host_fatal(fn, "needs at least one mode");
    int rmax = 1;
    for (int k = 0; k < d; k++) {
        if (idx[k] < 0 || idx[k] >= n[k]) {
            fprintf(stderr, "%s: index %d of mode %d out of range [0, %d)\n", fn, idx[k], k,
                    n[k]);
            exit(1);
        }
        if (r[k] < 1) ptt_host_fatal(fn, "ranks must be >= 1");
        rmax = r[k] > rmax ? r[k] : rmax;
    }
    rmax = r[d] > rmax ? r[d] : rmax;
    double *vec = (double *)ptt_hip_scratch(PTT_WS_SMALL2, sizeof(double) * (size_t)rmax);
    double v = 0.0;
    ptt_hip_tt_entry(d, n, r, cores, idx, vec, &v);
    return v;
}

double paratt_entry(paratt *x_ptt, int *mIdx) {
    attac_tt *x = x_ptt->x;
    return tt_value("paratt_entry", x->numDim, x->dim, x->ttRank, x->coreList, mIdx);
}

double pattacValue(const MPI_Comm comm, const int *index, const int *procHolder, const int d,
                   const int *n, const int *rx, double *const *x, double *work) {
    (void)comm; (void)procHolder; (void)work;
    return tt_value("pattacValue", d, n, rx, x, index);
}

/* Size in bytes of an MPICH built-in datatype handle (the size is encoded in
 * bits 8..15 of built-in handles, whose kind bits 30..31 are 01). */
static size_t builtin_type_size(MPI_Datatype t) {
    uint32_t h = (uint32_t)t;
    if ((h >> 30) != 1u) ptt_host_fatal("send_recv", "only built-in MPI datatypes are supported");
    return (size_t)((h >> 8) & 0xffu);
}

/* At one rank the cyclic neighbour above and below is the rank itself: the
 * message arrives at once and both requests are already complete. */
static void self_exchange(void *send_buff, void *recv_buff, int msg_sz, MPI_Datatype dat_t,
                          MPI_Request *sendReq, MPI_Request *recvReq) {
    if (msg_sz > 0) ptt_hip_copy(recv_buff, send_buff, (size_t)msg_sz * builtin_type_size(dat_t));
    if (sendReq) *sendReq = MPI_REQUEST_NULL;
    if (recvReq) *recvReq = MPI_REQUEST_NULL;
}

void send_recv_below(void *send_buff, void *recv_buff, int msg_sz, MPI_Datatype dat_t,
                     MPI_Request *sendReq, MPI_Request *recvReq) {
    self_exchange(send_buff, recv_buff, msg_sz, dat_t, sendReq, recvReq);
}

void send_recv_above(void *send_buff, void *recv_buff, int msg_sz, MPI_Datatype dat_t,
                     MPI_Request *sendReq, MPI_Request *recvReq) {
    self_exchange(send_buff, recv_buff, msg_sz, dat_t, sendReq, recvReq);
}
