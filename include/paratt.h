/*
 * paratt.h -- MI355X-native tensor-train library, drop-in for the
 * reference's source/paratt.h (declarations and struct layouts identical to
 * source/paratt.h:25-63 and :75-346, text of this project's own).
 *
 * What differs from the reference, on purpose:
 *  - Cores live in HBM.  `coreList[k]` holds a DEVICE address of
 *    r_k*n_k*r_{k+1} doubles (column-major, element (a,j,b) at
 *    a + r_k*(j + n_k*b), exactly the reference's layout).  Host code moves
 *    core data with paratt_amd_core_upload/_download (paratt_amd.h) instead
 *    of dereferencing coreList from the CPU.
 *  - One process drives one GPU; the reference's MPI split of every mode
 *    collapses to one rank (mpi_rk = 0, mpi_sz = 1, distrib_dim_mat = dim).
 *  - `wk`/`wksz` of attac_tt are unused (device workspace is pooled in the
 *    library); `orthog_wk` points at the opaque device workspace.
 *
 * Error behaviour follows the reference: functions return void and fatal
 * conditions print to stderr and exit(1).  A GPU runtime error is fatal.
 */
#ifndef PARATT_PARALLEL_TT_H
#define PARATT_PARALLEL_TT_H

#include <stdlib.h>
#include "pattac.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields, order and types as source/paratt.h:25-51. */
typedef struct {
    int numDim;        /* number of modes d */
    int *dim;          /* mode sizes n_k, d entries (host) */
    int *ttRank;       /* bond ranks r_0..r_d, r_0 = r_d = 1 (host) */
    double **coreList; /* d device pointers */
    int *coreMem;      /* capacity of each core in doubles (host) */
    double *wk;        /* unused on the GPU build (kept for layout) */
    int wksz;          /* unused on the GPU build (kept for layout) */
} attac_tt;

/* Same fields, order and types as source/paratt.h:53-63. */
typedef struct {
    attac_tt *x;
    int mpi_rk;            /* always 0 */
    int mpi_sz;            /* always 1 */
    int *distrib_dim_mat;  /* local mode sizes; equals dim at one rank */
    int *dim;              /* global mode sizes */
    int numDim;
    int *iwk;              /* host integer workspace (as the reference) */
    int iwksz;
    tsqr_tree *orthog_wk;  /* opaque device workspace */
} paratt;

/* Allocation (paratt.c:73-189).  Cores are device allocations; their
 * contents are undefined until written. */
attac_tt *alloc_tt(int ndims, int *d, int *ttR);
paratt *alloc_ptt(int ndims, int *d, int *ttR);
attac_tt *rank_change_tt(attac_tt *x_tt, int *ttR);
void copy_rank_change_tt(attac_tt *dest_tt, attac_tt *src_tt);
void free_tt(attac_tt *x_tt);
void free_ptt(paratt *x_ptt);
/* Cyclic ghost exchange with the rank above / below (paratt.c:192-234).  At
 * one rank the neighbour is the rank itself: recv_buff receives a copy of
 * send_buff (host or device addresses) and both requests are returned as
 * MPI_REQUEST_NULL.  Built-in MPICH datatypes only. */
void send_recv_below(void *send_buff, void *recv_buff, int msg_sz,
                     MPI_Datatype dat_t, MPI_Request *sendReq,
                     MPI_Request *recvReq);
void send_recv_above(void *send_buff, void *recv_buff, int msg_sz,
                     MPI_Datatype dat_t, MPI_Request *sendReq,
                     MPI_Request *recvReq);

/* z = a*x (+) b*y, resizing z (paratt.c:237-268). z must not alias x or y. */
void attac_tt_axpby(double a, attac_tt *x_tt, double b, attac_tt *y_tt,
                    attac_tt *z_tt);
/* Scale core 0 by a (paratt.c:271-281). */
void attac_tt_scale(double a, attac_tt *x_tt);
/* TT rounding of x in the plain Frobenius norm: left-to-right
 * orthogonalisation, then right-to-left truncation with the absolute
 * tolerance tol (tail of each bond's spectrum below tol/sqrt(d-1)) and the
 * rank caps maxRank (numDim+1 entries, may alias ttRank).  The intent of
 * paratt.c:283-313 (attacTruncate, attac.c:909-925); no grid scaling. */
void attac_tt_truncate(attac_tt *x_tt, double tol, int *maxRank);
/* ||x||_F; leaves x left-orthogonalised (paratt.c:345-366, attac.c:231-238). */
double attac_tt_norm(attac_tt *x_tt);
/* TT rounding in the grid-normalised norm (paratt.c:316-342).  maxRank has
 * numDim+1 entries and may alias x_ptt->x->ttRank. */
void paratt_truncate(paratt *x_ptt, double tol, int *maxRank);
/* ||x||_F / sqrt(prod n_k); leaves x left-orthogonalised (paratt.c:369-409). */
double paratt_norm(paratt *x_ptt);
/* <x, y> / prod n_k (paratt.c:412-469). */
double paratt_dotprod(paratt *x_ptt, paratt *y_ptt);

/* Rank-1 TT from host vectors vec[k][0..d[k]) (paratt.c:502-517). */
void attac_set_rank1(attac_tt **x_tt_ptr, double **vec, int ndims, int *d);
/* Sum of all_ttR rank-1 terms by repeated formal sums (paratt.c:520-568). */
attac_tt *attac_sum_of_product(double ***vec, int ndims, int *d, int all_ttR);

/* Core-wise operators along one mode (paratt.c:571-1250).  Coefficient
 * arrays (vec, sten, wave speeds, source_coef) are HOST arrays, as in the
 * reference; the operator is applied on the device. */
void attac_mult_diag_mat(attac_tt *f_tt, int mode, double *vec);
void attac_apply_periodic_stencil(attac_tt *f_tt, int mode, double *sten,
                                  int slen);
void pattac_apply_periodic_stencil(paratt *f_ptt, int mode, double *sten,
                                   int slen);
void pattac_apply_cdf_stencil(paratt *f_ptt, int mode, double *sten,
                              int slen);
void pattac_apply_cdf_lf_flux(paratt *f_ptt, int mode, double *wave_speed,
                              double *gc_wave_speed, double *source_coef,
                              double dt, double dx);
void pattac_apply_cdf_lw_rhs(paratt *f_ptt, int mode, double *wave_speed,
                             double *gc_wave_speed, double *bdry_wave_speed,
                             double *source_coef, double dt, double dx);

/* Entry f[mIdx[0], ..., mIdx[d-1]] (paratt.c:471-499).  Syncs; an index out
 * of range is fatal. */
double paratt_entry(paratt *x_ptt, int *mIdx);

/* Quadrature along one mode (paratt.c:1252-1618), in place on the device,
 * each fiber f[a, :, b] summed sequentially in the reference's order.  The
 * weight vectors are host arrays of the mode's length.
 *  apply_marginal: new TT (*marginal_ptt_ptr, which must be NULL on entry,
 *    else fatal) with `mode` contracted against row_vector;
 *  marginal_cusum: g[j] = sum_{j'<=j} row_vector[j'] f[j'];
 *  marginal_trap: trapezoid rule, g[0] = 0,
 *    g[j] = g[j-1] + 0.5 row_vector[j] (f[j-1] + f[j]);
 *  scale_weight_sum: the trapezoid rule on sum_w plus f[j] * scale[j]. */
void pattac_apply_marginal(paratt *f_ptt, paratt **marginal_ptt_ptr, int mode,
                           double *row_vector);
void pattac_marginal_cusum(paratt *f_ptt, int mode, double *row_vector);
void pattac_marginal_trap(paratt *f_ptt, int mode, double *row_vector);
void pattac_scale_weight_sum(paratt *f_ptt, int mode, double *scale,
                             double *sum_w);

/* TT folders (paratt.c:1621-1892): folder/core_<k+1>.cr holds core k as a
 * header line "r_k,n_k,r_{k+1}" and then one "%+1.16e" value per line in
 * column-major order.  Cores are downloaded / uploaded once each.
 *  write_tt_folder: creates the folder if missing; a file that cannot be
 *    written is fatal.
 *  load_tt_folder: allocates a TT of ndims modes from the headers; rank
 *    mismatches between files are reported on stderr and the files' sizes
 *    win; returns NULL (after a message) if a header cannot be read.
 *  read_core_size: header of one file; on error prints and leaves the
 *    outputs unchanged.
 *  load_core_slice: mode indices [N_beg, N_end) of one file into the DEVICE
 *    buffer *cr of capacity *cr_sz doubles (reallocated if short; sizes
 *    written to *rl, *N, *rr); out-of-range slices print and return. */
void pattac_write_tt_folder(char *folder_name, paratt *f_ptt);
paratt *pattac_load_tt_folder(char *folder_name, int ndims);
void read_core_size(char *fname, int *rl, int *N, int *rr);
void load_core_slice(char *fname, double **cr, int *cr_sz, int *rl, int *N,
                     int *rr, int N_beg, int N_end);

#ifdef __cplusplus
}
#endif
#endif
