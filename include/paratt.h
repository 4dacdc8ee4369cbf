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

/* z = a*x (+) b*y, resizing z (paratt.c:237-268). z must not alias x or y. */
void attac_tt_axpby(double a, attac_tt *x_tt, double b, attac_tt *y_tt,
                    attac_tt *z_tt);
/* Scale core 0 by a (paratt.c:271-281). */
void attac_tt_scale(double a, attac_tt *x_tt);
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

#ifdef __cplusplus
}
#endif
#endif
