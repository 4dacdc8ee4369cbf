/*
 * paratt_amd.h -- additions of the MI355X build that have no counterpart in
 * the reference: moving core data between host and HBM (the reference's
 * cores are host memory that user code dereferences), seeded device-side
 * input generation, statistics of the last rounding, and timing hooks.
 */
#ifndef PARATT_AMD_EXT_H
#define PARATT_AMD_EXT_H

#include <stdint.h>
#include "paratt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Copy core k between a host array and HBM.  The host array holds
 * ttRank[k]*dim[k]*ttRank[k+1] doubles in the reference layout. */
void paratt_amd_core_upload(attac_tt *x, int k, const double *host);
void paratt_amd_core_download(const attac_tt *x, int k, double *host);

/* Fill every core with N(0,1)*scale from the project's counter-based
 * generator (seed, core index, entry index): same stream as the CPU
 * generator up to the last bits of log/cos/sin. */
void paratt_amd_fill_gaussian(attac_tt *x, uint64_t seed, double scale);

/* Rounding with a relative tolerance: tol = eps_rel * ||x||_n, where the
 * grid-normalised norm comes out of the same left-orthogonalisation sweep
 * (equal to calling paratt_norm on a copy first, without the extra sweep). */
void paratt_amd_truncate_rel(paratt *x_ptt, double eps_rel, int *maxRank);

/* Statistics of the last rounding on this process. */
typedef struct {
    int d;
    int ltr_cholqr_passes[64];   /* Cholesky-QR passes per left-to-right core */
    int rtl_cholqr_passes[64];   /* Cholesky-QR passes per right-to-left core */
    int svd_sweeps[64];          /* Jacobi sweeps per bond */
    int numrank[64];             /* numerical rank per bond */
    double norm;                 /* ||x||_n computed by the sweep */
    double sig[64][8];           /* sigma[rho-1..rho+2] around each cut (scaled) */
    double kappa_ltr[64];        /* condition estimate of each column-scaled core */
    double kappa_rtl[64];
} paratt_amd_round_stats;
void paratt_amd_last_round_stats(paratt_amd_round_stats *out);

/* Block until all queued device work of the library is done. */
void paratt_amd_sync(void);

#ifdef __cplusplus
}
#endif
#endif
