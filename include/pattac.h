/*
 * pattac.h -- MI355X-native tensor-train library: the distributed-TT entry
 * points of the reference's mpi_attac layer that lie on the TT hot path.
 *
 * Single-GPU build: every call is "collective" over a communicator of one
 * rank, so the communicator argument is accepted and ignored.  Declarations
 * match the reference so user code relinks unchanged:
 *   source/mpi_attac_src/mpi_attac/pattac.h:86-148,
 *   source/mpi_attac_src/serial_attac/attac.h:38-67.
 *
 * Memory: every `double*` core pointer handed to these functions is a DEVICE
 * address (HBM of the current GPU), laid out exactly as the reference's core
 * (column-major r_k x n_k x r_{k+1}, element (a,j,b) at a + r_k*(j + n_k*b)).
 */
#ifndef PARATT_AMD_PATTAC_H
#define PARATT_AMD_PATTAC_H

#include <stdlib.h>

/* Like the reference headers (pattac.h:21-24), pull in MPI when the
 * compiler can see it, so MPI user code keeps compiling against these
 * headers unchanged; define PARATT_AMD_NO_MPI to use the one-rank handles
 * below instead.  The library itself is built without MPI. */
#if !defined(MPI_VERSION) && !defined(PARATT_AMD_NO_MPI) && defined(__has_include)
#if __has_include(<mpi.h>)
#include <mpi.h>
#endif
#endif

#if defined(MPI_VERSION)
/* The caller already included an MPI header: use its communicator type. */
#else
/* No MPI in this build: a communicator is a plain integer handle (the same
 * size and passing convention as MPICH's MPI_Comm). */
typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Request;
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#endif
#ifndef MPI_REQUEST_NULL
#define MPI_REQUEST_NULL ((MPI_Request)0x2c000000)
#endif
#ifndef MPI_DOUBLE
#define MPI_DOUBLE ((MPI_Datatype)0x4c00080b)
#endif
#ifndef MPI_INT
#define MPI_INT ((MPI_Datatype)0x4c000405)
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque per-tensor device workspace.  Occupies the reference's
 * `tsqr_tree*` slot (mpi_attac/tsqr.h:65-82): on one GPU the TSQR reduction
 * tree collapses to an on-device R-factor computation, so only the workspace
 * pointer survives. */
typedef struct ptt_workspace tsqr_tree;

/* z = alpha*x (+) beta*y in formal (block) representation.
 * Replaces serial_attac/attac.c:81-112 (attacFormalAXPBY) and
 * mpi_attac/pattac.c:154-159 (pattacFormalAXPBY). rz is written by the call;
 * z[k] must be device buffers of at least rz[k]*n[k]*rz[k+1] doubles. */
void attacFormalAXPBY(const int d, const int *n, const double alpha,
                      const int *rx, double *const *x, const double beta,
                      const int *ry, double *const *y, int *rz, double **z);
void pattacFormalAXPBY(const int d, const int *n, const double alpha,
                       const int *rx, double *const *x, const double beta,
                       const int *ry, double *const *y, int *rz, double **z);
/* z = x (+) y.  Replaces attac.c:37-63 and pattac.c:137-139. */
void attacFormalSum(const int d, const int *n, const int *rx, double *const *x,
                    const int *ry, double *const *y, int *rz, double **z);
void pattacFormalSum(const int d, const int *n, const int *rx, double *const *x,
                     const int *ry, double *const *y, int *rz, double **z);
/* z = x .* y (Kronecker of every mode slice; z ranks rz = rx*ry are set by
 * the caller, as in the reference).  Replaces attac.c:128-136 with
 * utilities.c:112-129 (attacKron) and pattac.c:175-177. */
void attacHadamard(const int d, const int *n, const int *rx, double *const *x,
                   const int *ry, double *const *y, int *rz, double **z);
void pattacHadamard(const int d, const int *n, const int *rx, double *const *x,
                    const int *ry, double *const *y, int *rz, double **z);

/* Left-orthogonalise x in place (cores 0..d-2 become left-orthonormal,
 * rx[k+1] <- min(rx[k]*n[k], rx[k+1])).  Replaces pattac.c:244-294. */
void pattacLOrthogonalization(MPI_Comm comm, int d, int *distrib_dim_mat,
                              int *rx, double **x, tsqr_tree *qrTree,
                              int *int_wrk);
/* Right-to-left SVD truncation of a left-orthogonal TT (absolute tolerance
 * abs_tol, per-bond threshold abs_tol/sqrt(d-1), caps maxr[0..d]).
 * Replaces pattac.c:297-522.  flt_wrk_ptr/flt_sz_ptr are accepted for
 * signature compatibility; device workspace lives in qrTree. */
void pattacLTruncate(MPI_Comm comm, const int d, int *distrib_dim_mat, int *rx,
                     double **x, const int *maxr, const double abs_tol,
                     tsqr_tree *qrTree, int *int_wrk, double **flt_wrk_ptr,
                     int *flt_sz_ptr);
/* Full rounding in the grid-normalised norm: as pattac.c:590-675 (scale by
 * 1/sqrt(n_k), left-orthogonalise, truncate with abs_tol, unscale).  Only
 * ortho == 'R' is implemented, as in the reference's live path; 'L' aborts. */
void pattacTruncate(MPI_Comm comm, const int d, int *distrib_dim_mat, int *rx,
                    double **x, const int *maxr, const double abs_tol,
                    const double rel_tol, const char ortho, tsqr_tree *qrTree,
                    int *int_wrk, double **flt_wrk_ptr, int *flt_sz_ptr);
/* <x, y> by the left-to-right contraction sweep.  Replaces pattac.c:693-741.
 * work is ignored (device workspace is internal). */
double pattacInnerProduct(const MPI_Comm comm, const int *isRedundant,
                          const int d, const int *n, const int *rx,
                          double *const *x, const int *ry, double *const *y,
                          double *work);
/* Entry x[index[0], ..., index[d-1]] by the chain of matrix-vector products
 * from the last core to the first.  Replaces pattac.c:82-121; procHolder and
 * work are ignored (one rank; device workspace is internal).  An index out
 * of range is fatal. */
double pattacValue(const MPI_Comm comm, const int *index, const int *procHolder,
                   const int d, const int *n, const int *rx, double *const *x,
                   double *work);

#ifdef __cplusplus
}
#endif
#endif
