/*
 * pastix.h -- public C API of the MI355X PaStiX factor+solve engine.
 *
 * The entry points keep the names, argument meaning, return codes and step
 * ordering of the reference API (reference include/pastix.h:56-160,
 * common/api.c, sopalin/pastix_task_sopalin.c:654, sopalin/pastix_task_solve.c:758,
 * refinement/pastix_task_refine.c:236).  Numerical work (coeftab fill,
 * factorization, solve, SpMV, refinement) runs on the GPU; the integer
 * analysis (ordering, symbolic factorization, splitting, solver structure)
 * runs on the host.
 *
 * Extensions without a reference counterpart are prefixed `pastix_amd_`:
 * device-resident solve/refine on HBM buffers, structure export, and the
 * multi-system batched pass.
 */
#ifndef PASTIX_AMD_PASTIX_H
#define PASTIX_AMD_PASTIX_H

#include "pastix/datatypes.h"
#include "pastix/api.h"
#include "pastix/order.h"
#include "pastix/solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Legacy one-call interface (reference sopalin/pastix.c:103). */
int pastix( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm, pastix_int_t n,
            pastix_int_t *colptr, pastix_int_t *rowptr, void *values,
            pastix_int_t *perm, pastix_int_t *invp, void *B, pastix_int_t nrhs,
            pastix_int_t *iparm, double *dparm );

void pastixInitParam( pastix_int_t *iparm, double *dparm );
void pastixInit( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
                 pastix_int_t *iparm, double *dparm );
void pastixInitWithAffinity( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
                             pastix_int_t *iparm, double *dparm, const int *bindtab );
void pastixFinalize( pastix_data_t **pastix_data );

int pastix_task_analyze( pastix_data_t *pastix_data, const spmatrix_t *spm );
int pastix_task_numfact( pastix_data_t *pastix_data, spmatrix_t *spm );
int pastix_task_solve( pastix_data_t *pastix_data, pastix_int_t m, pastix_int_t nrhs,
                       void *B, pastix_int_t ldb );
int pastix_task_refine( pastix_data_t *pastix_data, pastix_int_t n, pastix_int_t nrhs,
                        void *B, pastix_int_t ldb, void *X, pastix_int_t ldx );
int pastix_task_solve_and_refine( pastix_data_t *pastix_data, pastix_int_t n, pastix_int_t nrhs,
                                  void *B, pastix_int_t ldb, void *X, pastix_int_t ldx );

int pastix_subtask_order( pastix_data_t *pastix_data, const spmatrix_t *spm,
                          pastix_order_t *myorder );
int pastix_subtask_symbfact( pastix_data_t *pastix_data );
int pastix_subtask_reordering( pastix_data_t *pastix_data );
int pastix_subtask_blend( pastix_data_t *pastix_data );

int pastix_subtask_spm2bcsc( pastix_data_t *pastix_data, spmatrix_t *spm );
int pastix_subtask_bcsc2ctab( pastix_data_t *pastix_data );
int pastix_subtask_sopalin( pastix_data_t *pastix_data );

/* Solve and refinement subtasks (reference include/pastix.h:133-152).  The
 * right-hand sides live in host memory, in the matrix's arithmetic; each
 * subtask copies them to the device and back. */
int pastix_subtask_applyorder( pastix_data_t *pastix_data, pastix_dir_t dir, pastix_int_t m,
                               pastix_int_t n, void *B, pastix_int_t ldb, pastix_rhs_t Bp );
int pastix_subtask_trsm( pastix_data_t *pastix_data, pastix_side_t side, pastix_uplo_t uplo,
                         pastix_trans_t trans, pastix_diag_t diag, pastix_rhs_t b );
int pastix_subtask_diag( pastix_data_t *pastix_data, pastix_rhs_t b );
int pastix_subtask_solve( pastix_data_t *pastix_data, pastix_rhs_t b );
int pastix_subtask_refine( pastix_data_t *pastix_data, pastix_rhs_t b, pastix_rhs_t x );
int pastix_subtask_solve_adv( pastix_data_t *pastix_data, pastix_trans_t transA,
                              pastix_rhs_t b );

/* Right-hand side handles (reference include/pastix.h:172-182) */
int pastixRhsInit( pastix_rhs_t *rhs );
int pastixRhsFinalize( pastix_rhs_t rhs );
int pastixRhsDoubletoSingle( const pastix_rhs_t dB, pastix_rhs_t sB );
int pastixRhsSingleToDouble( const pastix_rhs_t sB, pastix_rhs_t dB );

int  pastixCheckParam( const pastix_int_t *iparm, const double *dparm );
void pastixDumpParam( const pastix_data_t *pastix_data );

/* ---- factor I/O ----------------------------------------------------------- */
/* Human-readable dump of the factors in HBM (reference coeftab_zdump,
 * sopalin/coeftab_z.c:55, and cpucblk_zdump, sopalin/coeftab_zinit.c:54):
 * one file <prefix>_<cblk>.txt per column block holding the lines
 * "row col value [L]" and, for LU / LDL^t, "col row value [U]" of its nonzero
 * coefficients (lower triangle of the diagonal block).  The d / s variant
 * matches the factors' arithmetic; the prefix may hold a directory. */
void coeftab_ddump( pastix_data_t *pastix_data, const SolverMatrix *solvmtx, const char *prefix );
void coeftab_sdump( pastix_data_t *pastix_data, const SolverMatrix *solvmtx, const char *prefix );

/* ---- extensions (no reference counterpart) -------------------------------- */

/*
 * Solve op(A) X = B in place on a DEVICE buffer dB (n x nrhs, leading
 * dimension ldb, FP64) that is already resident in HBM: permutation, forward
 * and backward sweeps run on the GPU, nothing crosses PCIe.  transA is
 * PastixNoTrans or PastixTrans (the adjoint solve of the differentiable
 * wrapper).  Same algorithm as pastix_task_solve.
 */
int pastix_amd_solve_dev( pastix_data_t *pastix_data, pastix_trans_t transA,
                          pastix_int_t nrhs, void *dB, pastix_int_t ldb );

/*
 * Factor checkpoints.  pastix_amd_factors_save writes the factors held in
 * HBM (after pastix_task_numfact) to a binary file: a header with the
 * factorization kind, arithmetic, static pivot count, threshold and a
 * fingerprint of the block structure, then the coefficient arrays in the
 * 2D tile layout.  pastix_amd_factors_load, on an instance where
 * pastix_task_analyze (same matrix pattern, ordering and parameters) and
 * pastix_subtask_spm2bcsc have run, puts them back in HBM in place of
 * bcsc2ctab + sopalin: solves and refinements then run as after a
 * factorization.  A file of another structure, factorization kind or
 * precision is refused with PASTIX_ERR_BADPARAMETER.
 */
int pastix_amd_factors_save( pastix_data_t *pastix_data, const char *path );
int pastix_amd_factors_load( pastix_data_t *pastix_data, const char *path );

/*
 * pastix_amd_solve_dev without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, same kernels and same results; the permuted right-hand
 * sides, their FP32 copy for FP32 factors and the second array of the
 * triangular sweeps live in a static 128 MiB device array per GPU (16 bytes
 * per entry), so n * nrhs may be at most 8,388,576; larger calls return
 * PASTIX_ERR_BADPARAMETER before any launch.  Launches are eager (no graph)
 * and the call returns after the solve has finished.  The array is shared by
 * every handle on a GPU: calls that overlap on one GPU race.
 */
int pastix_amd_solve_dev_noalloc( pastix_data_t *pastix_data, pastix_trans_t transA,
                                  pastix_int_t nrhs, void *dB, pastix_int_t ldb );

/*
 * Refine X on the device: FP64 GMRES/CG/SR (IPARM_REFINEMENT) with the
 * factorization as preconditioner, device SpMV on the permuted CSC and device
 * norms.  dB and dX are device FP64 buffers in the user numbering.
 * Returns the number of iterations (>= 0) or -PASTIX_ERR_*.
 */
int pastix_amd_refine_dev( pastix_data_t *pastix_data, pastix_int_t nrhs,
                           const void *dB, pastix_int_t ldb, void *dX, pastix_int_t ldx );
/* Same with op(A) given explicitly (PastixNoTrans or PastixTrans) instead of
 * IPARM_TRANSPOSE_SOLVE: the refined adjoint solve A^T x = b of the
 * differentiable wrapper. */
int pastix_amd_refine_dev_adv( pastix_data_t *pastix_data, pastix_trans_t transA, pastix_int_t nrhs,
                               const void *dB, pastix_int_t ldb, void *dX, pastix_int_t ldx );

/* y = alpha*op(A)*x + beta*y with the user's matrix, x/y device FP64 vectors
 * in the user numbering. */
int pastix_amd_spmv_dev( pastix_data_t *pastix_data, pastix_trans_t trans, double alpha,
                         const void *dx, double beta, void *dy );

/*
 * New numerical values for the analysed pattern, from device memory: d_values
 * is a device FP64 array with one value per stored entry of the spm last
 * given to pastix_task_numfact / pastix_subtask_spm2bcsc, in that spm's
 * order.  The permuted CSC is rebuilt on the device and ||A||_F
 * (DPARM_A_NORM) recomputed there; then pastix_subtask_bcsc2ctab +
 * pastix_subtask_sopalin factor the new values.  Replaces re-entering
 * pastix_subtask_spm2bcsc with a host spm (reference
 * sopalin/pastix_task_sopalin.c:93-210).  The first call builds the value
 * map from that spm, which must still be alive then.
 */
int pastix_amd_values_dev( pastix_data_t *pastix_data, const void *d_values );

/*
 * Adjoint helper of the differentiable solve: out[e] = alpha * sum_r
 * u[row(e), r] * v[col(e), r] for every stored entry e of the user's CSC
 * pattern (colptr[n+1], rows[nnz]: pastix_int_t device arrays; u, v: device
 * FP64 n x nrhs, leading dimensions ldu, ldv; out: device FP64 [nnz]).
 * With u = lambda = A^{-T} g, v = x and alpha = -1 this is dLoss/dA on the
 * pattern.  Runs on the library stream and synchronises it.
 */
int pastix_amd_sampled_outer( pastix_data_t *pastix_data, const void *d_colptr, const void *d_rows,
                              pastix_int_t nrhs, const void *d_u, pastix_int_t ldu, const void *d_v,
                              pastix_int_t ldv, double alpha, void *d_out );

/*
 * Autotuner of the launch schedule: the GEMM tile classes (128-row tiles
 * from how many update rows, in FP64 too (tall64), tile width following the
 * column-group width)
 * and the fused narrow-cblk kernels.  Each of the PASTIX_AMD_TUNE_NCAND
 * candidates that differs for this problem factors the current matrix
 * (after pastix_task_numfact) and is timed on the device; the fastest is
 * kept and its factors left in place.  times (may be NULL) receives each
 * candidate's factorization seconds (0: not applicable), params[4] (may be
 * NULL) the chosen {tm_min, narrow, small, tall64}.  Returns the chosen
 * candidate's index, or -PASTIX_ERR_*.  No reference counterpart (the
 * reference leaves these choices to its StarPU/PaRSEC runtime).
 */
#define PASTIX_AMD_TUNE_NCAND 8
int pastix_amd_tune( pastix_data_t *pastix_data, double *times, pastix_int_t *params );
/* Select a schedule chosen earlier (params from pastix_amd_tune) for the
 * next factorizations. */
int pastix_amd_set_tuning( pastix_data_t *pastix_data, const pastix_int_t *params );

/* Synchronise the library stream and return the elapsed seconds of the last
 * factorization measured with device events (0 if none). */
double pastix_amd_last_fact_seconds( pastix_data_t *pastix_data );

/*
 * Kernel profiling of the factorization with HIP events on the library
 * stream.  enable = 1 starts (and resets) accumulation, 0 stops, -1 only
 * reads.  out[9] (may be NULL): accumulated milliseconds of the diagonal,
 * TRSM and scatter-GEMM kernels, their launch counts, and the algorithmic
 * flops of each class for one factorization (reference flop model).
 */
int pastix_amd_profile( pastix_data_t *pastix_data, int enable, double *out );

/*
 * Structure export for tests and tooling.  Fills `dims`:
 *  [0] n, [1] cblknbr, [2] bloknbr, [3] coefnbr (per factor), [4] browtab size,
 *  [5] number of levels of the batched schedule, [6] factotype, [7] flttype
 * and returns PASTIX_SUCCESS once the analysis has been done.
 */
int pastix_amd_get_dims( const pastix_data_t *pastix_data, pastix_int_t *dims );
/*
 * Copies the structure into caller arrays (any may be NULL):
 *  cblk[5*cblknbr]: fcolnum, lcolnum, fblok index, stride, coefficient offset
 *  blok[5*bloknbr]: frownum, lrownum, lcblknm, fcblknm, coefind
 *  perm[n], invp[n]
 */
int pastix_amd_get_structure( const pastix_data_t *pastix_data, pastix_int_t *cblk,
                              pastix_int_t *blok, pastix_int_t *perm, pastix_int_t *invp );
/* The instance's SolverMatrix (after the analysis), e.g. for solverSave. */
const SolverMatrix *pastix_amd_get_solver( const pastix_data_t *pastix_data );
/*
 * The static schedule of the CPU factorization computed by the analysis
 * (proportional mapping + simulation over IPARM_THREAD_NBR threads):
 * *thrdnbr threads; when tptr is given (thrdnbr+1 entries), thread t runs
 * the cblks ttsk[tptr[t] .. tptr[t+1]) (cblknbr entries in all) in order.
 */
int pastix_amd_get_schedule( const pastix_data_t *pastix_data, pastix_int_t *thrdnbr,
                             pastix_int_t *tptr, pastix_int_t *ttsk );
/*
 * Copies a factor (side 0 = L, 1 = U^T) from HBM into `host` (coefnbr
 * entries, FP64 even when the factorization ran in FP32).
 */
int pastix_amd_get_coeftab( const pastix_data_t *pastix_data, int side, double *host );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_PASTIX_H */
