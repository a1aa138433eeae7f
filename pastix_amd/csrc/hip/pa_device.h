/*
 * pa_device.h -- the thin C-ABI between the C host library (API, analysis,
 * refinement drivers) and the HIP device layer.  Every pointer named d_* is
 * device memory; vectors of the refinement live in the factorization
 * (permuted) numbering.
 */
#ifndef PASTIX_AMD_DEVICE_H
#define PASTIX_AMD_DEVICE_H

#include <stdint.h>
#include "pastix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* factor arithmetic */
#define PA_PREC_FLOAT  2
#define PA_PREC_DOUBLE 3

int    pa_dev_open( pastix_data_t *pastix_data );
void   pa_dev_close( pastix_data_t *pastix_data );
/* upload the SolverMatrix, permutation and the level schedule */
int    pa_dev_setup_structure( pastix_data_t *pastix_data );
/* schedule autotuning: number of candidates, whether candidate c differs
 * from the default for this problem, select candidate c (or explicit
 * params {tm_min, narrow, small} when c < 0) and rebuild the task lists,
 * read the current choice (params[4]) */
int    pa_dev_tune_ncand( void );
int    pa_dev_tune_applicable( const pastix_data_t *pastix_data, int c );
int    pa_dev_set_tune( pastix_data_t *pastix_data, int c, const int32_t *params );
void   pa_dev_get_tune( const pastix_data_t *pastix_data, int32_t *params );
/* upload the permuted CSC (pattern + FP64 values) */
int    pa_dev_upload_bcsc( pastix_data_t *pastix_data );
/* upload the value map of the permuted CSC (see pa_valmap_t) */
int    pa_dev_upload_valmap( pastix_data_t *pastix_data, const pastix_int_t *Lptr, const pastix_int_t *Lsrc,
                             const pastix_int_t *Uptr, const pastix_int_t *Usrc, int64_t nnz );
/* gather new FP64 user values (device, user CSC order) into the permuted
 * device CSC through the value map; *normF gets ||A||_F */
int    pa_dev_set_values( pastix_data_t *pastix_data, const double *d_values, double *normF );
/* allocate the coeftab in HBM (prec) and fill it from the device CSC */
int    pa_dev_coeftab_init( pastix_data_t *pastix_data, int prec, int lu, int32_t *overflow );
/* level-batched numerical factorization; returns PASTIX_SUCCESS */
int    pa_dev_factorize( pastix_data_t *pastix_data, double criterion, int32_t *npiv,
                         int32_t *nneg, double *seconds );
/* in-place solve of op(A) x = b on a device FP64 buffer in the factorization
 * numbering (nrhs columns, leading dimension ld) */
int    pa_dev_solve_perm( pastix_data_t *pastix_data, pastix_trans_t trans, int32_t nrhs,
                          double *d_x, int64_t ld );
/* pa_dev_solve_perm plus the permutations, on a user-numbering device FP64
 * buffer, with no device allocation (the device's static scratch array);
 * pa_dev_noalloc_max gives the largest n * nrhs it takes */
int     pa_dev_solve_noalloc( pastix_data_t *pastix_data, pastix_trans_t trans, int32_t nrhs, double *d_b,
                              int64_t ldb );
int64_t pa_dev_noalloc_max( pastix_data_t *pastix_data );
/* factor arrays to and from host memory: what HBM holds (arithmetic, two
 * arrays, factorization kind, coefficients per array); raw copies of count
 * coefficients at offset off of array side (0: L, 1: second array) in the
 * factors' arithmetic; begin allocates the coeftab for the current
 * IPARM_FACTORIZATION without filling it, end builds what the solve derives
 * from the factors */
/* flops the factorization kernels perform (GEMM tiles with their padding) */
double pa_dev_real_flops( const pastix_data_t *pastix_data );
int    pa_dev_factor_info( const pastix_data_t *pastix_data, int *prec, int *two, int *fact, int64_t *coefnbr );
int    pa_dev_factor_get( const pastix_data_t *pastix_data, int side, int64_t off, int64_t count, void *host );
int    pa_dev_factor_begin( pastix_data_t *pastix_data, int prec, int lu );
int    pa_dev_factor_put( pastix_data_t *pastix_data, int side, int64_t off, int64_t count, const void *host );
int    pa_dev_factor_end( pastix_data_t *pastix_data );
/* one triangular sweep on a device FP64 buffer in the factorization
 * numbering: upper = 0 sweeps L, upper = 1 the stored U^T; transposed
 * selects op(factor) = factor^T; unit ignores the stored diagonal */
int    pa_dev_trsm_perm( pastix_data_t *pastix_data, int upper, int transposed, int unit, int32_t nrhs,
                         double *d_x, int64_t ld );
/* x <- D^{-1} x with D the diagonal of the L diagonal blocks */
int    pa_dev_diag_perm( pastix_data_t *pastix_data, int32_t nrhs, double *d_x, int64_t ld );
/* y = alpha op(A) x + beta y, factorization numbering */
int    pa_dev_spmv_perm( pastix_data_t *pastix_data, pastix_trans_t trans, double alpha,
                         const double *d_x, double beta, double *d_y );
/* user numbering <-> factorization numbering (nrhs columns) */
int    pa_dev_apply_perm( pastix_data_t *pastix_data, int forward, int32_t nrhs,
                          const double *d_src, int64_t lds, double *d_dst, int64_t ldd );

/* FP64 vector helpers on the library stream */
void  *pa_dev_malloc( size_t bytes );
void   pa_dev_free( void *d_ptr );
/* work buffers of the solve/refinement drivers: a block of at least `bytes`
 * from the handle's pool (hipMalloc only when no free block is large
 * enough), returned with pa_dev_wput; the pool is released at finalize.
 * The same call sequence gets the same addresses, which keeps the cached
 * solve graphs valid. */
void  *pa_dev_wget( pastix_data_t *pastix_data, size_t bytes );
void   pa_dev_wput( pastix_data_t *pastix_data, void *d_ptr );
int    pa_dev_memcpy_h2d( pastix_data_t *pastix_data, void *d_dst, const void *h_src, size_t bytes );
int    pa_dev_memcpy_d2h( pastix_data_t *pastix_data, void *h_dst, const void *d_src, size_t bytes );
int    pa_dev_memcpy_d2d( pastix_data_t *pastix_data, void *d_dst, const void *d_src, size_t bytes );
double pa_dev_dot( pastix_data_t *pastix_data, int64_t n, const double *d_x, const double *d_y );
double pa_dev_nrm2( pastix_data_t *pastix_data, int64_t n, const double *d_x );
/* out[j] = X[:, j] . v for j < m (X is n x m, ld n) */
int    pa_dev_mdot( pastix_data_t *pastix_data, int64_t n, int32_t m, const double *d_X,
                    const double *d_v, double *out );
/* y = a x + b y */
int    pa_dev_axpby( pastix_data_t *pastix_data, int64_t n, double a, const double *d_x,
                     double b, double *d_y );
/* y += alpha X c, X n x m (ld n), c on the host */
int    pa_dev_gemv( pastix_data_t *pastix_data, int64_t n, int32_t m, const double *d_X,
                    const double *c, double alpha, double *d_y );
int    pa_dev_sync( pastix_data_t *pastix_data );
/* copy a factor from HBM to the host as FP64 (side 0: L, 1: U^T) */
int    pa_dev_get_coeftab( const pastix_data_t *pastix_data, int side, double *host );
pastix_int_t pa_dev_nlevels( const pastix_data_t *pastix_data );
/* enable (1) / disable (0) / keep (-1) per-class kernel timing; out[9] gets
 * the accumulated ms of the diag/trsm/gemm kernels, their launch counts and
 * the algorithmic flops of each class for one factorization */
int    pa_dev_profile( pastix_data_t *pastix_data, int enable, double *out );
double pa_dev_last_fact_event_seconds( const pastix_data_t *pastix_data );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_DEVICE_H */
