/*
 * pastix_hip.h -- GPU kernel entry points with the reference's signatures,
 * for HIP on MI355X (gfx950).
 *
 * These replace the CUDA entry points of the reference
 * (kernels/pastix_zcuda.h:32-58, generated to the d/s precisions) with the
 * same arguments, hipStream_t in place of cudaStream_t.  A, B and C are device
 * pointers as in the reference: the start of the panel (cblk entry points) or
 * of the first block concerned (blok entry points); cblk/blok/fcblk are host
 * structures read by the host wrapper.  Calls are asynchronous on `stream`
 * and return the flop count of the operation, as the reference does; a call
 * with arguments outside the supported cases returns -1.
 *
 * gpucblk_{d,s}potrfsp and gpucblk_{d,s}getrfsp have no GPU counterpart in
 * the reference (its diagonal factorizations are CPU-only codelets,
 * sopalin/starpu/codelet_cblk_zpotrfsp.c:200); they follow the CPU kernels
 * cpucblk_zpotrfsp1d_potrf / cpucblk_zgetrfsp1d_getrf
 * (kernels/core_zpotrfsp.c:220, kernels/core_zgetrfsp.c:217).
 */
#ifndef PASTIX_AMD_HIP_H
#define PASTIX_AMD_HIP_H

#include <hip/hip_runtime_api.h>
#include "pastix/solver.h"

#ifdef __cplusplus
extern "C" {
#endif

struct pastix_lr_s;
typedef struct pastix_lr_s pastix_lr_t;

/* Replaces gpucblk_zgemmsp (reference kernels/gpu_zgemmsp.c:139,
 * kernels/pastix_zcuda.h:32): update of fcblk by block `blok` of cblk and
 * every block below it (C -= A_i B_blok^T, scattered into the 2D tiles). */
pastix_fixdbl_t gpucblk_dgemmsp( pastix_coefside_t sideA, pastix_trans_t trans,
                                 const SolverCblk *cblk, const SolverBlok *blok, SolverCblk *fcblk,
                                 const double *A, const double *B, double *C,
                                 const pastix_lr_t *lowrank, hipStream_t stream );
pastix_fixdbl_t gpucblk_sgemmsp( pastix_coefside_t sideA, pastix_trans_t trans,
                                 const SolverCblk *cblk, const SolverBlok *blok, SolverCblk *fcblk,
                                 const float *A, const float *B, float *C,
                                 const pastix_lr_t *lowrank, hipStream_t stream );

/* Replaces gpublok_zgemmsp (reference kernels/gpu_zgemmsp.c:486,
 * kernels/pastix_zcuda.h:37): 2D-task update between the blocks of cblk facing
 * the same cblk as blok_mk / blok_nk, into block blok_mn of fcblk. */
pastix_fixdbl_t gpublok_dgemmsp( pastix_trans_t trans, const SolverCblk *cblk, SolverCblk *fcblk,
                                 pastix_int_t blok_mk, pastix_int_t blok_nk, pastix_int_t blok_mn,
                                 const void *A, const void *B, void *C,
                                 const pastix_lr_t *lowrank, hipStream_t stream );
pastix_fixdbl_t gpublok_sgemmsp( pastix_trans_t trans, const SolverCblk *cblk, SolverCblk *fcblk,
                                 pastix_int_t blok_mk, pastix_int_t blok_nk, pastix_int_t blok_mn,
                                 const void *A, const void *B, void *C,
                                 const pastix_lr_t *lowrank, hipStream_t stream );

/* Replaces gpublok_ztrsmsp (reference kernels/gpu_ztrsmsp.c:189,
 * kernels/pastix_zcuda.h:43): C_b <- C_b op(A)^{-1} for the blocks of cblk
 * facing the same cblk as blok_m.  Supported: (Right, Lower, Trans|ConjTrans,
 * NonUnit) and (Right, Upper, NoTrans, NonUnit|Unit). */
pastix_fixdbl_t gpublok_dtrsmsp( pastix_side_t side, pastix_uplo_t uplo, pastix_trans_t trans,
                                 pastix_diag_t diag, const SolverCblk *cblk, pastix_int_t blok_m,
                                 const void *A, void *C, const pastix_lr_t *lowrank, hipStream_t stream );
pastix_fixdbl_t gpublok_strsmsp( pastix_side_t side, pastix_uplo_t uplo, pastix_trans_t trans,
                                 pastix_diag_t diag, const SolverCblk *cblk, pastix_int_t blok_m,
                                 const void *A, void *C, const pastix_lr_t *lowrank, hipStream_t stream );

/* Diagonal block factorizations with static pivoting on the GPU.  L (and U)
 * point at the cblk panel; d_nbpivots points at two device int32 counters:
 * [0] += replaced pivots, [1] += negative Cholesky diagonal terms. */
pastix_fixdbl_t gpucblk_dpotrfsp( const SolverCblk *cblk, void *L, double criterion,
                                  int32_t *d_nbpivots, hipStream_t stream );
pastix_fixdbl_t gpucblk_spotrfsp( const SolverCblk *cblk, void *L, double criterion,
                                  int32_t *d_nbpivots, hipStream_t stream );
pastix_fixdbl_t gpucblk_dgetrfsp( const SolverCblk *cblk, void *L, void *U, double criterion,
                                  int32_t *d_nbpivots, hipStream_t stream );
pastix_fixdbl_t gpucblk_sgetrfsp( const SolverCblk *cblk, void *L, void *U, double criterion,
                                  int32_t *d_nbpivots, hipStream_t stream );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_HIP_H */
