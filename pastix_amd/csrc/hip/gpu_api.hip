/*
 * gpu_api.hip -- reference-signature GPU entry points (see include/pastix_hip.h).
 *
 * Each entry point turns the host SolverCblk/SolverBlok description of one
 * update into by-value row tables (pa_call_rows, at most PA_CALL_MAXBLK
 * blocks per launch) and runs the same MFMA tile core and TRSM/diagonal
 * device functions as the level-batched schedule.
 */
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include "kernels.hpp"

extern "C" {
#include "pastix_hip.h"


}

/* The kernels of these entry points stage up to the device's LDS per
 * workgroup: raise their dynamic limit once per process (the current device) */
template <typename K>
static void
pa_api_lds_attr( K kernel, int bytes )
{
    hipFuncAttributes fa;
    int               stat = 0;
    if ( hipFuncGetAttributes( &fa, reinterpret_cast<const void *>( kernel ) ) == hipSuccess ) stat = (int)fa.sharedSizeBytes;
    (void)hipGetLastError();
    if ( hipFuncSetAttribute( reinterpret_cast<const void *>( kernel ), hipFuncAttributeMaxDynamicSharedMemorySize,
                              bytes - stat ) != hipSuccess )
        (void)hipGetLastError();
}

static void
pa_api_lds_limits( void )
{
    static bool done = false;
    if ( done ) return;
    done = true;
    int dev = 0, bytes = 65536;
    if ( hipGetDevice( &dev ) == hipSuccess ) {
        int v = 0;
        if ( hipDeviceGetAttribute( &v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev ) == hipSuccess && v > 0 ) bytes = v;
    }
    (void)hipGetLastError();
    pa_api_lds_attr( k_diag_one<double, true>, bytes ); pa_api_lds_attr( k_diag_one<double, false>, bytes );
    pa_api_lds_attr( k_diag_one<float, true>, bytes );  pa_api_lds_attr( k_diag_one<float, false>, bytes );
    pa_api_lds_attr( k_trsm_call<double, 0>, bytes ); pa_api_lds_attr( k_trsm_call<double, 1>, bytes );
    pa_api_lds_attr( k_trsm_call<double, 2>, bytes );
    pa_api_lds_attr( k_trsm_call<float, 0>, bytes );  pa_api_lds_attr( k_trsm_call<float, 1>, bytes );
    pa_api_lds_attr( k_trsm_call<float, 2>, bytes );
}

static inline pastix_int_t
pa_tile_ld( const SolverCblk *cblk, const SolverBlok *blok )
{
    return ( cblk->cblktype & CBLK_LAYOUT_2D ) ? blok_rownbr( blok ) : cblk->stride;
}

static inline double pa_fl_gemm( double m, double n, double k ) { return 2. * m * n * k; }
static inline double pa_fl_trsm( double m, double n ) { return m * n * ( n + 1. ); }

/* launch one chunk of rho blocks against the sigma run */
template <typename T>
static int
pa_launch_gemm_call( const T *A, const T *B, T *C, int32_t K, const pa_call_rows &Ar, const pa_call_rows &Br,
                     int32_t ccol0, hipStream_t stream )
{
    int32_t mt = Ar.pref[Ar.nblk], nt = Br.pref[Br.nblk];
    if ( mt <= 0 || nt <= 0 ) return 0;
    dim3 grid( ( mt + PA_GT - 1 ) / PA_GT, ( nt + PA_GT - 1 ) / PA_GT );
    k_gemm_call<T><<<grid, 256, 0, stream>>>( A, B, C, K, Ar, Br, ccol0 );
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
static pastix_fixdbl_t
pa_gpucblk_gemmsp( pastix_coefside_t sideA, pastix_trans_t trans, const SolverCblk *cblk,
                   const SolverBlok *blok, SolverCblk *fcblk, const T *A, const T *B, T *C,
                   hipStream_t stream )
{
    if ( ( cblk->cblktype & CBLK_COMPRESSED ) || ( fcblk->cblktype & CBLK_COMPRESSED ) || trans == PastixNoTrans ) {
        fprintf( stderr, "gpucblk_gemmsp: unsupported arguments (compressed cblk or NoTrans)\n" );
        return -1.;
    }
    const int shift = ( sideA == PastixUCoef ) ? 1 : 0;
    const int32_t K = (int32_t)cblk_colnbr( cblk );
    const int32_t N = (int32_t)blok_rownbr( blok );
    const SolverBlok *lblok = cblk[1].fblokptr;
    const SolverBlok *fblok = fcblk->fblokptr;
    const SolverBlok *fend  = fcblk[1].fblokptr;

    pa_call_rows Br;
    memset( &Br, 0, sizeof( Br ) );
    Br.nblk    = 1;
    Br.pref[0] = 0;
    Br.pref[1] = N;
    Br.off[0]  = blok->coefind;
    Br.ld[0]   = (int32_t)pa_tile_ld( cblk, blok );
    Br.toff[0] = 0;
    const int32_t ccol0 = (int32_t)( blok->frownum - fcblk->fcolnum );

    pa_call_rows Ar;
    memset( &Ar, 0, sizeof( Ar ) );
    int32_t rows = 0;
    for ( const SolverBlok *it = blok + shift; it < lblok; it++ ) {
        while ( !is_block_inside_fblock( it, fblok ) ) {
            fblok++;
            if ( fblok >= fend ) {
                fprintf( stderr, "gpucblk_gemmsp: block rows %ld-%ld have no facing block\n",
                         (long)it->frownum, (long)it->lrownum );
                return -1.;
            }
        }
        int32_t b = Ar.nblk;
        Ar.pref[b] = rows;
        Ar.off[b]  = it->coefind;
        Ar.ld[b]   = (int32_t)pa_tile_ld( cblk, it );
        Ar.toff[b] = fblok->coefind + ( it->frownum - fblok->frownum );
        Ar.tld[b]  = (int32_t)pa_tile_ld( fcblk, fblok );
        rows += (int32_t)blok_rownbr( it );
        Ar.nblk++;
        if ( Ar.nblk == PA_CALL_MAXBLK ) {
            Ar.pref[Ar.nblk] = rows;
            if ( pa_launch_gemm_call<T>( A, B, C, K, Ar, Br, ccol0, stream ) ) return -1.;
            memset( &Ar, 0, sizeof( Ar ) );
            rows = 0;
        }
    }
    if ( Ar.nblk > 0 ) {
        Ar.pref[Ar.nblk] = rows;
        if ( pa_launch_gemm_call<T>( A, B, C, K, Ar, Br, ccol0, stream ) ) return -1.;
    }
    /* flops as in the reference (gpu_zgemmsp.c:232-240) */
    pastix_int_t m = cblk->stride;
    m -= ( cblk->cblktype & CBLK_LAYOUT_2D ) ? blok->coefind / K : blok->coefind;
    m -= shift ? N : 0;
    return pa_fl_gemm( (double)m, (double)N, (double)K );
}

extern "C" pastix_fixdbl_t
gpucblk_dgemmsp( pastix_coefside_t sideA, pastix_trans_t trans, const SolverCblk *cblk, const SolverBlok *blok,
                 SolverCblk *fcblk, const double *A, const double *B, double *C, const pastix_lr_t *lowrank,
                 hipStream_t stream )
{
    (void)lowrank;
    return pa_gpucblk_gemmsp<double>( sideA, trans, cblk, blok, fcblk, A, B, C, stream );
}

extern "C" pastix_fixdbl_t
gpucblk_sgemmsp( pastix_coefside_t sideA, pastix_trans_t trans, const SolverCblk *cblk, const SolverBlok *blok,
                 SolverCblk *fcblk, const float *A, const float *B, float *C, const pastix_lr_t *lowrank,
                 hipStream_t stream )
{
    (void)lowrank;
    return pa_gpucblk_gemmsp<float>( sideA, trans, cblk, blok, fcblk, A, B, C, stream );
}

template <typename T>
static pastix_fixdbl_t
pa_gpublok_gemmsp( pastix_trans_t trans, const SolverCblk *cblk, SolverCblk *fcblk, pastix_int_t blok_mk,
                   pastix_int_t blok_nk, pastix_int_t blok_mn, const T *A, const T *B, T *C, hipStream_t stream )
{
    if ( !( cblk->cblktype & CBLK_LAYOUT_2D ) || !( fcblk->cblktype & CBLK_LAYOUT_2D ) ||
         ( cblk->cblktype & CBLK_COMPRESSED ) || ( fcblk->cblktype & CBLK_COMPRESSED ) || trans == PastixNoTrans ) {
        fprintf( stderr, "gpublok_gemmsp: unsupported arguments\n" );
        return -1.;
    }
    const SolverBlok *fblokK = cblk[0].fblokptr, *lblokK = cblk[1].fblokptr;
    const SolverBlok *blokB  = fblokK + blok_nk, *blokA = fblokK + blok_mk;
    const SolverBlok *blokC  = fcblk[0].fblokptr + blok_mn, *lblokN = fcblk[1].fblokptr;
    const pastix_int_t offA = blokA->coefind, offB = blokB->coefind, offC = blokC->coefind;
    const pastix_int_t cblk_m = blokA->fcblknm, cblk_n = blokB->fcblknm;
    const int32_t K = (int32_t)cblk_colnbr( cblk );
    double flops = 0.;

    pa_call_rows Br;
    memset( &Br, 0, sizeof( Br ) );
    int32_t ncols = 0;
    for ( const SolverBlok *bB = blokB; bB < lblokK && bB->fcblknm == cblk_n; bB++ ) {
        if ( Br.nblk == PA_CALL_MAXBLK ) {
            fprintf( stderr, "gpublok_gemmsp: more than %d column blocks\n", PA_CALL_MAXBLK );
            return -1.;
        }
        Br.pref[Br.nblk] = ncols;
        Br.off[Br.nblk]  = bB->coefind - offB;
        Br.ld[Br.nblk]   = (int32_t)blok_rownbr( bB );
        Br.toff[Br.nblk] = bB->frownum - fcblk->fcolnum;
        ncols += (int32_t)blok_rownbr( bB );
        Br.nblk++;
    }
    Br.pref[Br.nblk] = ncols;

    pa_call_rows Ar;
    memset( &Ar, 0, sizeof( Ar ) );
    int32_t rows = 0;
    const SolverBlok *bC = blokC;
    for ( const SolverBlok *bA = blokA; bA < lblokK && bA->fcblknm == cblk_m; bA++ ) {
        while ( !is_block_inside_fblock( bA, bC ) ) {
            bC++;
            if ( bC >= lblokN ) {
                fprintf( stderr, "gpublok_gemmsp: no facing block\n" );
                return -1.;
            }
        }
        int32_t b = Ar.nblk;
        Ar.pref[b] = rows;
        Ar.off[b]  = bA->coefind - offA;
        Ar.ld[b]   = (int32_t)blok_rownbr( bA );
        Ar.toff[b] = bC->coefind - offC + ( bA->frownum - bC->frownum );
        Ar.tld[b]  = (int32_t)blok_rownbr( bC );
        rows += (int32_t)blok_rownbr( bA );
        flops += pa_fl_gemm( (double)blok_rownbr( bA ), (double)ncols, (double)K );
        Ar.nblk++;
        if ( Ar.nblk == PA_CALL_MAXBLK ) {
            Ar.pref[Ar.nblk] = rows;
            if ( pa_launch_gemm_call<T>( A, B, C, K, Ar, Br, 0, stream ) ) return -1.;
            memset( &Ar, 0, sizeof( Ar ) );
            rows = 0;
        }
    }
    if ( Ar.nblk > 0 ) {
        Ar.pref[Ar.nblk] = rows;
        if ( pa_launch_gemm_call<T>( A, B, C, K, Ar, Br, 0, stream ) ) return -1.;
    }
    return flops;
}

extern "C" pastix_fixdbl_t
gpublok_dgemmsp( pastix_trans_t trans, const SolverCblk *cblk, SolverCblk *fcblk, pastix_int_t blok_mk,
                 pastix_int_t blok_nk, pastix_int_t blok_mn, const void *A, const void *B, void *C,
                 const pastix_lr_t *lowrank, hipStream_t stream )
{
    (void)lowrank;
    return pa_gpublok_gemmsp<double>( trans, cblk, fcblk, blok_mk, blok_nk, blok_mn, (const double *)A,
                                      (const double *)B, (double *)C, stream );
}

extern "C" pastix_fixdbl_t
gpublok_sgemmsp( pastix_trans_t trans, const SolverCblk *cblk, SolverCblk *fcblk, pastix_int_t blok_mk,
                 pastix_int_t blok_nk, pastix_int_t blok_mn, const void *A, const void *B, void *C,
                 const pastix_lr_t *lowrank, hipStream_t stream )
{
    (void)lowrank;
    return pa_gpublok_gemmsp<float>( trans, cblk, fcblk, blok_mk, blok_nk, blok_mn, (const float *)A,
                                     (const float *)B, (float *)C, stream );
}

template <typename T>
static pastix_fixdbl_t
pa_gpublok_trsmsp( pastix_side_t side, pastix_uplo_t uplo, pastix_trans_t trans, pastix_diag_t diag,
                   const SolverCblk *cblk, pastix_int_t blok_m, const T *A, T *C, hipStream_t stream )
{
    int mode;
    if ( side == PastixRight && uplo == PastixLower && trans != PastixNoTrans && diag == PastixNonUnit ) mode = 0;
    else if ( side == PastixRight && uplo == PastixUpper && trans == PastixNoTrans && diag == PastixNonUnit ) mode = 1;
    else if ( side == PastixRight && uplo == PastixUpper && trans == PastixNoTrans && diag == PastixUnit ) mode = 2;
    else {
        fprintf( stderr, "gpublok_trsmsp: unsupported side/uplo/trans/diag combination\n" );
        return -1.;
    }
    if ( !( cblk->cblktype & CBLK_LAYOUT_2D ) || ( cblk->cblktype & CBLK_COMPRESSED ) ) {
        fprintf( stderr, "gpublok_trsmsp: the cblk must be stored in the 2D layout\n" );
        return -1.;
    }
    const int32_t N = (int32_t)cblk_colnbr( cblk );
    const SolverBlok *blok = cblk[0].fblokptr + blok_m, *lblok = cblk[1].fblokptr;
    const pastix_int_t offset = blok->coefind, cblk_m = blok->fcblknm;
    double flops = 0.;
    pa_call_rows R;
    memset( &R, 0, sizeof( R ) );
    int32_t rows = 0;
    size_t lds = pa_trsm_lds<T>( N );
    pa_api_lds_limits();
    for ( ; blok < lblok && blok->fcblknm == cblk_m; blok++ ) {
        int32_t b = R.nblk;
        R.pref[b] = rows;
        R.off[b]  = blok->coefind - offset;
        R.ld[b]   = (int32_t)blok_rownbr( blok );
        rows += (int32_t)blok_rownbr( blok );
        flops += pa_fl_trsm( (double)blok_rownbr( blok ), (double)N );
        R.nblk++;
        bool last = !( blok + 1 < lblok && blok[1].fcblknm == cblk_m );
        if ( R.nblk == PA_CALL_MAXBLK || last ) {
            R.pref[R.nblk] = rows;
            unsigned grid = ( rows + PA_TRSM_TR - 1 ) / PA_TRSM_TR;
            if ( mode == 0 )      k_trsm_call<T, 0><<<grid, 256, lds, stream>>>( C, A, N, R );
            else if ( mode == 1 ) k_trsm_call<T, 1><<<grid, 256, lds, stream>>>( C, A, N, R );
            else                  k_trsm_call<T, 2><<<grid, 256, lds, stream>>>( C, A, N, R );
            if ( hipGetLastError() != hipSuccess ) return -1.;
            /* offsets of the next chunk stay relative to the same C pointer */
            memset( &R, 0, sizeof( R ) );
            rows = 0;
        }
    }
    return flops;
}

extern "C" pastix_fixdbl_t
gpublok_dtrsmsp( pastix_side_t side, pastix_uplo_t uplo, pastix_trans_t trans, pastix_diag_t diag,
                 const SolverCblk *cblk, pastix_int_t blok_m, const void *A, void *C, const pastix_lr_t *lowrank,
                 hipStream_t stream )
{
    (void)lowrank;
    return pa_gpublok_trsmsp<double>( side, uplo, trans, diag, cblk, blok_m, (const double *)A, (double *)C, stream );
}

extern "C" pastix_fixdbl_t
gpublok_strsmsp( pastix_side_t side, pastix_uplo_t uplo, pastix_trans_t trans, pastix_diag_t diag,
                 const SolverCblk *cblk, pastix_int_t blok_m, const void *A, void *C, const pastix_lr_t *lowrank,
                 hipStream_t stream )
{
    (void)lowrank;
    return pa_gpublok_trsmsp<float>( side, uplo, trans, diag, cblk, blok_m, (const float *)A, (float *)C, stream );
}

/* FLOPS_DPOTRF / FLOPS_DGETRF of reference common/flops.h:91-92, 80-83 */
static double
pa_fl_potrf( double n )
{
    return n * ( ( ( 1. / 6. ) * n + 0.5 ) * n + ( 1. / 3. ) ) + n * ( ( ( 1. / 6. ) * n ) * n - ( 1. / 6. ) );
}
static double
pa_fl_getrf( double n )
{
    return ( 0.5 * n * ( n * ( n - ( 1. / 3. ) * n - 1. ) + n ) + ( 2. / 3. ) * n ) +
           ( 0.5 * n * ( n * ( n - ( 1. / 3. ) * n ) - n ) + ( 1. / 6. ) * n );
}

template <typename T, bool LU>
static pastix_fixdbl_t
pa_gpucblk_diag( const SolverCblk *cblk, void *L, void *U, double criterion, int32_t *d_npiv, hipStream_t stream )
{
    if ( !( cblk->cblktype & CBLK_LAYOUT_2D ) || ( cblk->cblktype & CBLK_COMPRESSED ) ) {
        fprintf( stderr, "gpucblk_potrfsp/getrfsp: the cblk must be stored in the 2D layout\n" );
        return -1.;
    }
    int32_t nc = (int32_t)cblk_colnbr( cblk );
    /* the panel, and for LU also the row panel U12 */
    size_t lds = (size_t)pa_diag_nb<T>::value * nc * sizeof( T ) * ( LU ? 2 : 1 );
    pa_api_lds_limits();
    /* negative diagonal terms are counted in the second word of the counter */
    if ( !LU && sizeof( T ) == 4 && nc > 64 && nc <= 16 * PA_DREG_NT )
        k_diag_potrf_reg_one<<<1, 1024, 0, stream>>>( (float *)L, nc, criterion, d_npiv, d_npiv + 1 );
    else
        k_diag_one<T, LU><<<1, 512, lds, stream>>>( (T *)L, (T *)U, nc, criterion, d_npiv, d_npiv + 1 );
    if ( hipGetLastError() != hipSuccess ) return -1.;
    return LU ? pa_fl_getrf( nc ) : pa_fl_potrf( nc );
}

extern "C" pastix_fixdbl_t
gpucblk_dpotrfsp( const SolverCblk *cblk, void *L, double criterion, int32_t *d_nbpivots, hipStream_t stream )
{
    return pa_gpucblk_diag<double, false>( cblk, L, nullptr, criterion, d_nbpivots, stream );
}
extern "C" pastix_fixdbl_t
gpucblk_spotrfsp( const SolverCblk *cblk, void *L, double criterion, int32_t *d_nbpivots, hipStream_t stream )
{
    return pa_gpucblk_diag<float, false>( cblk, L, nullptr, criterion, d_nbpivots, stream );
}
extern "C" pastix_fixdbl_t
gpucblk_dgetrfsp( const SolverCblk *cblk, void *L, void *U, double criterion, int32_t *d_nbpivots, hipStream_t stream )
{
    return pa_gpucblk_diag<double, true>( cblk, L, U, criterion, d_nbpivots, stream );
}
extern "C" pastix_fixdbl_t
gpucblk_sgetrfsp( const SolverCblk *cblk, void *L, void *U, double criterion, int32_t *d_nbpivots, hipStream_t stream )
{
    return pa_gpucblk_diag<float, true>( cblk, L, U, criterion, d_nbpivots, stream );
}
