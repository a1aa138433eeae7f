/*
 * refine.c -- iterative refinement in FP64 on the GPU: GMRES(m), conjugate
 * gradient and simple iterative refinement, preconditioned by the supernodal
 * solve (FP32 sweeps when the factorization is mixed precision).
 *
 * The algorithms restate the reference's refinement/z_refine_gmres.c:49-370,
 * z_refine_grad.c:49-175 and z_refine_pivot.c:48-157; the vector kernels
 * (SpMV, norms, dots, axpy) run on the device through pa_device.h in place of
 * the reference's function table (refinement/z_refine_functions.c:121-153).
 * Vectors are kept in the factorization numbering.
 */
#include <math.h>
#include "internal.h"
#include "../hip/pa_device.h"

typedef struct pa_refine_out_s {
    double resid_b;   /* last ||r||_2 / ||b||_2 estimate */
    pastix_int_t iters;
} pa_refine_out_t;

static int
pa_spsv( pastix_data_t *pd, double *d_x )
{
    return pa_dev_solve_perm( pd, (pastix_trans_t)pd->iparm[IPARM_TRANSPOSE_SOLVE], 1, d_x, pd->solvmatr->nodenbr );
}

static int
pa_spmv( pastix_data_t *pd, double alpha, const double *x, double beta, double *y )
{
    return pa_dev_spmv_perm( pd, (pastix_trans_t)pd->iparm[IPARM_TRANSPOSE_SOLVE], alpha, x, beta, y );
}

/* Right-preconditioned GMRES(im) (Saad, Iterative Methods, alg. 9.5), with the
 * Gram-Schmidt loop, Givens rotations and stopping rules of the reference. */
static int
pa_gmres( pastix_data_t *pd, const double *b, double *x, pa_refine_out_t *out )
{
    pastix_int_t n = pd->solvmatr->nodenbr;
    pastix_int_t im = pd->iparm[IPARM_GMRES_IM], im1 = im + 1;
    pastix_int_t itermax = pd->iparm[IPARM_ITERMAX];
    double eps = pd->dparm[DPARM_EPSILON_REFINEMENT];
    double *gmV, *gmW, *H, *cs, *sn, *G;
    double normb, normx, resid, resid_b = 0.;
    pastix_int_t i, j, iters = 0;
    int outflag = 1, rc = 0;

    gmV = pa_dev_wget( pd, (size_t)n * im1 * sizeof( double ) );
    gmW = pa_dev_wget( pd, (size_t)n * im * sizeof( double ) );
    H   = calloc( im * im1, sizeof( double ) );
    cs  = calloc( im, sizeof( double ) );
    sn  = calloc( im, sizeof( double ) );
    G   = calloc( im1, sizeof( double ) );
    if ( !gmV || !gmW ) { rc = PASTIX_ERR_OUTOFMEMORY; goto end; }

    normb = pa_dev_nrm2( pd, n, b );
    if ( normb == 0. ) normb = 1.;
    normx = pa_dev_nrm2( pd, n, x );

    while ( outflag ) {
        double *Vi = gmV;
        pa_dev_memcpy_d2d( pd, Vi, b, n * sizeof( double ) );
        if ( normx > 0. ) pa_spmv( pd, -1., x, 1., Vi );
        resid   = pa_dev_nrm2( pd, n, Vi );
        resid_b = resid / normb;
        if ( resid_b <= eps ) break;
        pa_dev_axpby( pd, n, 0., Vi, 1. / resid, Vi );
        G[0] = resid;
        for ( j = 1; j < im1; j++ ) G[j] = 0.;

        i = -1;
        for ( ;; ) {
            double *Hi, *Wi, norm, tmp;
            i++;
            Hi = H + i * im1;
            Wi = gmW + i * n;
            Vi = gmV + i * n;
            pa_dev_memcpy_d2d( pd, Wi, Vi, n * sizeof( double ) );
            if ( ( rc = pa_spsv( pd, Wi ) ) ) goto end;
            pa_spmv( pd, 1., Wi, 0., Vi + n );
            Vi += n;
            for ( j = 0; j <= i; j++ ) {
                Hi[j] = pa_dev_dot( pd, n, Vi, gmV + j * n );
                pa_dev_axpby( pd, n, -Hi[j], gmV + j * n, 1., Vi );
            }
            norm      = pa_dev_nrm2( pd, n, Vi );
            Hi[i + 1] = norm;
            if ( norm > 1e-50 ) pa_dev_axpby( pd, n, 0., Vi, 1. / norm, Vi );
            for ( j = 0; j < i; j++ ) {
                tmp       = Hi[j];
                Hi[j]     = cs[j] * tmp + sn[j] * Hi[j + 1];
                Hi[j + 1] = cs[j] * Hi[j + 1] - sn[j] * tmp;
            }
            tmp = sqrt( Hi[i] * Hi[i] + Hi[i + 1] * Hi[i + 1] );
            if ( fabs( tmp ) <= eps ) tmp = eps;
            cs[i] = Hi[i] / tmp;
            sn[i] = Hi[i + 1] / tmp;
            G[i + 1] = -sn[i] * G[i];
            G[i]     = cs[i] * G[i];
            Hi[i]    = cs[i] * Hi[i] + sn[i] * Hi[i + 1];
            resid    = fabs( G[i + 1] );
            resid_b  = resid / normb;
            iters++;
            if ( pd->iparm[IPARM_VERBOSE] > PastixVerboseNot ) {
                fprintf( stdout, "   GMRES iter %ld: ||r||/||b|| ~ %e\n", (long)iters, resid_b );
            }
            if ( ( i + 1 >= im ) || ( resid_b <= eps ) || ( iters >= itermax ) ) break;
        }
        /* y = H^{-1} G (upper triangular, leading dimension im1) */
        for ( j = i; j >= 0; j-- ) {
            pastix_int_t l;
            double s = G[j];
            for ( l = j + 1; l <= i; l++ ) s -= H[j + l * im1] * G[l];
            G[j] = s / H[j + j * im1];
        }
        pa_dev_gemv( pd, n, (int32_t)( i + 1 ), gmW, G, 1., x );
        normx = 1.;
        if ( iters >= itermax ) outflag = 0;
    }
end:
    pa_dev_sync( pd );
    pa_dev_wput( pd, gmV );
    pa_dev_wput( pd, gmW );
    free( H ); free( cs ); free( sn ); free( G );
    out->resid_b = resid_b;
    out->iters   = iters;
    return rc;
}

/* Preconditioned conjugate gradient (reference z_refine_grad.c:49-175). */
static int
pa_cg( pastix_data_t *pd, const double *b, double *x, pa_refine_out_t *out )
{
    pastix_int_t n = pd->solvmatr->nodenbr, itermax = pd->iparm[IPARM_ITERMAX], it = 0;
    double eps = pd->dparm[DPARM_EPSILON_REFINEMENT];
    double *r = pa_dev_wget( pd, n * sizeof( double ) ), *p = pa_dev_wget( pd, n * sizeof( double ) );
    double *z = pa_dev_wget( pd, n * sizeof( double ) ), *q = pa_dev_wget( pd, n * sizeof( double ) );
    double normb, normx, normr, resid_b, alpha, beta;
    int rc = 0;
    if ( !r || !p || !z || !q ) { rc = PASTIX_ERR_OUTOFMEMORY; goto end; }
    normb = pa_dev_nrm2( pd, n, b );
    if ( normb == 0. ) normb = 1.;
    normx = pa_dev_nrm2( pd, n, x );
    pa_dev_memcpy_d2d( pd, r, b, n * sizeof( double ) );
    if ( normx > 0. ) pa_spmv( pd, -1., x, 1., r );
    normr   = pa_dev_nrm2( pd, n, r );
    resid_b = normr / normb;
    pa_dev_memcpy_d2d( pd, z, r, n * sizeof( double ) );
    if ( ( rc = pa_spsv( pd, z ) ) ) goto end;
    pa_dev_memcpy_d2d( pd, p, z, n * sizeof( double ) );
    while ( resid_b > eps && it < itermax ) {
        it++;
        pa_spmv( pd, 1., p, 0., q );
        beta  = pa_dev_dot( pd, n, r, z );
        alpha = beta / pa_dev_dot( pd, n, q, p );
        pa_dev_axpby( pd, n, alpha, p, 1., x );
        pa_dev_axpby( pd, n, -alpha, q, 1., r );
        pa_dev_memcpy_d2d( pd, z, r, n * sizeof( double ) );
        if ( ( rc = pa_spsv( pd, z ) ) ) goto end;
        alpha = pa_dev_dot( pd, n, r, z );
        beta  = alpha / beta;
        pa_dev_axpby( pd, n, 1., z, beta, p );
        normr   = pa_dev_nrm2( pd, n, r );
        resid_b = normr / normb;
        if ( pd->iparm[IPARM_VERBOSE] > PastixVerboseNot ) {
            fprintf( stdout, "   CG iter %ld: ||r||/||b|| = %e\n", (long)it, resid_b );
        }
    }
    out->resid_b = resid_b;
    out->iters   = it;
end:
    pa_dev_sync( pd );
    pa_dev_wput( pd, r ); pa_dev_wput( pd, p ); pa_dev_wput( pd, z ); pa_dev_wput( pd, q );
    return rc;
}

/* Simple iterative refinement, stopped when the error no longer halves
 * (reference z_refine_pivot.c:48-157). */
static int
pa_sr( pastix_data_t *pd, const double *b, double *x, pa_refine_out_t *out )
{
    pastix_int_t n = pd->solvmatr->nodenbr, itermax = pd->iparm[IPARM_ITERMAX], iter = 0;
    double eps = pd->dparm[DPARM_EPSILON_REFINEMENT];
    double *r = pa_dev_wget( pd, n * sizeof( double ) );
    double normb, berr = 0., last_berr = 0.;
    int rc = 0;
    if ( !r ) return PASTIX_ERR_OUTOFMEMORY;
    normb = pa_dev_nrm2( pd, n, b );
    if ( normb == 0. ) normb = 1.;
    for ( ;; ) {
        pa_dev_memcpy_d2d( pd, r, b, n * sizeof( double ) );
        pa_spmv( pd, -1., x, 1., r );
        berr = pa_dev_nrm2( pd, n, r ) / normb;
        if ( iter == 0 ) last_berr = 3. * berr;
        if ( iter < itermax && berr > eps && berr <= last_berr / 2. ) {
            if ( ( rc = pa_spsv( pd, r ) ) ) break;
            pa_dev_axpby( pd, n, 1., r, 1., x );
            last_berr = berr;
        }
        else {
            break;
        }
        iter++;
    }
    pa_dev_sync( pd );
    pa_dev_wput( pd, r );
    out->resid_b = berr;
    out->iters   = iter;
    return rc;
}

/* Refine nrhs columns held on the device in the factorization numbering. */
static int
pa_refine_perm( pastix_data_t *pd, pastix_int_t nrhs, const double *b, pastix_int_t ldb, double *x,
                pastix_int_t ldx )
{
    pastix_int_t r;
    int rc = 0;
    if ( pd->dparm[DPARM_EPSILON_REFINEMENT] < 0. ) {
        pd->dparm[DPARM_EPSILON_REFINEMENT] = ( pd->iparm[IPARM_FLOAT] == SpmFloat ) ? 1e-6 : 1e-12;
    }
    pd->iparm[IPARM_NBITER] = 0;
    for ( r = 0; r < nrhs && !rc; r++ ) {
        pa_refine_out_t out = { 0., 0 };
        switch ( pd->iparm[IPARM_REFINEMENT] ) {
        case PastixRefineCG: rc = pa_cg( pd, b + r * ldb, x + r * ldx, &out ); break;
        case PastixRefineSR: rc = pa_sr( pd, b + r * ldb, x + r * ldx, &out ); break;
        default:             rc = pa_gmres( pd, b + r * ldb, x + r * ldx, &out ); break;
        }
        pd->iparm[IPARM_NBITER] = pa_imax( pd->iparm[IPARM_NBITER], out.iters );
        pd->dparm[DPARM_RELATIVE_ERROR] = out.resid_b;
    }
    return rc;
}

int
pastix_amd_refine_dev( pastix_data_t *pd, pastix_int_t nrhs, const void *dB, pastix_int_t ldb, void *dX,
                       pastix_int_t ldx )
{
    pastix_int_t n;
    double *bp, *xp, t0;
    int rc;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_task_refine: the factorization has to be done first" );
        return -PASTIX_ERR_BADPARAMETER;
    }
    n  = pd->solvmatr->nodenbr;
    t0 = pa_wtime();
    bp = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    xp = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !bp || !xp ) { pa_dev_wput( pd, bp ); pa_dev_wput( pd, xp ); return -PASTIX_ERR_OUTOFMEMORY; }
    pa_dev_apply_perm( pd, 1, (int32_t)nrhs, dB, ldb, bp, n );
    pa_dev_apply_perm( pd, 1, (int32_t)nrhs, dX, ldx, xp, n );
    rc = pa_refine_perm( pd, nrhs, bp, n, xp, n );
    pa_dev_apply_perm( pd, 0, (int32_t)nrhs, xp, n, dX, ldx );
    pa_dev_sync( pd );
    pa_dev_wput( pd, bp );
    pa_dev_wput( pd, xp );
    pd->dparm[DPARM_REFINE_TIME] = pa_wtime() - t0;
    if ( rc ) return -rc;
    pd->steps |= STEP_REFINE;
    return (int)pd->iparm[IPARM_NBITER];
}

int
pastix_task_refine( pastix_data_t *pd, pastix_int_t n, pastix_int_t nrhs, void *B, pastix_int_t ldb,
                    void *X, pastix_int_t ldx )
{
    double *dB, *dX;
    pastix_int_t j;
    int rc;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_task_refine: the factorization has to be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( n != pd->solvmatr->nodenbr || B == NULL || X == NULL ) return PASTIX_ERR_BADPARAMETER;
    dB = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    dX = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !dB || !dX ) { pa_dev_wput( pd, dB ); pa_dev_wput( pd, dX ); return PASTIX_ERR_OUTOFMEMORY; }
    for ( j = 0; j < nrhs; j++ ) {
        pa_dev_memcpy_h2d( pd, dB + j * n, (double *)B + j * ldb, n * sizeof( double ) );
        pa_dev_memcpy_h2d( pd, dX + j * n, (double *)X + j * ldx, n * sizeof( double ) );
    }
    rc = pastix_amd_refine_dev( pd, nrhs, dB, n, dX, n );
    for ( j = 0; rc >= 0 && j < nrhs; j++ ) {
        pa_dev_memcpy_d2h( pd, (double *)X + j * ldx, dX + j * n, n * sizeof( double ) );
    }
    pa_dev_wput( pd, dB );
    pa_dev_wput( pd, dX );
    return rc < 0 ? -rc : PASTIX_SUCCESS;
}

int
pastix_task_solve_and_refine( pastix_data_t *pd, pastix_int_t n, pastix_int_t nrhs, void *B,
                              pastix_int_t ldb, void *X, pastix_int_t ldx )
{
    pastix_int_t j;
    int rc;
    for ( j = 0; j < nrhs; j++ ) memcpy( (double *)X + j * ldx, (double *)B + j * ldb, n * sizeof( double ) );
    if ( ( rc = pastix_task_solve( pd, n, nrhs, X, ldx ) ) ) return rc;
    return pastix_task_refine( pd, n, nrhs, B, ldb, X, ldx );
}

int
pastix_amd_spmv_dev( pastix_data_t *pd, pastix_trans_t trans, double alpha, const void *dx, double beta,
                     void *dy )
{
    pastix_int_t n;
    double *xp, *yp;
    if ( pd == NULL || !( pd->steps & STEP_CSC2BCSC ) ) return PASTIX_ERR_BADPARAMETER;
    n  = pd->solvmatr->nodenbr;
    xp = pa_dev_wget( pd, n * sizeof( double ) );
    yp = pa_dev_wget( pd, n * sizeof( double ) );
    if ( !xp || !yp ) { pa_dev_wput( pd, xp ); pa_dev_wput( pd, yp ); return PASTIX_ERR_OUTOFMEMORY; }
    pa_dev_apply_perm( pd, 1, 1, dx, n, xp, n );
    pa_dev_apply_perm( pd, 1, 1, dy, n, yp, n );
    pa_dev_spmv_perm( pd, trans, alpha, xp, beta, yp );
    pa_dev_apply_perm( pd, 0, 1, yp, n, dy, n );
    pa_dev_sync( pd );
    pa_dev_wput( pd, xp );
    pa_dev_wput( pd, yp );
    return PASTIX_SUCCESS;
}
