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
    return pa_dev_solve_perm( pd, pd->refine_trans, 1, d_x, pd->solvmatr->nodenbr );
}

static int
pa_spmv( pastix_data_t *pd, double alpha, const double *x, double beta, double *y )
{
    return pa_dev_spmv_perm( pd, pd->refine_trans, alpha, x, beta, y );
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

/* Right-preconditioned BiCGSTAB with the steps and stopping rule of the
 * reference (refinement/z_refine_bicgstab.c:49-230): three preconditioner
 * applications per iteration, the third one for the M^{-1} t weighting of
 * omega. */
static int
pa_bicgstab( pastix_data_t *pd, const double *b, double *x, pa_refine_out_t *out )
{
    pastix_int_t n = pd->solvmatr->nodenbr, itermax = pd->iparm[IPARM_ITERMAX], it = 0;
    double eps = pd->dparm[DPARM_EPSILON_REFINEMENT];
    const size_t sz = (size_t)n * sizeof( double );
    double *r = pa_dev_wget( pd, sz ), *r2 = pa_dev_wget( pd, sz ), *p = pa_dev_wget( pd, sz );
    double *y = pa_dev_wget( pd, sz ), *v = pa_dev_wget( pd, sz ), *s = pa_dev_wget( pd, sz );
    double *z = pa_dev_wget( pd, sz ), *t = pa_dev_wget( pd, sz ), *g2 = pa_dev_wget( pd, sz );
    double normb, normx, normr, resid_b, alpha, beta, v1, v2, w;
    int rc = 0;
    if ( !r || !r2 || !p || !y || !v || !s || !z || !t || !g2 ) { rc = PASTIX_ERR_OUTOFMEMORY; goto end; }
    normb = pa_dev_nrm2( pd, n, b );
    if ( normb == 0. ) normb = 1.;
    normx = pa_dev_nrm2( pd, n, x );
    pa_dev_memcpy_d2d( pd, r, b, sz );
    if ( normx > 0. ) pa_spmv( pd, -1., x, 1., r );
    normr = pa_dev_nrm2( pd, n, r );
    pa_dev_memcpy_d2d( pd, r2, r, sz );
    pa_dev_memcpy_d2d( pd, p, r, sz );
    resid_b = normr / normb;
    while ( resid_b > eps && it < itermax ) {
        it++;
        /* y = M^{-1} p, v = A y */
        pa_dev_memcpy_d2d( pd, y, p, sz );
        if ( ( rc = pa_spsv( pd, y ) ) ) goto end;
        pa_spmv( pd, 1., y, 0., v );
        /* alpha = (r, r2) / (v, r2) */
        alpha = pa_dev_dot( pd, n, v, r2 );
        beta  = pa_dev_dot( pd, n, r, r2 );
        alpha = beta / alpha;
        /* s = r - alpha v, z = M^{-1} s, t = A z */
        pa_dev_memcpy_d2d( pd, s, r, sz );
        pa_dev_axpby( pd, n, -alpha, v, 1., s );
        pa_dev_memcpy_d2d( pd, z, s, sz );
        if ( ( rc = pa_spsv( pd, z ) ) ) goto end;
        pa_spmv( pd, 1., z, 0., t );
        /* w = (M^{-1} t, M^{-1} s) / (M^{-1} t, M^{-1} t) */
        pa_dev_memcpy_d2d( pd, g2, t, sz );
        if ( ( rc = pa_spsv( pd, g2 ) ) ) goto end;
        v1 = pa_dev_dot( pd, n, g2, z );
        v2 = pa_dev_dot( pd, n, g2, g2 );
        w  = v1 / v2;
        /* x += alpha y + w z; r = s - w t */
        pa_dev_axpby( pd, n, alpha, y, 1., x );
        pa_dev_axpby( pd, n, w, z, 1., x );
        pa_dev_memcpy_d2d( pd, r, s, sz );
        pa_dev_axpby( pd, n, -w, t, 1., r );
        /* beta = (r', r2) / (r, r2) * (alpha / w); p = r + beta (p - w v) */
        v1   = pa_dev_dot( pd, n, r, r2 );
        beta = v1 / beta * ( alpha / w );
        pa_dev_axpby( pd, n, -w, v, 1., p );
        pa_dev_axpby( pd, n, 1., r, beta, p );
        normr   = pa_dev_nrm2( pd, n, r );
        resid_b = normr / normb;
        if ( pd->iparm[IPARM_VERBOSE] > PastixVerboseNot ) {
            fprintf( stdout, "   BiCGSTAB iter %ld: ||r||/||b|| = %e\n", (long)it, resid_b );
        }
    }
    out->resid_b = resid_b;
    out->iters   = it;
end:
    pa_dev_sync( pd );
    pa_dev_wput( pd, r ); pa_dev_wput( pd, r2 ); pa_dev_wput( pd, p ); pa_dev_wput( pd, y ); pa_dev_wput( pd, v );
    pa_dev_wput( pd, s ); pa_dev_wput( pd, z ); pa_dev_wput( pd, t ); pa_dev_wput( pd, g2 );
    return rc;
}

/* Refine nrhs columns held on the device in the factorization numbering. */
static int
pa_refine_perm( pastix_data_t *pd, pastix_trans_t trans, pastix_int_t nrhs, const double *b, pastix_int_t ldb,
                double *x, pastix_int_t ldx )
{
    pastix_int_t r;
    int rc = 0;
    switch ( pd->iparm[IPARM_REFINEMENT] ) {
    case PastixRefineGMRES: case PastixRefineCG: case PastixRefineSR: case PastixRefineBiCGSTAB: break;
    default:
        pa_error( "refinement: unknown method %ld", (long)pd->iparm[IPARM_REFINEMENT] );
        return PASTIX_ERR_BADPARAMETER;
    }
    pd->refine_trans = trans;
    if ( pd->dparm[DPARM_EPSILON_REFINEMENT] < 0. ) {
        pd->dparm[DPARM_EPSILON_REFINEMENT] = ( pd->iparm[IPARM_FLOAT] == SpmFloat ) ? 1e-6 : 1e-12;
    }
    pd->iparm[IPARM_NBITER] = 0;
    for ( r = 0; r < nrhs && !rc; r++ ) {
        pa_refine_out_t out = { 0., 0 };
        switch ( pd->iparm[IPARM_REFINEMENT] ) {
        case PastixRefineCG:       rc = pa_cg( pd, b + r * ldb, x + r * ldx, &out ); break;
        case PastixRefineSR:       rc = pa_sr( pd, b + r * ldb, x + r * ldx, &out ); break;
        case PastixRefineBiCGSTAB: rc = pa_bicgstab( pd, b + r * ldb, x + r * ldx, &out ); break;
        default:                   rc = pa_gmres( pd, b + r * ldb, x + r * ldx, &out ); break;
        }
        pd->iparm[IPARM_NBITER] = pa_imax( pd->iparm[IPARM_NBITER], out.iters );
        pd->dparm[DPARM_RELATIVE_ERROR] = out.resid_b;
    }
    return rc;
}

int
pastix_amd_refine_dev_adv( pastix_data_t *pd, pastix_trans_t transA, pastix_int_t nrhs, const void *dB,
                           pastix_int_t ldb, void *dX, pastix_int_t ldx )
{
    pastix_int_t n;
    double *bp, *xp, t0;
    int rc;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_task_refine: the factorization has to be done first" );
        return -PASTIX_ERR_BADPARAMETER;
    }
    if ( transA == PastixConjTrans ) transA = PastixTrans;
    if ( transA != PastixNoTrans && transA != PastixTrans ) return -PASTIX_ERR_BADPARAMETER;
    n = pd->solvmatr->nodenbr;
    if ( nrhs < 1 || ldb < n || ldx < n || dB == NULL || dX == NULL ) return -PASTIX_ERR_BADPARAMETER;
    t0 = pa_wtime();
    bp = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    xp = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !bp || !xp ) { pa_dev_wput( pd, bp ); pa_dev_wput( pd, xp ); return -PASTIX_ERR_OUTOFMEMORY; }
    pa_dev_apply_perm( pd, 1, (int32_t)nrhs, dB, ldb, bp, n );
    pa_dev_apply_perm( pd, 1, (int32_t)nrhs, dX, ldx, xp, n );
    rc = pa_refine_perm( pd, transA, nrhs, bp, n, xp, n );
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
pastix_amd_refine_dev( pastix_data_t *pd, pastix_int_t nrhs, const void *dB, pastix_int_t ldb, void *dX,
                       pastix_int_t ldx )
{
    if ( pd == NULL ) return -PASTIX_ERR_BADPARAMETER;
    return pastix_amd_refine_dev_adv( pd, (pastix_trans_t)pd->iparm[IPARM_TRANSPOSE_SOLVE], nrhs, dB, ldb, dX, ldx );
}

/* copies between a host RHS of the caller's arithmetic and a device FP64 array */
static int
pa_vec_h2d( pastix_data_t *pd, int flt, const void *H, pastix_int_t ldh, pastix_int_t n, pastix_int_t nrhs, double *D )
{
    pastix_int_t i, j;
    int rc = 0;
    double *tmp = NULL;
    if ( flt == PastixFloat && !( tmp = malloc( (size_t)pa_imax( n, 1 ) * sizeof( double ) ) ) ) return PASTIX_ERR_ALLOC;
    for ( j = 0; j < nrhs && !rc; j++ ) {
        if ( flt == PastixFloat ) {
            for ( i = 0; i < n; i++ ) tmp[i] = ( (const float *)H )[j * ldh + i];
            rc = pa_dev_memcpy_h2d( pd, D + j * n, tmp, n * sizeof( double ) );
        }
        else {
            rc = pa_dev_memcpy_h2d( pd, D + j * n, (const double *)H + j * ldh, n * sizeof( double ) );
        }
    }
    free( tmp );
    return rc;
}

static int
pa_vec_d2h( pastix_data_t *pd, int flt, void *H, pastix_int_t ldh, pastix_int_t n, pastix_int_t nrhs, const double *D )
{
    pastix_int_t i, j;
    int rc = 0;
    double *tmp = NULL;
    if ( flt == PastixFloat && !( tmp = malloc( (size_t)pa_imax( n, 1 ) * sizeof( double ) ) ) ) return PASTIX_ERR_ALLOC;
    for ( j = 0; j < nrhs && !rc; j++ ) {
        if ( flt == PastixFloat ) {
            rc = pa_dev_memcpy_d2h( pd, tmp, D + j * n, n * sizeof( double ) );
            for ( i = 0; !rc && i < n; i++ ) ( (float *)H )[j * ldh + i] = (float)tmp[i];
        }
        else {
            rc = pa_dev_memcpy_d2h( pd, (double *)H + j * ldh, D + j * n, n * sizeof( double ) );
        }
    }
    free( tmp );
    return rc;
}

static int
pa_user_flt( const pastix_data_t *pd )
{
    return ( pd->iparm[IPARM_FLOAT] == SpmFloat ) ? PastixFloat : PastixDouble;
}

/* Refinement of permuted host vectors (reference pastix_subtask_refine,
 * refinement/pastix_task_refine.c:106-185): one right-hand side after the
 * other, IPARM_NBITER the largest iteration count. */
int
pastix_subtask_refine( pastix_data_t *pd, pastix_rhs_t Bp, pastix_rhs_t Xp )
{
    pastix_int_t n, nrhs;
    double *db, *dx, t0;
    int rc;
    if ( pd == NULL || Bp == NULL || Xp == NULL || Bp->b == NULL || Xp->b == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_subtask_refine: the factorization has to be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    n    = pd->solvmatr->nodenbr;
    nrhs = Bp->n;
    if ( Bp->m != n || Xp->m != n || Xp->n != nrhs || Bp->ld < n || Xp->ld < n || Bp->flttype != Xp->flttype ||
         ( Bp->flttype != PastixDouble && Bp->flttype != PastixFloat ) )
        return PASTIX_ERR_BADPARAMETER;
    if ( nrhs == 0 ) return PASTIX_SUCCESS;
    t0 = pa_wtime();
    db = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    dx = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !db || !dx ) { pa_dev_wput( pd, db ); pa_dev_wput( pd, dx ); return PASTIX_ERR_OUTOFMEMORY; }
    rc = pa_vec_h2d( pd, Bp->flttype, Bp->b, Bp->ld, n, nrhs, db );
    if ( !rc ) rc = pa_vec_h2d( pd, Xp->flttype, Xp->b, Xp->ld, n, nrhs, dx );
    if ( !rc ) rc = pa_refine_perm( pd, (pastix_trans_t)pd->iparm[IPARM_TRANSPOSE_SOLVE], nrhs, db, n, dx, n );
    if ( !rc ) rc = pa_vec_d2h( pd, Xp->flttype, Xp->b, Xp->ld, n, nrhs, dx );
    pa_dev_wput( pd, db );
    pa_dev_wput( pd, dx );
    pd->dparm[DPARM_REFINE_TIME] = pa_wtime() - t0;
    if ( !rc ) pd->steps |= STEP_REFINE;
    return rc;
}

int
pastix_task_refine( pastix_data_t *pd, pastix_int_t n, pastix_int_t nrhs, void *B, pastix_int_t ldb,
                    void *X, pastix_int_t ldx )
{
    double *dB, *dX;
    int rc, flt;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_task_refine: the factorization has to be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( n != pd->solvmatr->nodenbr || B == NULL || X == NULL || nrhs < 1 || ldb < n || ldx < n )
        return PASTIX_ERR_BADPARAMETER;
    flt = pa_user_flt( pd );
    dB  = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    dX  = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !dB || !dX ) { pa_dev_wput( pd, dB ); pa_dev_wput( pd, dX ); return PASTIX_ERR_OUTOFMEMORY; }
    rc = pa_vec_h2d( pd, flt, B, ldb, n, nrhs, dB );
    if ( !rc ) rc = pa_vec_h2d( pd, flt, X, ldx, n, nrhs, dX );
    if ( !rc ) {
        rc = pastix_amd_refine_dev( pd, nrhs, dB, n, dX, n );
        rc = rc < 0 ? -rc : PASTIX_SUCCESS;
    }
    if ( !rc ) rc = pa_vec_d2h( pd, flt, X, ldx, n, nrhs, dX );
    pa_dev_wput( pd, dB );
    pa_dev_wput( pd, dX );
    return rc;
}

/* Solve then refine without the intermediate permutations (reference
 * refinement/pastix_task_refine.c:359-450).  X receives the solution; its
 * entry values are not read (the reference expects a copy of B there). */
int
pastix_task_solve_and_refine( pastix_data_t *pd, pastix_int_t n, pastix_int_t nrhs, void *B,
                              pastix_int_t ldb, void *X, pastix_int_t ldx )
{
    double *dB, *dX;
    int rc, flt;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_task_solve_and_refine: the factorization has to be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( n != pd->solvmatr->nodenbr || B == NULL || X == NULL || nrhs < 1 || ldb < n || ldx < n )
        return PASTIX_ERR_BADPARAMETER;
    flt = pa_user_flt( pd );
    dB  = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    dX  = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !dB || !dX ) { pa_dev_wput( pd, dB ); pa_dev_wput( pd, dX ); return PASTIX_ERR_OUTOFMEMORY; }
    rc = pa_vec_h2d( pd, flt, B, ldb, n, nrhs, dB );
    if ( !rc ) rc = pa_dev_memcpy_d2d( pd, dX, dB, (size_t)n * nrhs * sizeof( double ) );
    if ( !rc ) rc = pastix_amd_solve_dev( pd, (pastix_trans_t)pd->iparm[IPARM_TRANSPOSE_SOLVE], nrhs, dX, n );
    if ( !rc ) {
        rc = pastix_amd_refine_dev( pd, nrhs, dB, n, dX, n );
        rc = rc < 0 ? -rc : PASTIX_SUCCESS;
    }
    if ( !rc ) rc = pa_vec_d2h( pd, flt, X, ldx, n, nrhs, dX );
    pa_dev_wput( pd, dB );
    pa_dev_wput( pd, dX );
    return rc;
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
