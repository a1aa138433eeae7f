/*
 * oracle.c -- CPU restatement of PaStiX's sequential full-rank numerical path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker of the GPU engine: only
 * the tests, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product never calls it and has no CPU fallback.
 *
 * It restates, in plain C with its own dense loops (no BLAS library):
 *   - coeftab fill from the permuted CSC     reference kernels/cpucblk_zinit.c:395-480
 *   - sequential Cholesky driver             reference sopalin/sequential_zpotrf.c:51-92
 *   - sequential LU driver                   reference sopalin/sequential_zgetrf.c:51
 *   - diagonal POTRF, 64-blocked, static pivoting
 *                                            reference kernels/core_zpotrfsp.c:68-190
 *   - diagonal GETRF, 64-blocked, static pivoting, geadd/getmo
 *                                            reference kernels/core_zgetrfsp.c:65-272
 *   - sequential LDL^t driver, diagonal SYTRF (64-blocked, static
 *     pivoting), panel TRSM against D L^t and the scaled copy L D
 *                                            reference sopalin/sequential_zsytrf.c,
 *                                            kernels/core_zsytrfsp.c:28-380,
 *                                            kernels/core_zscalo.c
 *   - panel TRSM on 2D tiles                 reference kernels/core_ztrsmsp.c:149-180
 *   - 2D scatter GEMM                        reference kernels/core_zgemmsp.c:404-481
 *   - supernodal forward/backward solve      reference kernels/solve_ztrsmsp.c:302-661
 *                                            and sopalin/pastix_task_solve.c:543-649
 * It is compiled twice: FP64 (oracle_d*) and, with ORACLE_SINGLE, FP32 (oracle_s*).
 *
 * For the CPU-baseline leg only, the caller may route the dense GEMM/TRSM
 * through a BLAS library (set_blas: Fortran-convention pointers, as the
 * reference's per-task BLAS calls) and run the level-parallel driver
 * potrf_par on several threads (OpenMP; updates into a shared facing cblk
 * serialised by a per-cblk lock, as the reference's pastix_cblk_lock).  The
 * tests use neither.
 *
 * The structure is passed as the arrays exported by pastix_amd_get_structure:
 *   cblk[5*k]: fcolnum, lcolnum, first block, stride, coefficient offset
 *   blok[5*b]: frownum, lrownum, lcblknm, fcblknm, coefind
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <pthread.h>
#include <sched.h>

#if defined( ORACLE_SINGLE )
#define REAL float
#define PFX( name ) oracle_s##name
#else
#define REAL double
#define PFX( name ) oracle_d##name
#endif

typedef int64_t oint;

#define C_FCOL( k ) cblk[5 * ( k ) + 0]
#define C_LCOL( k ) cblk[5 * ( k ) + 1]
#define C_FBLK( k ) cblk[5 * ( k ) + 2]
#define C_STRD( k ) cblk[5 * ( k ) + 3]
#define C_COEF( k ) cblk[5 * ( k ) + 4]
#define B_FROW( b ) blok[5 * ( b ) + 0]
#define B_LROW( b ) blok[5 * ( b ) + 1]
#define B_LCBK( b ) blok[5 * ( b ) + 2]
#define B_FCBK( b ) blok[5 * ( b ) + 3]
#define B_COEF( b ) blok[5 * ( b ) + 4]
#define B_NROW( b ) ( B_LROW( b ) - B_FROW( b ) + 1 )
#define C_NCOL( k ) ( C_LCOL( k ) - C_FCOL( k ) + 1 )

#define MAXSIZEOFBLOCKS 64

/* ---- optional BLAS (CPU-baseline leg only) ------------------------------ */
typedef void ( *blas_gemm_t )( const char *, const char *, const int *, const int *, const int *, const REAL *,
                               const REAL *, const int *, const REAL *, const int *, const REAL *, REAL *,
                               const int * );
typedef void ( *blas_trsm_t )( const char *, const char *, const char *, const char *, const int *, const int *,
                               const REAL *, const REAL *, const int *, REAL *, const int * );
static blas_gemm_t blas_gemm;
static blas_trsm_t blas_trsm;

void
PFX( set_blas )( void *gemm, void *trsm )
{
    blas_gemm = (blas_gemm_t)gemm;
    blas_trsm = (blas_trsm_t)trsm;
}

/* ---- dense helpers (column major) --------------------------------------- */

/* C(m x n) -= A(m x k) * op(B),  op(B) = B^T with B (n x k) */
static void
gemm_nt_minus( oint m, oint n, oint k, const REAL *A, oint lda, const REAL *B, oint ldb, REAL *C, oint ldc )
{
    oint i, j, l;
    if ( blas_gemm && m > 0 && n > 0 && k > 0 ) {
        const int  im = (int)m, in = (int)n, ik = (int)k, ia = (int)lda, ib = (int)ldb, ic = (int)ldc;
        const REAL mone = (REAL)-1, one = (REAL)1;
        blas_gemm( "N", "T", &im, &in, &ik, &mone, A, &ia, B, &ib, &one, C, &ic );
        return;
    }
    for ( j = 0; j < n; j++ ) {
        REAL *c = C + j * ldc;
        for ( l = 0; l < k; l++ ) {
            REAL bjl = B[j + l * ldb];
            const REAL *a = A + l * lda;
            if ( bjl == (REAL)0 ) continue;
            for ( i = 0; i < m; i++ ) c[i] -= a[i] * bjl;
        }
    }
}

/* C(m x n) -= A(m x k) * B(k x n) */
static void
gemm_nn_minus( oint m, oint n, oint k, const REAL *A, oint lda, const REAL *B, oint ldb, REAL *C, oint ldc )
{
    oint i, j, l;
    if ( blas_gemm && m > 0 && n > 0 && k > 0 ) {
        const int  im = (int)m, in = (int)n, ik = (int)k, ia = (int)lda, ib = (int)ldb, ic = (int)ldc;
        const REAL mone = (REAL)-1, one = (REAL)1;
        blas_gemm( "N", "N", &im, &in, &ik, &mone, A, &ia, B, &ib, &one, C, &ic );
        return;
    }
    for ( j = 0; j < n; j++ ) {
        REAL *c = C + j * ldc;
        for ( l = 0; l < k; l++ ) {
            REAL blj = B[l + j * ldb];
            const REAL *a = A + l * lda;
            if ( blj == (REAL)0 ) continue;
            for ( i = 0; i < m; i++ ) c[i] -= a[i] * blj;
        }
    }
}

/* X(m x n) <- X * L^{-T}, L lower non-unit (Right, Lower, Trans, NonUnit) */
static void
trsm_rlt( oint m, oint n, const REAL *L, oint lda, REAL *X, oint ldx )
{
    oint i, j, l;
    if ( blas_trsm && m > 0 && n > 0 ) {
        const int  im = (int)m, in = (int)n, ia = (int)lda, ix = (int)ldx;
        const REAL one = (REAL)1;
        blas_trsm( "R", "L", "T", "N", &im, &in, &one, L, &ia, X, &ix );
        return;
    }
    for ( j = 0; j < n; j++ ) {
        REAL *xj = X + j * ldx;
        for ( l = 0; l < j; l++ ) {
            REAL c = L[j + l * lda];
            const REAL *xl = X + l * ldx;
            for ( i = 0; i < m; i++ ) xj[i] -= xl[i] * c;
        }
        {
            REAL d = L[j + j * lda];
            for ( i = 0; i < m; i++ ) xj[i] /= d;
        }
    }
}

/* X(m x n) <- X * U^{-1}, U upper (Right, Upper, NoTrans, unit or not) */
static void
trsm_run( oint m, oint n, const REAL *U, oint lda, int unit, REAL *X, oint ldx )
{
    oint i, j, l;
    if ( blas_trsm && m > 0 && n > 0 ) {
        const int  im = (int)m, in = (int)n, ia = (int)lda, ix = (int)ldx;
        const REAL one = (REAL)1;
        blas_trsm( "R", "U", "N", unit ? "U" : "N", &im, &in, &one, U, &ia, X, &ix );
        return;
    }
    for ( j = 0; j < n; j++ ) {
        REAL *xj = X + j * ldx;
        for ( l = 0; l < j; l++ ) {
            REAL c = U[l + j * lda];
            const REAL *xl = X + l * ldx;
            for ( i = 0; i < m; i++ ) xj[i] -= xl[i] * c;
        }
        if ( !unit ) {
            REAL d = U[j + j * lda];
            for ( i = 0; i < m; i++ ) xj[i] /= d;
        }
    }
}

/* X(m x n) <- L^{-1} X, L lower unit (Left, Lower, NoTrans, Unit) */
static void
trsm_llnu( oint m, oint n, const REAL *L, oint lda, REAL *X, oint ldx )
{
    oint i, j, l;
    if ( blas_trsm && m > 0 && n > 0 ) {
        const int  im = (int)m, in = (int)n, ia = (int)lda, ix = (int)ldx;
        const REAL one = (REAL)1;
        blas_trsm( "L", "L", "N", "U", &im, &in, &one, L, &ia, X, &ix );
        return;
    }
    for ( j = 0; j < n; j++ ) {
        REAL *x = X + j * ldx;
        for ( l = 0; l < m; l++ ) {
            REAL v = x[l];
            for ( i = l + 1; i < m; i++ ) x[i] -= L[i + l * lda] * v;
        }
    }
}

/* ---- diagonal factorizations ------------------------------------------- */

/* unblocked Cholesky of an n x n block with static pivoting
 * (reference kernels/core_zpotrfsp.c:68-107) */
static void
potf2sp( oint n, REAL *A, oint lda, oint *npiv, oint *nneg, double crit )
{
    oint k, i, j;
    for ( k = 0; k < n; k++ ) {
        REAL *akk = A + k * ( lda + 1 );
        REAL  alpha;
        if ( fabs( (double)*akk ) < crit ) {
            *akk = (REAL)crit;
            ( *npiv )++;
        }
        if ( *akk < (REAL)0 ) ( *nneg )++;
        *akk  = (REAL)sqrt( (double)*akk );
        alpha = (REAL)1 / *akk;
        for ( i = k + 1; i < n; i++ ) akk[i - k] *= alpha;
        /* syr: A(k+1:n, k+1:n) -= x x^T (lower) */
        for ( j = k + 1; j < n; j++ ) {
            REAL xj = A[j + k * lda];
            for ( i = j; i < n; i++ ) A[i + j * lda] -= A[i + k * lda] * xj;
        }
    }
}

/* reference kernels/core_zpotrfsp.c:144-190 */
static void
potrfsp( oint n, REAL *A, oint lda, oint *npiv, oint *nneg, double crit )
{
    oint k, blocknbr = ( n + MAXSIZEOFBLOCKS - 1 ) / MAXSIZEOFBLOCKS;
    for ( k = 0; k < blocknbr; k++ ) {
        oint bs = n - k * MAXSIZEOFBLOCKS < MAXSIZEOFBLOCKS ? n - k * MAXSIZEOFBLOCKS : MAXSIZEOFBLOCKS;
        REAL *tmp = A + ( k * MAXSIZEOFBLOCKS ) * ( lda + 1 );
        potf2sp( bs, tmp, lda, npiv, nneg, crit );
        if ( k * MAXSIZEOFBLOCKS + bs < n ) {
            REAL *tmp1 = tmp + bs, *tmp2 = tmp1 + bs * lda;
            oint  ms = n - ( k * MAXSIZEOFBLOCKS + bs ), i, j, l;
            trsm_rlt( ms, bs, tmp, lda, tmp1, lda );
            /* syrk lower: tmp2 -= tmp1 tmp1^T */
            if ( blas_gemm ) {
                gemm_nt_minus( ms, ms, bs, tmp1, lda, tmp1, lda, tmp2, lda );
            }
            else {
                for ( j = 0; j < ms; j++ ) {
                    for ( l = 0; l < bs; l++ ) {
                        REAL b = tmp1[j + l * lda];
                        for ( i = j; i < ms; i++ ) tmp2[i + j * lda] -= tmp1[i + l * lda] * b;
                    }
                }
            }
        }
    }
}

/* unblocked LU of an m x n panel with static pivoting
 * (reference kernels/core_zgetrfsp.c:65-105) */
static void
getf2sp( oint m, oint n, REAL *A, oint lda, oint *npiv, double crit )
{
    oint k, i, j, minmn = m < n ? m : n;
    for ( k = 0; k < minmn; k++ ) {
        REAL *akk = A + k * ( lda + 1 ), alpha;
        if ( fabs( (double)*akk ) < crit ) {
            *akk = ( *akk < (REAL)0 ) ? (REAL)( -crit ) : (REAL)crit;
            ( *npiv )++;
        }
        alpha = (REAL)1 / *akk;
        for ( i = k + 1; i < m; i++ ) A[i + k * lda] *= alpha;
        for ( j = k + 1; j < n; j++ ) {
            REAL ukj = A[k + j * lda];
            for ( i = k + 1; i < m; i++ ) A[i + j * lda] -= A[i + k * lda] * ukj;
        }
    }
}

/* reference kernels/core_zgetrfsp.c:138-185 */
static void
getrfsp( oint n, REAL *A, oint lda, oint *npiv, double crit )
{
    oint k, blocknbr = ( n + MAXSIZEOFBLOCKS - 1 ) / MAXSIZEOFBLOCKS;
    REAL *Akk = A;
    for ( k = 0; k < blocknbr; k++ ) {
        oint tempm = n - k * MAXSIZEOFBLOCKS;
        oint bs    = tempm < MAXSIZEOFBLOCKS ? tempm : MAXSIZEOFBLOCKS;
        REAL *Lik = Akk + bs, *Ukj = Akk + bs * lda, *Aij = Ukj + bs;
        oint ms;
        getf2sp( tempm, bs, Akk, lda, npiv, crit );
        ms = tempm - bs;
        if ( ms > 0 ) {
            trsm_llnu( bs, ms, Akk, lda, Ukj, lda );
            gemm_nn_minus( ms, ms, bs, Lik, lda, Ukj, lda, Aij, lda );
        }
        Akk += bs * ( lda + 1 );
    }
}

/* X(m x n) <- X * L^{-T}, L lower unit (Right, Lower, Trans, Unit) */
static void
trsm_rltu( oint m, oint n, const REAL *L, oint lda, REAL *X, oint ldx )
{
    oint i, j, l;
    if ( blas_trsm && m > 0 && n > 0 ) {
        const int  im = (int)m, in = (int)n, ia = (int)lda, ix = (int)ldx;
        const REAL one = (REAL)1;
        blas_trsm( "R", "L", "T", "U", &im, &in, &one, L, &ia, X, &ix );
        return;
    }
    for ( j = 0; j < n; j++ ) {
        REAL *xj = X + j * ldx;
        for ( l = 0; l < j; l++ ) {
            REAL c = L[j + l * lda];
            const REAL *xl = X + l * ldx;
            for ( i = 0; i < m; i++ ) xj[i] -= xl[i] * c;
        }
    }
}

/* unblocked LDL^t of an n x n block with static pivoting (reference
 * kernels/core_zsytrfsp.c:68-120): the pivot keeps its sign when replaced;
 * the unscaled column is copied into the row (D L^t) before the column is
 * scaled by 1/d_k, then the trailing lower part takes -d_k l l^T */
static void
sytf2sp( oint n, REAL *A, oint lda, oint *npiv, double crit )
{
    oint k, i, j;
    for ( k = 0; k < n; k++ ) {
        REAL *akk = A + k * ( lda + 1 ), alpha, d;
        oint  m = n - k - 1;
        if ( fabs( (double)*akk ) < crit ) {
            *akk = ( *akk < (REAL)0 ) ? (REAL)( -crit ) : (REAL)crit;
            ( *npiv )++;
        }
        d     = *akk;
        alpha = (REAL)1 / d;
        for ( i = 1; i <= m; i++ ) akk[i * lda] = akk[i];
        for ( i = 1; i <= m; i++ ) akk[i] *= alpha;
        for ( j = 1; j <= m; j++ ) {
            REAL t = -d * akk[j];
            for ( i = j; i <= m; i++ ) akk[i + j * lda] += akk[i] * t;
        }
    }
}

/* reference kernels/core_zsytrfsp.c:150-215: 64-column blocks; below each
 * factored block, X L_kk^{-T} (unit), then row copy and 1/d scaling column
 * by column, and the trailing square -= (L D) L^T */
static void
sytrfsp( oint n, REAL *A, oint lda, oint *npiv, double crit )
{
    oint k, blocknbr = ( n + MAXSIZEOFBLOCKS - 1 ) / MAXSIZEOFBLOCKS;
    for ( k = 0; k < blocknbr; k++ ) {
        oint bs = n - k * MAXSIZEOFBLOCKS < MAXSIZEOFBLOCKS ? n - k * MAXSIZEOFBLOCKS : MAXSIZEOFBLOCKS;
        REAL *Akk = A + ( k * MAXSIZEOFBLOCKS ) * ( lda + 1 );
        REAL *Amk = Akk + bs, *Akm = Akk + bs * lda, *Amm = Amk + bs * lda;
        sytf2sp( bs, Akk, lda, npiv, crit );
        if ( k * MAXSIZEOFBLOCKS + bs < n ) {
            oint ms = n - ( k * MAXSIZEOFBLOCKS + bs ), col, i;
            trsm_rltu( ms, bs, Akk, lda, Amk, lda );
            for ( col = 0; col < bs; col++ ) {
                REAL alpha = (REAL)1 / Akk[col * ( lda + 1 )];
                for ( i = 0; i < ms; i++ ) Akm[col + i * lda] = Amk[i + col * lda];
                for ( i = 0; i < ms; i++ ) Amk[i + col * lda] *= alpha;
            }
            gemm_nn_minus( ms, ms, bs, Amk, lda, Akm, lda, Amm, lda );
        }
    }
}

/* ---- supernodal kernels ------------------------------------------------- */

/* 2D scatter GEMM of block `bj` of cblk k into its facing cblk
 * (reference kernels/core_zgemmsp.c:404-481); shift 1 on the U side */
static void
gemmsp_2d2d( const oint *cblk, const oint *blok, oint k, oint bj, int shift, const REAL *A, const REAL *B,
             REAL *C )
{
    oint f = B_FCBK( bj ), fb = C_FBLK( f ), lb = C_FBLK( k + 1 ), i;
    oint K = C_NCOL( k ), N = B_NROW( bj );
    const REAL *Bj = B + C_COEF( k ) + B_COEF( bj );
    REAL *Cf = C + C_COEF( f );
    for ( i = bj + shift; i < lb; i++ ) {
        oint M = B_NROW( i ), ldc;
        while ( !( B_FROW( i ) >= B_FROW( fb ) && B_LROW( i ) <= B_LROW( fb ) ) ) fb++;
        ldc = B_NROW( fb );
        gemm_nt_minus( M, N, K, A + C_COEF( k ) + B_COEF( i ), M, Bj, N,
                       Cf + B_COEF( fb ) + ( B_FROW( i ) - B_FROW( fb ) ) + ( B_FROW( bj ) - C_FCOL( f ) ) * ldc,
                       ldc );
    }
}

/* coeftab fill (reference kernels/cpucblk_zinit.c:395-480): lower part of the
 * permuted CSC into L, strictly lower part of the transpose into U */
int
PFX( fill )( oint cblknbr, const oint *cblk, const oint *blok, const oint *colptr, const oint *rowtab,
             const double *Lv, const double *Uv, REAL *L, REAL *U )
{
    oint k, c, e;
    for ( k = 0; k < cblknbr; k++ ) {
        oint nc = C_NCOL( k ), lb = C_FBLK( k + 1 );
        for ( c = C_FCOL( k ); c <= C_LCOL( k ); c++ ) {
            oint b = C_FBLK( k );
            for ( e = colptr[c]; e < colptr[c + 1]; e++ ) {
                oint r = rowtab[e], idx;
                if ( r < c ) continue;
                while ( b < lb && ( B_LROW( b ) < r || B_FROW( b ) > r ) ) b++;
                if ( b >= lb ) continue;
                idx = C_COEF( k ) + B_COEF( b ) + ( r - B_FROW( b ) ) + ( c - C_FCOL( k ) ) * B_NROW( b );
                L[idx] = (REAL)Lv[e];
                if ( U && r > c ) U[idx] = (REAL)Uv[e];
            }
        }
        (void)nc;
    }
    return 0;
}

/* sequential LL^t (reference sopalin/sequential_zpotrf.c:51-92 with
 * cpucblk_zpotrfsp1d, kernels/core_zpotrfsp.c:220-367) */
int
PFX( potrf )( oint cblknbr, const oint *cblk, const oint *blok, REAL *L, double crit, oint *npiv, oint *nneg )
{
    oint k, b;
    *npiv = 0;
    *nneg = 0;
    for ( k = 0; k < cblknbr; k++ ) {
        oint nc = C_NCOL( k );
        REAL *D = L + C_COEF( k );
        potrfsp( nc, D, nc, npiv, nneg, crit );
        for ( b = C_FBLK( k ) + 1; b < C_FBLK( k + 1 ); b++ ) {
            trsm_rlt( B_NROW( b ), nc, D, nc, L + C_COEF( k ) + B_COEF( b ), B_NROW( b ) );
        }
        for ( b = C_FBLK( k ) + 1; b < C_FBLK( k + 1 ); b++ ) {
            gemmsp_2d2d( cblk, blok, k, b, 0, L, L, L );
        }
    }
    return 0;
}

/* Level-parallel LL^t on nthreads threads (CPU-baseline leg): the cblks of
 * one elimination-tree level are factored concurrently (diagonal + panel
 * TRSM), then all (cblk, block) scatter-GEMMs of the level run concurrently,
 * each under the lock of its facing cblk.  Same kernels and results (up to
 * the order of the updates) as the sequential driver. */
int
PFX( potrf_par )( oint cblknbr, const oint *cblk, const oint *blok, REAL *L, double crit, oint *npiv, oint *nneg,
                  int nthreads )
{
    oint       k, l, nlev = 0, *lev, *order, *lptr, *pairs, npairs = 0;
    omp_lock_t *locks;
    oint        piv = 0, neg = 0;
    lev   = calloc( cblknbr + 1, sizeof( oint ) );
    order = malloc( ( cblknbr + 1 ) * sizeof( oint ) );
    locks = malloc( ( cblknbr + 1 ) * sizeof( omp_lock_t ) );
    /* height of each cblk in the tree (children have smaller indices) */
    for ( k = 0; k < cblknbr; k++ ) {
        omp_init_lock( locks + k );
        if ( C_FBLK( k ) + 1 < C_FBLK( k + 1 ) ) {
            oint f = B_FCBK( C_FBLK( k ) + 1 );
            if ( lev[f] < lev[k] + 1 ) lev[f] = lev[k] + 1;
        }
        if ( lev[k] + 1 > nlev ) nlev = lev[k] + 1;
    }
    lptr = calloc( nlev + 2, sizeof( oint ) );
    for ( k = 0; k < cblknbr; k++ ) lptr[lev[k] + 1]++;
    for ( l = 0; l < nlev; l++ ) lptr[l + 1] += lptr[l];
    {
        oint *pos = malloc( ( nlev + 1 ) * sizeof( oint ) );
        memcpy( pos, lptr, ( nlev + 1 ) * sizeof( oint ) );
        for ( k = 0; k < cblknbr; k++ ) order[pos[lev[k]]++] = k;
        free( pos );
    }
    for ( k = 0; k < cblknbr; k++ ) npairs += C_FBLK( k + 1 ) - C_FBLK( k ) - 1;
    pairs = malloc( ( npairs + 1 ) * 2 * sizeof( oint ) );
    for ( l = 0; l < nlev; l++ ) {
        oint b0 = lptr[l], b1 = lptr[l + 1], np = 0, i;
#pragma omp parallel for num_threads( nthreads ) schedule( dynamic, 1 ) reduction( + : piv, neg )
        for ( i = b0; i < b1; i++ ) {
            oint kk = order[i], b, nc = C_NCOL( kk ), p = 0, q = 0;
            REAL *D = L + C_COEF( kk );
            potrfsp( nc, D, nc, &p, &q, crit );
            piv += p;
            neg += q;
            for ( b = C_FBLK( kk ) + 1; b < C_FBLK( kk + 1 ); b++ )
                trsm_rlt( B_NROW( b ), nc, D, nc, L + C_COEF( kk ) + B_COEF( b ), B_NROW( b ) );
        }
        for ( i = b0; i < b1; i++ ) {
            oint kk = order[i], b;
            for ( b = C_FBLK( kk ) + 1; b < C_FBLK( kk + 1 ); b++ ) {
                pairs[2 * np]     = kk;
                pairs[2 * np + 1] = b;
                np++;
            }
        }
#pragma omp parallel for num_threads( nthreads ) schedule( dynamic, 1 )
        for ( i = 0; i < np; i++ ) {
            oint kk = pairs[2 * i], b = pairs[2 * i + 1], f = B_FCBK( b );
            omp_set_lock( locks + f );
            gemmsp_2d2d( cblk, blok, kk, b, 0, L, L, L );
            omp_unset_lock( locks + f );
        }
    }
    for ( k = 0; k < cblknbr; k++ ) omp_destroy_lock( locks + k );
    *npiv = piv;
    *nneg = neg;
    free( lev ); free( order ); free( locks ); free( lptr ); free( pairs );
    return 0;
}

/* sequential LU with static pivoting (reference sopalin/sequential_zgetrf.c,
 * cpucblk_zgetrfsp1d, kernels/core_zgetrfsp.c:217-395) */
int
PFX( getrf )( oint cblknbr, const oint *cblk, const oint *blok, REAL *L, REAL *U, double crit, oint *npiv )
{
    oint k, b, i, j;
    *npiv = 0;
    for ( k = 0; k < cblknbr; k++ ) {
        oint nc = C_NCOL( k ), lb = C_FBLK( k + 1 );
        REAL *D = L + C_COEF( k ), *Du = U + C_COEF( k );
        /* geadd(Trans): D += Du^T */
        for ( j = 0; j < nc; j++ )
            for ( i = 0; i < nc; i++ ) D[i + j * nc] += Du[j + i * nc];
        getrfsp( nc, D, nc, npiv, crit );
        /* getmo: Du = D^T */
        for ( j = 0; j < nc; j++ )
            for ( i = 0; i < nc; i++ ) Du[j + i * nc] = D[i + j * nc];
        for ( b = C_FBLK( k ) + 1; b < lb; b++ ) {
            trsm_run( B_NROW( b ), nc, D, nc, 0, L + C_COEF( k ) + B_COEF( b ), B_NROW( b ) );
            trsm_run( B_NROW( b ), nc, Du, nc, 1, U + C_COEF( k ) + B_COEF( b ), B_NROW( b ) );
        }
        for ( b = C_FBLK( k ) + 1; b < lb; b++ ) {
            gemmsp_2d2d( cblk, blok, k, b, 0, L, U, L );
            if ( b + 1 < lb ) gemmsp_2d2d( cblk, blok, k, b, 1, U, L, U );
        }
    }
    return 0;
}

/* sequential LDL^t with static pivoting (reference
 * sopalin/sequential_zsytrf.c with cpucblk_zsytrfsp1d,
 * kernels/core_zsytrfsp.c:296-380): diagonal SYTRF, the panel solved against
 * the D L^t stored in the upper part of the diagonal block, its copy L D
 * into DLt (cpucblk_zscalo), and the updates C -= L_i (L D)_j^T */
int
PFX( sytrf )( oint cblknbr, const oint *cblk, const oint *blok, REAL *L, REAL *DLt, double crit, oint *npiv )
{
    oint k, b, i, j;
    *npiv = 0;
    for ( k = 0; k < cblknbr; k++ ) {
        oint nc = C_NCOL( k ), lb = C_FBLK( k + 1 );
        REAL *D = L + C_COEF( k );
        sytrfsp( nc, D, nc, npiv, crit );
        for ( b = C_FBLK( k ) + 1; b < lb; b++ ) {
            oint  M = B_NROW( b );
            REAL *X = L + C_COEF( k ) + B_COEF( b ), *Y = DLt + C_COEF( k ) + B_COEF( b );
            trsm_run( M, nc, D, nc, 0, X, M );
            for ( j = 0; j < nc; j++ )
                for ( i = 0; i < M; i++ ) Y[i + j * M] = X[i + j * M] * D[j + j * nc];
        }
        for ( b = C_FBLK( k ) + 1; b < lb; b++ ) gemmsp_2d2d( cblk, blok, k, b, 0, L, DLt, L );
    }
    return 0;
}

/* Level-parallel LU (CPU-baseline leg), the scheme of potrf_par with the
 * LU kernels of the sequential driver. */
int
PFX( getrf_par )( oint cblknbr, const oint *cblk, const oint *blok, REAL *L, REAL *U, double crit, oint *npiv,
                  int nthreads )
{
    oint        k, l, nlev = 0, *lev, *order, *lptr, *pairs, npairs = 0, piv = 0;
    omp_lock_t *locks;
    lev   = calloc( cblknbr + 1, sizeof( oint ) );
    order = malloc( ( cblknbr + 1 ) * sizeof( oint ) );
    locks = malloc( ( cblknbr + 1 ) * sizeof( omp_lock_t ) );
    for ( k = 0; k < cblknbr; k++ ) {
        omp_init_lock( locks + k );
        if ( C_FBLK( k ) + 1 < C_FBLK( k + 1 ) ) {
            oint f = B_FCBK( C_FBLK( k ) + 1 );
            if ( lev[f] < lev[k] + 1 ) lev[f] = lev[k] + 1;
        }
        if ( lev[k] + 1 > nlev ) nlev = lev[k] + 1;
    }
    lptr = calloc( nlev + 2, sizeof( oint ) );
    for ( k = 0; k < cblknbr; k++ ) lptr[lev[k] + 1]++;
    for ( l = 0; l < nlev; l++ ) lptr[l + 1] += lptr[l];
    {
        oint *pos = malloc( ( nlev + 1 ) * sizeof( oint ) );
        memcpy( pos, lptr, ( nlev + 1 ) * sizeof( oint ) );
        for ( k = 0; k < cblknbr; k++ ) order[pos[lev[k]]++] = k;
        free( pos );
    }
    for ( k = 0; k < cblknbr; k++ ) npairs += C_FBLK( k + 1 ) - C_FBLK( k ) - 1;
    pairs = malloc( ( npairs + 1 ) * 2 * sizeof( oint ) );
    for ( l = 0; l < nlev; l++ ) {
        oint b0 = lptr[l], b1 = lptr[l + 1], np = 0, i;
#pragma omp parallel for num_threads( nthreads ) schedule( dynamic, 1 ) reduction( + : piv )
        for ( i = b0; i < b1; i++ ) {
            oint kk = order[i], b, ii, jj, nc = C_NCOL( kk ), p = 0;
            REAL *D = L + C_COEF( kk ), *Du = U + C_COEF( kk );
            for ( jj = 0; jj < nc; jj++ )
                for ( ii = 0; ii < nc; ii++ ) D[ii + jj * nc] += Du[jj + ii * nc];
            getrfsp( nc, D, nc, &p, crit );
            piv += p;
            for ( jj = 0; jj < nc; jj++ )
                for ( ii = 0; ii < nc; ii++ ) Du[jj + ii * nc] = D[ii + jj * nc];
            for ( b = C_FBLK( kk ) + 1; b < C_FBLK( kk + 1 ); b++ ) {
                trsm_run( B_NROW( b ), nc, D, nc, 0, L + C_COEF( kk ) + B_COEF( b ), B_NROW( b ) );
                trsm_run( B_NROW( b ), nc, Du, nc, 1, U + C_COEF( kk ) + B_COEF( b ), B_NROW( b ) );
            }
        }
        for ( i = b0; i < b1; i++ ) {
            oint kk = order[i], b;
            for ( b = C_FBLK( kk ) + 1; b < C_FBLK( kk + 1 ); b++ ) {
                pairs[2 * np]     = kk;
                pairs[2 * np + 1] = b;
                np++;
            }
        }
#pragma omp parallel for num_threads( nthreads ) schedule( dynamic, 1 )
        for ( i = 0; i < np; i++ ) {
            oint kk = pairs[2 * i], b = pairs[2 * i + 1], f = B_FCBK( b );
            omp_set_lock( locks + f );
            gemmsp_2d2d( cblk, blok, kk, b, 0, L, U, L );
            if ( b + 1 < C_FBLK( kk + 1 ) ) gemmsp_2d2d( cblk, blok, kk, b, 1, U, L, U );
            omp_unset_lock( locks + f );
        }
    }
    for ( k = 0; k < cblknbr; k++ ) omp_destroy_lock( locks + k );
    *npiv = piv;
    free( lev ); free( order ); free( locks ); free( lptr ); free( pairs );
    return 0;
}

/* ---- static scheduling (CPU-baseline leg) -------------------------------- */

/* Threads that follow a static schedule: thread t runs the cblk tasks
 * ttsk[tptr[t] .. tptr[t+1]) in that order (reference thread_zpotrf_static,
 * sopalin/sequential_zpotrf.c:104-145, and its getrf twin).  A task first
 * waits until all the block updates facing its cblk have landed (the cblk's
 * counter, initialised to its number of incoming blocks, reaches zero:
 * cpucblk_zincoming_deps, kernels/cpucblk_zmpi_coeftab.c:543-600), factors
 * the cblk, then applies each block's update to the facing cblk under that
 * cblk's lock and decrements its counter (cpucblk_zrelease_deps, :602-624;
 * the lock is the reference's pastix_cblk_lock of core_zgemmsp.c).  One
 * POSIX thread per list, so every list has its own thread. */
typedef struct {
    oint             cblknbr;
    const oint      *cblk, *blok, *tptr, *ttsk;
    REAL            *L, *U;
    double           crit;
    int              lu;
    volatile oint   *ctrb;
    pthread_mutex_t *locks;
    oint             npiv, nneg;
    int              tid;
} PFX( static_arg );

static void *
PFX( static_thread )( void *argp )
{
    PFX( static_arg ) *a = (PFX( static_arg ) *)argp;
    const oint *cblk = a->cblk, *blok = a->blok;
    REAL       *L = a->L, *U = a->U;
    oint        t, i, j, b;
    a->npiv = 0;
    a->nneg = 0;
    for ( t = a->tptr[a->tid]; t < a->tptr[a->tid + 1]; t++ ) {
        oint  k = a->ttsk[t], nc = C_NCOL( k ), lb = C_FBLK( k + 1 );
        REAL *D = L + C_COEF( k );
        while ( __atomic_load_n( &a->ctrb[k], __ATOMIC_ACQUIRE ) > 0 ) sched_yield();
        if ( a->lu ) {
            REAL *Du = U + C_COEF( k );
            for ( j = 0; j < nc; j++ )
                for ( i = 0; i < nc; i++ ) D[i + j * nc] += Du[j + i * nc];
            getrfsp( nc, D, nc, &a->npiv, a->crit );
            for ( j = 0; j < nc; j++ )
                for ( i = 0; i < nc; i++ ) Du[j + i * nc] = D[i + j * nc];
            for ( b = C_FBLK( k ) + 1; b < lb; b++ ) {
                trsm_run( B_NROW( b ), nc, D, nc, 0, L + C_COEF( k ) + B_COEF( b ), B_NROW( b ) );
                trsm_run( B_NROW( b ), nc, Du, nc, 1, U + C_COEF( k ) + B_COEF( b ), B_NROW( b ) );
            }
        }
        else {
            potrfsp( nc, D, nc, &a->npiv, &a->nneg, a->crit );
            for ( b = C_FBLK( k ) + 1; b < lb; b++ )
                trsm_rlt( B_NROW( b ), nc, D, nc, L + C_COEF( k ) + B_COEF( b ), B_NROW( b ) );
        }
        for ( b = C_FBLK( k ) + 1; b < lb; b++ ) {
            oint f = B_FCBK( b );
            pthread_mutex_lock( a->locks + f );
            if ( a->lu ) {
                gemmsp_2d2d( cblk, blok, k, b, 0, L, U, L );
                if ( b + 1 < lb ) gemmsp_2d2d( cblk, blok, k, b, 1, U, L, U );
            }
            else {
                gemmsp_2d2d( cblk, blok, k, b, 0, L, L, L );
            }
            pthread_mutex_unlock( a->locks + f );
            __atomic_fetch_sub( &a->ctrb[f], 1, __ATOMIC_RELEASE );
        }
    }
    return NULL;
}

/* lu = 0: LL^t in L; lu = 1: LU in L and U.  Returns -1 when a list does not
 * hold every cblk exactly once. */
int
PFX( factor_static )( oint cblknbr, const oint *cblk, const oint *blok, REAL *L, REAL *U, int lu, double crit,
                      oint *npiv, oint *nneg, int nthreads, const oint *tptr, const oint *ttsk )
{
    PFX( static_arg ) *args;
    pthread_t       *th;
    pthread_mutex_t *locks;
    volatile oint   *ctrb;
    char            *seen;
    oint             k, b;
    int              t, rc = 0;
    if ( nthreads < 1 || tptr[0] != 0 || tptr[nthreads] != cblknbr ) return -1;
    seen = calloc( cblknbr + 1, 1 );
    for ( k = 0; k < cblknbr; k++ ) {
        if ( ttsk[k] < 0 || ttsk[k] >= cblknbr || seen[ttsk[k]] ) {
            free( seen );
            return -1;
        }
        seen[ttsk[k]] = 1;
    }
    free( seen );
    args  = calloc( nthreads, sizeof( *args ) );
    th    = calloc( nthreads, sizeof( pthread_t ) );
    locks = malloc( ( cblknbr + 1 ) * sizeof( pthread_mutex_t ) );
    ctrb  = calloc( cblknbr + 1, sizeof( oint ) );
    for ( k = 0; k < cblknbr; k++ ) pthread_mutex_init( locks + k, NULL );
    for ( k = 0; k < cblknbr; k++ )
        for ( b = C_FBLK( k ) + 1; b < C_FBLK( k + 1 ); b++ ) ctrb[B_FCBK( b )]++;
    for ( t = 0; t < nthreads; t++ ) {
        args[t].cblknbr = cblknbr; args[t].cblk = cblk; args[t].blok = blok;
        args[t].tptr = tptr; args[t].ttsk = ttsk; args[t].L = L; args[t].U = U;
        args[t].crit = crit; args[t].lu = lu; args[t].ctrb = ctrb; args[t].locks = locks; args[t].tid = t;
        if ( pthread_create( th + t, NULL, PFX( static_thread ), args + t ) != 0 ) {
            /* a missing thread would leave its tasks' dependents waiting */
            fprintf( stderr, "oracle factor_static: cannot start thread %d\n", t );
            abort();
        }
    }
    *npiv = 0;
    if ( nneg ) *nneg = 0;
    for ( t = 0; t < nthreads; t++ ) {
        pthread_join( th[t], NULL );
        *npiv += args[t].npiv;
        if ( nneg ) *nneg += args[t].nneg;
    }
    for ( k = 0; k < cblknbr; k++ ) pthread_mutex_destroy( locks + k );
    free( args ); free( th ); free( locks ); free( (void *)ctrb );
    return rc;
}

/* Supernodal solve in the factorization numbering, x is n x nrhs (ld).
 * fwdC/bwdC select the factor tiles of each sweep and unit_f/unit_b the
 * diagonal (reference sopalin/pastix_task_solve.c:543-649,
 * kernels/solve_ztrsmsp.c:302-661). */
static void
solve_sweeps( oint cblknbr, const oint *cblk, const oint *blok, const REAL *fwdC, int unit_f, const REAL *bwdC,
              int unit_b, const REAL *dscale, oint nrhs, REAL *x, oint ld )
{
    oint k, b, r, i, j;
    for ( k = 0; k < cblknbr; k++ ) {
        oint nc = C_NCOL( k ), f0 = C_FCOL( k );
        const REAL *D = fwdC + C_COEF( k );
        for ( r = 0; r < nrhs; r++ ) {
            REAL *xk = x + r * ld + f0;
            for ( j = 0; j < nc; j++ ) {
                if ( !unit_f ) xk[j] /= D[j + j * nc];
                for ( i = j + 1; i < nc; i++ ) xk[i] -= D[i + j * nc] * xk[j];
            }
            for ( b = C_FBLK( k ) + 1; b < C_FBLK( k + 1 ); b++ ) {
                const REAL *A = fwdC + C_COEF( k ) + B_COEF( b );
                oint M = B_NROW( b );
                REAL *y = x + r * ld + B_FROW( b );
                for ( j = 0; j < nc; j++ )
                    for ( i = 0; i < M; i++ ) y[i] -= A[i + j * M] * xk[j];
            }
        }
    }
    /* D^{-1} between the sweeps (LDL^t, reference solve_cblk_zdiag,
     * kernels/solve_ztrsmsp.c:664-716): the diagonal of each diagonal block */
    for ( k = 0; dscale && k < cblknbr; k++ ) {
        oint nc = C_NCOL( k ), f0 = C_FCOL( k );
        for ( r = 0; r < nrhs; r++ )
            for ( j = 0; j < nc; j++ ) x[r * ld + f0 + j] /= dscale[C_COEF( k ) + j * ( nc + 1 )];
    }
    for ( k = cblknbr - 1; k >= 0; k-- ) {
        oint nc = C_NCOL( k ), f0 = C_FCOL( k );
        const REAL *D = bwdC + C_COEF( k );
        for ( r = 0; r < nrhs; r++ ) {
            REAL *xk = x + r * ld + f0;
            for ( b = C_FBLK( k ) + 1; b < C_FBLK( k + 1 ); b++ ) {
                const REAL *A = bwdC + C_COEF( k ) + B_COEF( b );
                oint M = B_NROW( b );
                const REAL *y = x + r * ld + B_FROW( b );
                for ( j = 0; j < nc; j++ ) {
                    REAL s = 0;
                    for ( i = 0; i < M; i++ ) s += A[i + j * M] * y[i];
                    xk[j] -= s;
                }
            }
            for ( j = nc - 1; j >= 0; j-- ) {
                REAL s = xk[j];
                for ( i = j + 1; i < nc; i++ ) s -= D[i + j * nc] * xk[i];
                xk[j] = unit_b ? s : s / D[j + j * nc];
            }
        }
    }
}

/* factotype: 0 (or 3) = LL^t, 1 (or 4) = LDL^t, 2 = LU; trans: 0 = A x = b,
 * 1 = A^T x = b */
int
PFX( solve )( oint cblknbr, const oint *cblk, const oint *blok, int factotype, int trans, const REAL *L,
              const REAL *U, oint nrhs, REAL *x, oint ld )
{
    if ( factotype == 2 ) {
        if ( !trans ) solve_sweeps( cblknbr, cblk, blok, L, 1, U, 0, NULL, nrhs, x, ld );
        else          solve_sweeps( cblknbr, cblk, blok, U, 0, L, 1, NULL, nrhs, x, ld );
    }
    else if ( factotype == 1 || factotype == 4 ) {
        solve_sweeps( cblknbr, cblk, blok, L, 1, L, 1, L, nrhs, x, ld );
    }
    else {
        solve_sweeps( cblknbr, cblk, blok, L, 0, L, 0, NULL, nrhs, x, ld );
    }
    return 0;
}

/* Single update of the reference-signature GEMM (for per-call parity of
 * gpucblk_*gemmsp): cpucblk_zgemmsp on one (cblk, blok) pair. */
int
PFX( gemmsp_one )( const oint *cblk, const oint *blok, oint k, oint bj, int shift, const REAL *A, const REAL *B,
                   REAL *C )
{
    gemmsp_2d2d( cblk, blok, k, bj, shift, A, B, C );
    return 0;
}

/* Single diagonal factorization (per-call parity of gpucblk_*potrfsp / getrfsp). */
int
PFX( potrf_diag )( oint nc, REAL *D, double crit, oint *npiv, oint *nneg )
{
    *npiv = 0;
    *nneg = 0;
    potrfsp( nc, D, nc, npiv, nneg, crit );
    return 0;
}
int
PFX( getrf_diag )( oint nc, REAL *D, REAL *Du, double crit, oint *npiv )
{
    oint i, j;
    *npiv = 0;
    for ( j = 0; j < nc; j++ )
        for ( i = 0; i < nc; i++ ) D[i + j * nc] += Du[j + i * nc];
    getrfsp( nc, D, nc, npiv, crit );
    for ( j = 0; j < nc; j++ )
        for ( i = 0; i < nc; i++ ) Du[j + i * nc] = D[i + j * nc];
    return 0;
}
int
PFX( sytrf_diag )( oint nc, REAL *D, double crit, oint *npiv )
{
    *npiv = 0;
    sytrfsp( nc, D, nc, npiv, crit );
    return 0;
}
/* Panel TRSM of one cblk (per-call parity of gpublok_*trsmsp):
 * mode 0: X L^{-T} (POTRF), 1: X U^{-1} (LU, L side), 2: X V^{-1} unit (LU, U side) */
int
PFX( trsm_cblk )( const oint *cblk, const oint *blok, oint k, int mode, const REAL *D, REAL *panel )
{
    oint b, nc = C_NCOL( k );
    for ( b = C_FBLK( k ) + 1; b < C_FBLK( k + 1 ); b++ ) {
        REAL *X = panel + B_COEF( b );
        if ( mode == 0 ) trsm_rlt( B_NROW( b ), nc, D, nc, X, B_NROW( b ) );
        else             trsm_run( B_NROW( b ), nc, D, nc, mode == 2, X, B_NROW( b ) );
    }
    return 0;
}
