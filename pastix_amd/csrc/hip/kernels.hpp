/*
 * kernels.hpp -- CDNA4 (gfx950) device kernels of the supernodal factor+solve
 * engine.  All kernels read the supernodal structure (SolverMatrix in
 * structure-of-arrays form, see pa_dstruct) from HBM and operate on the coeftab
 * (the 2D-tiled L and U^T panels, reference sopalin/coeftab.c:98-140) that
 * stays resident in HBM for the whole factorization and solve.
 *
 * Precision is a template parameter: T = double (FP64 path) or float (FP32
 * factor of the mixed-precision path).  FP32 and FP64 matrix products run on
 * v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double pa_d4 __attribute__( ( ext_vector_type( 4 ) ) );
typedef float  pa_f4 __attribute__( ( ext_vector_type( 4 ) ) );

/* Supernodal structure on the device (one entry per cblk / per block). */
struct pa_dstruct {
    int32_t        cblknbr;
    int32_t        bloknbr;
    const int32_t *c_fcol;   /* [cblknbr+1] first column                         */
    const int32_t *c_ncol;   /* [cblknbr]   number of columns                    */
    const int32_t *c_stride; /* [cblknbr]   number of rows (all blocks)          */
    const int32_t *c_fblok;  /* [cblknbr+1] first block (the diagonal one)       */
    const int64_t *c_coef;   /* [cblknbr+1] offset of the cblk panel in coeftab  */
    const int32_t *b_frow;   /* [bloknbr]   first row                            */
    const int32_t *b_lrow;   /* [bloknbr]   last row (inclusive)                 */
    const int32_t *b_fcblk;  /* [bloknbr]   faced cblk                           */
    const int32_t *b_rowoff; /* [bloknbr]   rows of the blocks above in the cblk */
};

/* One output tile of the scatter-GEMM of a level. */
struct pa_gemm_task {
    int32_t k;      /* source cblk                                             */
    int32_t F;      /* target (faced) cblk of the sigma rows                   */
    int32_t s0;     /* first sigma row (below-diagonal row index of cblk k)    */
    int32_t r0;     /* first rho row                                           */
    int32_t slen;   /* number of sigma rows (<= TN)                            */
    int32_t flags;  /* PA_TASK_ATOMIC | PA_TASK_USIDE                          */
    int32_t pad0, pad1;
};
#define PA_TASK_ATOMIC 1
#define PA_TASK_USIDE  2

/* Kernels get internal linkage: each translation unit that launches them
 * carries its own copy. */
namespace {

/* ------------------------------------------------------------------------- */
/* helpers                                                                   */
/* ------------------------------------------------------------------------- */
template <typename T> __device__ __forceinline__ T pa_abs( T x ) { return x < T( 0 ) ? -x : x; }

__device__ __forceinline__ void pa_atomic_add( double *p, double v ) { unsafeAtomicAdd( p, v ); }
__device__ __forceinline__ void pa_atomic_add( float *p, float v ) { unsafeAtomicAdd( p, v ); }

/* Block of cblk k (blocks [lo, hi]) holding row index x of the cblk, where a
 * block starts at b_rowoff.  Blocks are sorted by rowoff. */
__device__ __forceinline__ int32_t
pa_find_rowblk( const int32_t *rowoff, int32_t lo, int32_t hi, int32_t x )
{
    while ( lo < hi ) {
        int32_t mid = ( lo + hi + 1 ) >> 1;
        if ( rowoff[mid] <= x ) lo = mid;
        else                    hi = mid - 1;
    }
    return lo;
}

/* Block of a cblk (blocks [lo, hi]) whose global row range holds `grow`. */
__device__ __forceinline__ int32_t
pa_find_globrow( const int32_t *frow, int32_t lo, int32_t hi, int32_t grow )
{
    while ( lo < hi ) {
        int32_t mid = ( lo + hi + 1 ) >> 1;
        if ( frow[mid] <= grow ) lo = mid;
        else                     hi = mid - 1;
    }
    return lo;
}

/* ------------------------------------------------------------------------- */
/* coeftab fill: scatter the permuted CSC into the L (and U^T) tiles          */
/* (reference kernels/cpucblk_zinit.c:395-480, cpucblk_zcinit.c:80-331).     */
/* One thread per column.                                                    */
/* ------------------------------------------------------------------------- */
template <typename T>
__global__ void
k_fill( pa_dstruct S, int32_t n, const int64_t *colptr, const int32_t *rowtab,
        const double *Lv, const double *Uv, T *L, T *U, const int32_t *col2cblk,
        int32_t *overflow )
{
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if ( c >= n ) return;
    int32_t k    = col2cblk[c];
    int32_t fcol = S.c_fcol[k], nc = S.c_ncol[k];
    int32_t b = S.c_fblok[k], lb = S.c_fblok[k + 1];
    int64_t base = S.c_coef[k];
    int32_t nover = 0;
    for ( int64_t e = colptr[c]; e < colptr[c + 1]; e++ ) {
        int32_t r = rowtab[e];
        if ( r < c ) continue;
        while ( b < lb && S.b_lrow[b] < r ) b++;
        if ( b >= lb || S.b_frow[b] > r ) continue; /* outside the structure */
        int32_t M   = S.b_lrow[b] - S.b_frow[b] + 1;
        int64_t idx = base + (int64_t)S.b_rowoff[b] * nc + ( r - S.b_frow[b] ) + (int64_t)( c - fcol ) * M;
        double  v   = Lv[e];
        if ( sizeof( T ) == 4 && pa_abs( v ) > 3.4028234663852886e38 ) nover++;
        L[idx] = (T)v;
        if ( U != nullptr && r > c ) {
            double vu = Uv[e];
            if ( sizeof( T ) == 4 && pa_abs( vu ) > 3.4028234663852886e38 ) nover++;
            U[idx] = (T)vu;
        }
    }
    if ( nover ) atomicAdd( overflow, nover );
}

/* ------------------------------------------------------------------------- */
/* Diagonal block factorization, one workgroup per cblk of the level.        */
/*                                                                           */
/* POTRF: Cholesky with static pivoting: a pivot below `crit` in magnitude   */
/* is replaced by `crit` and counted (reference kernels/core_zpotrfsp.c:     */
/* 68-107).  GETRF: LU without row exchange, a small pivot is replaced by    */
/* +-crit keeping its sign (reference kernels/core_zgetrfsp.c:65-185); the   */
/* diagonal block of U^T is first merged into L and the factored block is    */
/* transposed back into U^T (core_zgeadd / core_zgetmo, :253-264).           */
/*                                                                           */
/* Right-looking over column panels of NB columns held in LDS; the trailing  */
/* update of the (L2-resident) diagonal tile is done by all 256 threads.     */
/* ------------------------------------------------------------------------- */
#define PA_DIAG_NB 16

template <typename T, bool LU>
__device__ void
pa_diag_factor( T *D, T *Du, int32_t nc, double crit, int32_t *npiv, int32_t *nneg, T *P )
{
    int     t  = threadIdx.x;
    int32_t mypiv = 0, myneg = 0;
    const T tcrit = (T)crit;

    if ( LU ) {
        /* L += U^T on the full square (strict lower of U^T holds the original
         * strict upper of A, never updated; L holds the rest) */
        for ( int32_t idx = t; idx < nc * nc; idx += 256 ) {
            int32_t i = idx % nc, j = idx / nc;
            if ( i < j ) D[i + (int64_t)j * nc] += Du[j + (int64_t)i * nc];
        }
        __syncthreads();
    }
    for ( int32_t p = 0; p < nc; p += PA_DIAG_NB ) {
        int32_t m = nc - p;
        int32_t w = m < PA_DIAG_NB ? m : PA_DIAG_NB;

        for ( int32_t idx = t; idx < w * m; idx += 256 ) {
            int32_t j = idx / m, i = idx % m;
            P[j * m + i] = D[( p + i ) + (int64_t)( p + j ) * nc];
        }
        __syncthreads();

        for ( int32_t j = 0; j < w; j++ ) {
            if ( t == 0 ) {
                T d = P[j * m + j];
                if ( pa_abs( d ) < tcrit ) {
                    d = ( LU && d < T( 0 ) ) ? -tcrit : tcrit;
                    mypiv++;
                }
                if ( !LU ) {
                    if ( d < T( 0 ) ) myneg++;
                    d = sqrt( d );
                }
                P[j * m + j] = d;
            }
            __syncthreads();
            T d = P[j * m + j];
            for ( int32_t i = j + 1 + t; i < m; i += 256 ) P[j * m + i] /= d;
            __syncthreads();
            int32_t nl = w - j - 1;
            if ( nl > 0 ) {
                int32_t ni = m - j - 1;
                for ( int32_t idx = t; idx < nl * ni; idx += 256 ) {
                    int32_t l = j + 1 + idx / ni, i = j + 1 + idx % ni;
                    if ( LU ) {
                        P[l * m + i] -= P[j * m + i] * P[l * m + j];
                    }
                    else if ( i >= l ) {
                        P[l * m + i] -= P[j * m + i] * P[j * m + l];
                    }
                }
            }
            __syncthreads();
        }

        for ( int32_t idx = t; idx < w * m; idx += 256 ) {
            int32_t j = idx / m, i = idx % m;
            D[( p + i ) + (int64_t)( p + j ) * nc] = P[j * m + i];
        }
        int32_t q = m - w; /* trailing size */
        if ( q > 0 ) {
            if ( LU ) {
                /* U12 = L11^{-1} A12, unit lower, one column per thread */
                for ( int32_t c = p + w + t; c < nc; c += 256 ) {
                    T u[PA_DIAG_NB];
                    for ( int32_t jj = 0; jj < w; jj++ ) {
                        T x = D[( p + jj ) + (int64_t)c * nc];
                        for ( int32_t ii = 0; ii < jj; ii++ ) x -= P[ii * m + jj] * u[ii];
                        u[jj] = x;
                    }
                    for ( int32_t jj = 0; jj < w; jj++ ) D[( p + jj ) + (int64_t)c * nc] = u[jj];
                }
            }
            __syncthreads();
            /* A22 -= L21 * (L21^T or U12) */
            for ( int32_t idx = t; idx < q * q; idx += 256 ) {
                int32_t i = idx % q, c = idx / q;
                if ( !LU && i < c ) continue;
                T acc = T( 0 );
                if ( LU ) {
                    for ( int32_t j = 0; j < w; j++ )
                        acc += P[j * m + w + i] * D[( p + j ) + (int64_t)( p + w + c ) * nc];
                }
                else {
                    for ( int32_t j = 0; j < w; j++ ) acc += P[j * m + w + i] * P[j * m + w + c];
                }
                D[( p + w + i ) + (int64_t)( p + w + c ) * nc] -= acc;
            }
        }
        __syncthreads();
    }

    if ( LU ) {
        for ( int32_t idx = t; idx < nc * nc; idx += 256 ) {
            int32_t i = idx % nc, j = idx / nc;
            Du[j + (int64_t)i * nc] = D[i + (int64_t)j * nc];
        }
    }
    if ( t == 0 ) {
        if ( mypiv ) atomicAdd( npiv, mypiv );
        if ( myneg ) atomicAdd( nneg, myneg );
    }
}

/* one workgroup per cblk of a level */
template <typename T, bool LU>
__global__ __launch_bounds__( 256 ) void
k_diag( pa_dstruct S, const int32_t *cblks, T *Lc, T *Uc, double crit, int32_t *npiv,
        int32_t *nneg )
{
    extern __shared__ unsigned char pa_smem_raw[];
    int32_t k = cblks[blockIdx.x];
    pa_diag_factor<T, LU>( Lc + S.c_coef[k], LU ? Uc + S.c_coef[k] : nullptr, S.c_ncol[k], crit, npiv, nneg,
                           reinterpret_cast<T *>( pa_smem_raw ) );
}

/* single diagonal block given by pointers (reference-signature entry points) */
template <typename T, bool LU>
__global__ __launch_bounds__( 256 ) void
k_diag_one( T *D, T *Du, int32_t nc, double crit, int32_t *npiv, int32_t *nneg )
{
    extern __shared__ unsigned char pa_smem_raw[];
    pa_diag_factor<T, LU>( D, Du, nc, crit, npiv, nneg, reinterpret_cast<T *>( pa_smem_raw ) );
}

/* ------------------------------------------------------------------------- */
/* Panel TRSM on the off-diagonal rows of a cblk (reference                   */
/* kernels/core_ztrsmsp.c:149-180):                                          */
/*   MODE 0 (POTRF):  X <- X L^{-T}, L lower non-unit diagonal tile          */
/*   MODE 1 (LU, L):  X <- X U^{-1}, U upper non-unit part of the L tile     */
/*   MODE 2 (LU, U):  X <- X V^{-1}, V upper unit part of the U^T tile       */
/* A task is (cblk, first below-diagonal row); TR rows per workgroup are     */
/* staged in LDS and solved with a column-blocked algorithm.                 */
/* ------------------------------------------------------------------------- */
#define PA_TRSM_TR 16
#define PA_TRSM_CB 16

/* Solve `rows` (<= TR) rows whose offsets/leading dimensions inside `panel`
 * are xoff/xld (in LDS); D is the nc x nc diagonal tile. */
template <typename T, int MODE>
__device__ void
pa_trsm_rows( T *panel, const T *D, int32_t nc, int32_t rows, const int64_t *xoff, const int32_t *xld, T *Xs )
{
    int t = threadIdx.x;
    for ( int32_t idx = t; idx < nc * PA_TRSM_TR; idx += 256 ) {
        int32_t r = idx % PA_TRSM_TR, j = idx / PA_TRSM_TR;
        Xs[j * PA_TRSM_TR + r] = ( r < rows ) ? panel[xoff[r] + (int64_t)j * xld[r]] : T( 0 );
    }
    __syncthreads();

    for ( int32_t jb = 0; jb < nc; jb += PA_TRSM_CB ) {
        int32_t cw = nc - jb < PA_TRSM_CB ? nc - jb : PA_TRSM_CB;
        {
            int32_t r = t % PA_TRSM_TR, jj = t / PA_TRSM_TR;
            if ( jj < cw && jb > 0 ) {
                int32_t j = jb + jj;
                T acc = T( 0 );
                if ( MODE == 0 ) {
                    for ( int32_t l = 0; l < jb; l++ ) acc += Xs[l * PA_TRSM_TR + r] * D[j + (int64_t)l * nc];
                }
                else {
                    const T *col = D + (int64_t)j * nc;
                    for ( int32_t l = 0; l < jb; l++ ) acc += Xs[l * PA_TRSM_TR + r] * col[l];
                }
                Xs[j * PA_TRSM_TR + r] -= acc;
            }
        }
        __syncthreads();
        if ( t < rows ) {
            int32_t r = t;
            for ( int32_t jj = 0; jj < cw; jj++ ) {
                int32_t j = jb + jj;
                T x = Xs[j * PA_TRSM_TR + r];
                for ( int32_t l = jb; l < j; l++ ) {
                    T cf = ( MODE == 0 ) ? D[j + (int64_t)l * nc] : D[l + (int64_t)j * nc];
                    x -= Xs[l * PA_TRSM_TR + r] * cf;
                }
                if ( MODE != 2 ) x /= D[j + (int64_t)j * nc];
                Xs[j * PA_TRSM_TR + r] = x;
            }
        }
        __syncthreads();
    }
    for ( int32_t idx = t; idx < nc * PA_TRSM_TR; idx += 256 ) {
        int32_t r = idx % PA_TRSM_TR, j = idx / PA_TRSM_TR;
        if ( r < rows ) panel[xoff[r] + (int64_t)j * xld[r]] = Xs[j * PA_TRSM_TR + r];
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__( 256 ) void
k_trsm( pa_dstruct S, const int2 *tasks, T *Lc, T *Uc )
{
    extern __shared__ unsigned char pa_smem_raw[];
    __shared__ int64_t xoff[PA_TRSM_TR];
    __shared__ int32_t xld[PA_TRSM_TR];

    int2    task = tasks[blockIdx.x];
    int32_t k = task.x, r0 = task.y;
    int32_t nc = S.c_ncol[k];
    int32_t R  = S.c_stride[k] - nc;
    int32_t rows = R - r0 < PA_TRSM_TR ? R - r0 : PA_TRSM_TR;
    T      *panel = ( MODE == 2 ? Uc : Lc ) + S.c_coef[k];
    int t = threadIdx.x;

    if ( t < PA_TRSM_TR ) {
        if ( t < rows ) {
            int32_t x = nc + r0 + t;
            int32_t b = pa_find_rowblk( S.b_rowoff, S.c_fblok[k] + 1, S.c_fblok[k + 1] - 1, x );
            int32_t loc = x - S.b_rowoff[b];
            xoff[t] = (int64_t)S.b_rowoff[b] * nc + loc;
            xld[t]  = S.b_lrow[b] - S.b_frow[b] + 1;
        }
        else {
            xoff[t] = 0;
            xld[t]  = 0;
        }
    }
    __syncthreads();
    pa_trsm_rows<T, MODE>( panel, panel, nc, rows, xoff, xld, reinterpret_cast<T *>( pa_smem_raw ) );
}

/* Per-call description of a run of blocks (reference-signature entry points):
 * block i holds rows [pref[i], pref[i+1]) of the concatenated rows, starts at
 * off[i] with leading dimension ld[i]; for targets, toff/tld describe where
 * the rows land. */
#define PA_CALL_MAXBLK 40
struct pa_call_rows {
    int32_t nblk;
    int32_t pref[PA_CALL_MAXBLK + 1];
    int32_t ld[PA_CALL_MAXBLK];
    int32_t tld[PA_CALL_MAXBLK];
    int64_t off[PA_CALL_MAXBLK];
    int64_t toff[PA_CALL_MAXBLK];
};

__device__ __forceinline__ int32_t
pa_call_find( const pa_call_rows &R, int32_t v )
{
    int32_t lo = 0, hi = R.nblk - 1;
    while ( lo < hi ) {
        int32_t mid = ( lo + hi + 1 ) >> 1;
        if ( R.pref[mid] <= v ) lo = mid;
        else                    hi = mid - 1;
    }
    return lo;
}

/* TRSM over the rows of a run of blocks given by value; one workgroup per TR rows. */
template <typename T, int MODE>
__global__ __launch_bounds__( 256 ) void
k_trsm_call( T *C, const T *D, int32_t nc, pa_call_rows R )
{
    extern __shared__ unsigned char pa_smem_raw[];
    __shared__ int64_t xoff[PA_TRSM_TR];
    __shared__ int32_t xld[PA_TRSM_TR];
    int32_t total = R.pref[R.nblk];
    int32_t r0 = blockIdx.x * PA_TRSM_TR;
    int32_t rows = total - r0 < PA_TRSM_TR ? total - r0 : PA_TRSM_TR;
    int t = threadIdx.x;
    if ( t < PA_TRSM_TR ) {
        if ( t < rows ) {
            int32_t b = pa_call_find( R, r0 + t );
            xoff[t] = R.off[b] + ( r0 + t - R.pref[b] );
            xld[t]  = R.ld[b];
        }
        else {
            xoff[t] = 0;
            xld[t]  = 0;
        }
    }
    __syncthreads();
    pa_trsm_rows<T, MODE>( C, D, nc, rows, xoff, xld, reinterpret_cast<T *>( pa_smem_raw ) );
}

/* ------------------------------------------------------------------------- */
/* Scatter-GEMM (reference kernels/core_zgemmsp.c:404-481): for a cblk k,    */
/* for every pair of its off-diagonal blocks i >= j (i > j on the U side),   */
/*     C_f[rows(i), rows(j) - fcol(F)] -= A_i B_j^T                          */
/* where F = fcblk(j) and f is the block of F holding rows(i).  The update   */
/* of a cblk is one virtual (R x R) product over its R below-diagonal rows,  */
/* tiled into 64 x 64 output tiles (sigma = columns, rho = rows).  At tile   */
/* start the row map of the tile -- source offsets and leading dimensions,   */
/* and the target offset of each rho row inside F -- is staged in LDS; the   */
/* product runs on MFMA 16x16x4 and the epilogue writes straight into F.     */
/* The product is computed transposed (sigma x rho) so that the 16 lanes of  */
/* an MFMA output row hold 16 consecutive rows of one target column: the     */
/* scatter store is coalesced.                                               */
/* ------------------------------------------------------------------------- */
#define PA_GT  64
#define PA_GKC 16

template <typename T> struct pa_mfma;
/*
 * Operand layout of the 16x16x4 MFMAs (both precisions): lane l supplies
 * A[l % 16][l / 16] and B[l / 16][l % 16].  Output element r of lane l is
 * D[row][l % 16] with row = 4 * (l / 16) + r for f32 but row = l / 16 + 4 * r
 * for f64 (measured on gfx950).
 */
template <> struct pa_mfma<double> {
    typedef pa_d4 acc_t;
    static __device__ __forceinline__ acc_t run( double a, double b, acc_t c ) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64( a, b, c, 0, 0, 0 );
    }
    static __device__ __forceinline__ int row( int lane, int r ) { return ( lane >> 4 ) + 4 * r; }
};
template <> struct pa_mfma<float> {
    typedef pa_f4 acc_t;
    static __device__ __forceinline__ acc_t run( float a, float b, acc_t c ) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32( a, b, c, 0, 0, 0 );
    }
    static __device__ __forceinline__ int row( int lane, int r ) { return 4 * ( lane >> 4 ) + r; }
};

/*
 * Tile core shared by the level-batched kernel and the reference-signature
 * entry points.  Row maps (in LDS) must be filled by the caller:
 *   a_off/a_ld: source of the TM rho rows inside Asrc (a_ld = 0: empty row)
 *   b_off/b_ld: source of the TN sigma rows inside Bsrc
 *   c_off/c_ld: target row base and leading dimension for each rho row
 *   c_col:      target column of each sigma row
 *   a_blk/b_blk: block index of each row for the i >= j (+shift) mask
 */
template <typename T>
__device__ __forceinline__ void
pa_gemm_tile( const T *__restrict__ Asrc, const T *__restrict__ Bsrc, T *C, int32_t K,
              const int64_t *a_off, const int32_t *a_ld, const int64_t *b_off, const int32_t *b_ld,
              const int64_t *c_off, const int32_t *c_ld, const int32_t *c_col,
              const int32_t *a_blk, const int32_t *b_blk, int32_t shift, bool atomic,
              T ( *As )[PA_GT], T ( *Bs )[PA_GT] )
{
    typedef typename pa_mfma<T>::acc_t acc_t;
    const int t    = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm   = wave & 1;  /* rho half   */
    const int wn   = wave >> 1; /* sigma half */
    acc_t acc[2][2];
    for ( int i = 0; i < 2; i++ )
        for ( int j = 0; j < 2; j++ )
            for ( int r = 0; r < 4; r++ ) acc[i][j][r] = T( 0 );

    const int lrow = t & 63, lk = t >> 6;
    const int64_t aof = a_off[lrow], bof = b_off[lrow];
    const int32_t ald = a_ld[lrow], bld = b_ld[lrow];

    for ( int32_t k0 = 0; k0 < K; k0 += PA_GKC ) {
#pragma unroll
        for ( int q = 0; q < PA_GKC / 4; q++ ) {
            int32_t kc = lk + 4 * q;
            int32_t kk = k0 + kc;
            As[kc][lrow] = ( ald > 0 && kk < K ) ? Asrc[aof + (int64_t)kk * ald] : T( 0 );
            Bs[kc][lrow] = ( bld > 0 && kk < K ) ? Bsrc[bof + (int64_t)kk * bld] : T( 0 );
        }
        __syncthreads();
#pragma unroll
        for ( int ks = 0; ks < PA_GKC / 4; ks++ ) {
            int kk = ks * 4 + ( lane >> 4 );
            T xa[2], yb[2];
            for ( int mi = 0; mi < 2; mi++ ) xa[mi] = Bs[kk][wn * 32 + mi * 16 + ( lane & 15 )];
            for ( int mj = 0; mj < 2; mj++ ) yb[mj] = As[kk][wm * 32 + mj * 16 + ( lane & 15 )];
            for ( int mi = 0; mi < 2; mi++ )
                for ( int mj = 0; mj < 2; mj++ ) acc[mi][mj] = pa_mfma<T>::run( xa[mi], yb[mj], acc[mi][mj] );
        }
        __syncthreads();
    }

    for ( int mi = 0; mi < 2; mi++ ) {
        for ( int mj = 0; mj < 2; mj++ ) {
            int rl = wm * 32 + mj * 16 + ( lane & 15 );
            if ( a_ld[rl] == 0 ) continue;
            int64_t cbase = c_off[rl];
            int32_t cld   = c_ld[rl];
            int32_t ablk  = a_blk[rl];
            for ( int r = 0; r < 4; r++ ) {
                int sl = wn * 32 + mi * 16 + pa_mfma<T>::row( lane, r );
                if ( b_ld[sl] == 0 || ablk < b_blk[sl] + shift ) continue;
                T *dst = C + cbase + (int64_t)c_col[sl] * cld;
                T  v   = acc[mi][mj][r];
                if ( atomic ) pa_atomic_add( dst, -v );
                else          *dst -= v;
            }
        }
    }
}

/* Level-batched scatter-GEMM: one workgroup per pa_gemm_task. */
template <typename T>
__global__ __launch_bounds__( 256 ) void
k_gemm_level( pa_dstruct S, const pa_gemm_task *tasks, T *Lc, T *Uc )
{
    __shared__ T As[PA_GKC][PA_GT];
    __shared__ T Bs[PA_GKC][PA_GT];
    __shared__ int64_t a_off[PA_GT], b_off[PA_GT], c_off[PA_GT];
    __shared__ int32_t a_ld[PA_GT], b_ld[PA_GT], c_ld[PA_GT], c_col[PA_GT], a_blk[PA_GT], b_blk[PA_GT];

    const pa_gemm_task tk = tasks[blockIdx.x];
    const int32_t k = tk.k, F = tk.F, nc = S.c_ncol[k];
    const int32_t R = S.c_stride[k] - nc;
    const int32_t fb = S.c_fblok[k], lb = S.c_fblok[k + 1];
    const bool uside = ( tk.flags & PA_TASK_USIDE ) != 0;
    const int t = threadIdx.x;

    if ( t < PA_GT ) {
        /* rho row */
        int32_t v = tk.r0 + t;
        if ( v < R ) {
            int32_t x = nc + v;
            int32_t b = pa_find_rowblk( S.b_rowoff, fb + 1, lb - 1, x );
            int32_t loc  = x - S.b_rowoff[b];
            int32_t grow = S.b_frow[b] + loc;
            a_off[t] = (int64_t)S.b_rowoff[b] * nc + loc;
            a_ld[t]  = S.b_lrow[b] - S.b_frow[b] + 1;
            a_blk[t] = b;
            int32_t ncF = S.c_ncol[F];
            int32_t f   = pa_find_globrow( S.b_frow, S.c_fblok[F], S.c_fblok[F + 1] - 1, grow );
            c_off[t] = S.c_coef[F] + (int64_t)S.b_rowoff[f] * ncF + ( grow - S.b_frow[f] );
            c_ld[t]  = S.b_lrow[f] - S.b_frow[f] + 1;
            /* a row outside the target structure is never written */
            if ( grow < S.b_frow[f] || grow > S.b_lrow[f] ) a_ld[t] = 0;
        }
        else {
            a_off[t] = 0; a_ld[t] = 0; a_blk[t] = 0; c_off[t] = 0; c_ld[t] = 0;
        }
    }
    else if ( t < 2 * PA_GT ) {
        int32_t s = t - PA_GT;
        int32_t v = tk.s0 + s;
        if ( s < tk.slen && v < R ) {
            int32_t x = nc + v;
            int32_t b = pa_find_rowblk( S.b_rowoff, fb + 1, lb - 1, x );
            int32_t loc  = x - S.b_rowoff[b];
            int32_t grow = S.b_frow[b] + loc;
            b_off[s] = (int64_t)S.b_rowoff[b] * nc + loc;
            b_ld[s]  = S.b_lrow[b] - S.b_frow[b] + 1;
            b_blk[s] = b;
            c_col[s] = grow - S.c_fcol[F];
            /* a column outside the target cblk is never written */
            if ( c_col[s] < 0 || c_col[s] >= S.c_ncol[F] ) b_ld[s] = 0;
        }
        else {
            b_off[s] = 0; b_ld[s] = 0; b_blk[s] = 0; c_col[s] = 0;
        }
    }
    __syncthreads();

    const int64_t cbase = S.c_coef[k];
    const T *Asrc = ( uside ? Uc : Lc ) + cbase;
    const T *Bsrc = ( ( Uc != nullptr && !uside ) ? Uc : Lc ) + cbase;
    T *C = uside ? Uc : Lc;
    pa_gemm_tile<T>( Asrc, Bsrc, C, nc, a_off, a_ld, b_off, b_ld, c_off, c_ld, c_col, a_blk, b_blk,
                     uside ? 1 : 0, ( tk.flags & PA_TASK_ATOMIC ) != 0, As, Bs );
}

/*
 * Scatter-GEMM over two runs of blocks given by value (reference-signature
 * entry points gpucblk_*gemmsp / gpublok_*gemmsp): rho rows from `Ar`
 * (A blocks with their target offsets), sigma rows from `Br` (B blocks), the
 * target column of sigma row s being ccol0 + s.  Grid: (rho tiles, sigma tiles).
 */
template <typename T>
__global__ __launch_bounds__( 256 ) void
k_gemm_call( const T *A, const T *B, T *C, int32_t K, pa_call_rows Ar, pa_call_rows Br, int32_t ccol0 )
{
    __shared__ T As[PA_GKC][PA_GT];
    __shared__ T Bs[PA_GKC][PA_GT];
    __shared__ int64_t a_off[PA_GT], b_off[PA_GT], c_off[PA_GT];
    __shared__ int32_t a_ld[PA_GT], b_ld[PA_GT], c_ld[PA_GT], c_col[PA_GT], a_blk[PA_GT], b_blk[PA_GT];
    const int t = threadIdx.x;
    const int32_t r0 = blockIdx.x * PA_GT, s0 = blockIdx.y * PA_GT;
    if ( t < PA_GT ) {
        int32_t v = r0 + t;
        if ( v < Ar.pref[Ar.nblk] ) {
            int32_t b = pa_call_find( Ar, v );
            a_off[t] = Ar.off[b] + ( v - Ar.pref[b] );
            a_ld[t]  = Ar.ld[b];
            c_off[t] = Ar.toff[b] + ( v - Ar.pref[b] );
            c_ld[t]  = Ar.tld[b];
        }
        else {
            a_off[t] = 0; a_ld[t] = 0; c_off[t] = 0; c_ld[t] = 0;
        }
        a_blk[t] = 0;
    }
    else if ( t < 2 * PA_GT ) {
        int32_t s = t - PA_GT, v = s0 + s;
        if ( v < Br.pref[Br.nblk] ) {
            int32_t b = pa_call_find( Br, v );
            b_off[s] = Br.off[b] + ( v - Br.pref[b] );
            b_ld[s]  = Br.ld[b];
            c_col[s] = ccol0 + (int32_t)Br.toff[b] + ( v - Br.pref[b] );
        }
        else {
            b_off[s] = 0; b_ld[s] = 0; c_col[s] = 0;
        }
        b_blk[s] = 0;
    }
    __syncthreads();
    pa_gemm_tile<T>( A, B, C, K, a_off, a_ld, b_off, b_ld, c_off, c_ld, c_col, a_blk, b_blk, 0, false, As, Bs );
}

/* ------------------------------------------------------------------------- */
/* Supernodal triangular solve (reference kernels/solve_ztrsmsp.c:302-661).  */
/* x is an n x nrhs column-major array in the factorization numbering.       */
/* ------------------------------------------------------------------------- */

/* x_k <- D_kk^{-1} x_k (lower, forward) or D_kk^{-T} x_k (backward), one     */
/* workgroup per cblk of the level; x_k staged in LDS.                       */
template <typename T, bool BACKWARD, bool UNIT>
__global__ __launch_bounds__( 256 ) void
k_solve_diag( pa_dstruct S, const int32_t *cblks, const T *Coef, T *x, int32_t nrhs, int64_t ldx )
{
    extern __shared__ unsigned char pa_smem_raw[];
    T *xs = reinterpret_cast<T *>( pa_smem_raw ); /* [nrhs][nc] */
    int32_t k = cblks[blockIdx.x], nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const T *D = Coef + S.c_coef[k];
    int t = threadIdx.x;
    for ( int32_t idx = t; idx < nc * nrhs; idx += 256 ) {
        int32_t i = idx % nc, r = idx / nc;
        xs[r * nc + i] = x[fcol + i + r * ldx];
    }
    __syncthreads();
    if ( !BACKWARD ) {
        for ( int32_t j = 0; j < nc; j++ ) {
            if ( !UNIT ) {
                for ( int32_t r = t; r < nrhs; r += 256 ) xs[r * nc + j] /= D[j + (int64_t)j * nc];
                __syncthreads();
            }
            int32_t cnt = ( nc - j - 1 ) * nrhs;
            for ( int32_t idx = t; idx < cnt; idx += 256 ) {
                int32_t i = j + 1 + idx % ( nc - j - 1 ), r = idx / ( nc - j - 1 );
                xs[r * nc + i] -= D[i + (int64_t)j * nc] * xs[r * nc + j];
            }
            __syncthreads();
        }
    }
    else {
        for ( int32_t j = nc - 1; j >= 0; j-- ) {
            /* x_j -= sum_{i>j} D[i][j] x_i, then divide */
            for ( int32_t r = 0; r < nrhs; r++ ) {
                T part = T( 0 );
                for ( int32_t i = j + 1 + t; i < nc; i += 256 ) part += D[i + (int64_t)j * nc] * xs[r * nc + i];
                /* block reduction */
                for ( int off = 32; off > 0; off >>= 1 ) part += __shfl_down( part, off, 64 );
                __shared__ T red[4];
                if ( ( t & 63 ) == 0 ) red[t >> 6] = part;
                __syncthreads();
                if ( t == 0 ) {
                    T s = red[0] + red[1] + red[2] + red[3];
                    T v = xs[r * nc + j] - s;
                    if ( !UNIT ) v /= D[j + (int64_t)j * nc];
                    xs[r * nc + j] = v;
                }
                __syncthreads();
            }
        }
    }
    for ( int32_t idx = t; idx < nc * nrhs; idx += 256 ) {
        int32_t i = idx % nc, r = idx / nc;
        x[fcol + i + r * ldx] = xs[r * nc + i];
    }
}

/* Forward update: for the below-diagonal rows of a cblk, x[row] -= A_row x_k.
 * Task = (cblk, first below-diagonal row); one thread per row. */
#define PA_SOLVE_ROWS 256
template <typename T>
__global__ __launch_bounds__( 256 ) void
k_solve_fwd_update( pa_dstruct S, const int2 *tasks, const T *Coef, T *x, int32_t nrhs, int64_t ldx )
{
    int2    tk = tasks[blockIdx.x];
    int32_t k = tk.x, nc = S.c_ncol[k], fcol = S.c_fcol[k];
    int32_t R = S.c_stride[k] - nc;
    int32_t v = tk.y + threadIdx.x;
    if ( v >= R ) return;
    int32_t xr = nc + v;
    int32_t b  = pa_find_rowblk( S.b_rowoff, S.c_fblok[k] + 1, S.c_fblok[k + 1] - 1, xr );
    int32_t loc = xr - S.b_rowoff[b];
    int32_t M  = S.b_lrow[b] - S.b_frow[b] + 1;
    int32_t grow = S.b_frow[b] + loc;
    const T *A = Coef + S.c_coef[k] + (int64_t)S.b_rowoff[b] * nc + loc;
    for ( int32_t r = 0; r < nrhs; r++ ) {
        T acc = T( 0 );
        const T *xk = x + fcol + r * ldx;
        for ( int32_t c = 0; c < nc; c++ ) acc += A[(int64_t)c * M] * xk[c];
        pa_atomic_add( x + grow + r * ldx, -acc );
    }
}

/* Backward update: x_k[c] -= sum_rows A[row][c] x[row], partial sums over a
 * chunk of rows per workgroup, reduced in LDS then added atomically. */
template <typename T>
__global__ __launch_bounds__( 256 ) void
k_solve_bwd_update( pa_dstruct S, const int2 *tasks, const T *Coef, T *x, int32_t nrhs, int64_t ldx )
{
    __shared__ int64_t roff[PA_SOLVE_ROWS];
    __shared__ int32_t rld[PA_SOLVE_ROWS];
    __shared__ T xv[PA_SOLVE_ROWS];
    __shared__ T red[4];
    int2    tk = tasks[blockIdx.x];
    int32_t k = tk.x, nc = S.c_ncol[k], fcol = S.c_fcol[k];
    int32_t R = S.c_stride[k] - nc;
    int32_t rows = R - tk.y < PA_SOLVE_ROWS ? R - tk.y : PA_SOLVE_ROWS;
    int t = threadIdx.x;
    int32_t grow = 0;
    if ( t < rows ) {
        int32_t xr = nc + tk.y + t;
        int32_t b  = pa_find_rowblk( S.b_rowoff, S.c_fblok[k] + 1, S.c_fblok[k + 1] - 1, xr );
        int32_t loc = xr - S.b_rowoff[b];
        roff[t] = (int64_t)S.b_rowoff[b] * nc + loc;
        rld[t]  = S.b_lrow[b] - S.b_frow[b] + 1;
        grow    = S.b_frow[b] + loc;
    }
    const T *A = Coef + S.c_coef[k];
    for ( int32_t r = 0; r < nrhs; r++ ) {
        __syncthreads();
        xv[t] = ( t < rows ) ? x[grow + r * ldx] : T( 0 );
        __syncthreads();
        for ( int32_t c = 0; c < nc; c++ ) {
            T part = ( t < rows ) ? A[roff[t] + (int64_t)c * rld[t]] * xv[t] : T( 0 );
            for ( int off = 32; off > 0; off >>= 1 ) part += __shfl_down( part, off, 64 );
            if ( ( t & 63 ) == 0 ) red[t >> 6] = part;
            __syncthreads();
            if ( t == 0 ) pa_atomic_add( x + fcol + c + r * ldx, -( red[0] + red[1] + red[2] + red[3] ) );
            __syncthreads();
        }
    }
}

/* ------------------------------------------------------------------------- */
/* SpMV on the permuted CSC (reference bcsc/bcsc_zspmv.c:57-200):            */
/* y = alpha op(A) x + beta y.  With vals = Uvalues (A^T by columns) a       */
/* column of the CSC is a row of A, so each output is a gathered dot product */
/* (no atomics).  8 lanes per row.                                           */
/* ------------------------------------------------------------------------- */
__global__ __launch_bounds__( 256 ) void
k_spmv( int32_t n, const int64_t *colptr, const int32_t *rowtab, const double *vals, double alpha,
        const double *x, double beta, double *y )
{
    int32_t row = ( blockIdx.x * blockDim.x + threadIdx.x ) >> 3;
    int     sub = threadIdx.x & 7;
    double  acc = 0.;
    if ( row < n ) {
        for ( int64_t e = colptr[row] + sub; e < colptr[row + 1]; e += 8 ) acc += vals[e] * x[rowtab[e]];
    }
    acc += __shfl_down( acc, 4, 8 );
    acc += __shfl_down( acc, 2, 8 );
    acc += __shfl_down( acc, 1, 8 );
    if ( row < n && sub == 0 ) y[row] = ( beta == 0. ? 0. : beta * y[row] ) + alpha * acc;
}

/* ------------------------------------------------------------------------- */
/* BLAS-1 on FP64 vectors.  Reductions write one partial per workgroup; the  */
/* host sums the partials (deterministic).                                   */
/* ------------------------------------------------------------------------- */
__device__ __forceinline__ double
pa_block_sum( double v, double *red )
{
    for ( int off = 32; off > 0; off >>= 1 ) v += __shfl_down( v, off, 64 );
    if ( ( threadIdx.x & 63 ) == 0 ) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.;
    if ( threadIdx.x == 0 ) s = red[0] + red[1] + red[2] + red[3];
    return s;
}

/* partial[blockIdx.y * gridDim.x + blockIdx.x] = sum_i X[:, y] . v  (multi-dot
 * for classical Gram-Schmidt: X holds nvec vectors of length n, ld n) */
__global__ __launch_bounds__( 256 ) void
k_mdot( int64_t n, const double *X, int64_t ldx, const double *v, double *partial )
{
    __shared__ double red[4];
    const double *xv = X + blockIdx.y * ldx;
    double acc = 0.;
    for ( int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 ) acc += xv[i] * v[i];
    double s = pa_block_sum( acc, red );
    if ( threadIdx.x == 0 ) partial[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

/* y += X * c (n x m times m), ld n */
__global__ __launch_bounds__( 256 ) void
k_gemv_n( int64_t n, int32_t m, const double *X, int64_t ldx, const double *c, double alpha, double *y )
{
    for ( int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 ) {
        double acc = 0.;
        for ( int32_t j = 0; j < m; j++ ) acc += X[i + j * ldx] * c[j];
        y[i] += alpha * acc;
    }
}

__global__ void
k_axpby( int64_t n, double a, const double *x, double b, double *y )
{
    for ( int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 )
        y[i] = a * x[i] + ( b == 0. ? 0. : b * y[i] );
}

/* dst[i] = src[perm[i]] (gather) for nrhs columns */
template <typename TD, typename TS>
__global__ void
k_gather( int64_t n, int32_t nrhs, const int32_t *perm, const TS *src, int64_t lds, TD *dst, int64_t ldd )
{
    for ( int64_t i = blockIdx.x * 256 + threadIdx.x; i < n * nrhs; i += (int64_t)gridDim.x * 256 ) {
        int64_t r = i / n, j = i % n;
        dst[j + r * ldd] = (TD)src[perm[j] + r * lds];
    }
}
/* dst[perm[i]] = src[i] (scatter) */
template <typename TD, typename TS>
__global__ void
k_scatter( int64_t n, int32_t nrhs, const int32_t *perm, const TS *src, int64_t lds, TD *dst, int64_t ldd )
{
    for ( int64_t i = blockIdx.x * 256 + threadIdx.x; i < n * nrhs; i += (int64_t)gridDim.x * 256 ) {
        int64_t r = i / n, j = i % n;
        dst[perm[j] + r * ldd] = (TD)src[j + r * lds];
    }
}
template <typename TD, typename TS>
__global__ void
k_convert( int64_t n, const TS *src, TD *dst )
{
    for ( int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 ) dst[i] = (TD)src[i];
}

} /* anonymous namespace */
