/*
 * kernels.hpp -- CDNA4 (gfx950) device kernels of the supernodal factor+solve
 * engine.  All kernels read the supernodal structure (SolverMatrix in
 * structure-of-arrays form, see pa_dstruct) from HBM and operate on the coeftab
 * (the 2D-tiled L and U^T panels, reference sopalin/coeftab.c:98-140) that
 * stays resident in HBM for the whole factorization and solve.
 *
 * Precision is a template parameter: T = double (FP64 path) or float (FP32
 * factor of the mixed-precision path).  FP32 and FP64 matrix products run on
 * v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64, the FP32 scatter-GEMM tiles
 * with 32 x 32 wave blocks on v_mfma_f32_32x32x2f32.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double pa_d4 __attribute__( ( ext_vector_type( 4 ) ) );
typedef float  pa_f4 __attribute__( ( ext_vector_type( 4 ) ) );
typedef float  pa_f16 __attribute__( ( ext_vector_type( 16 ) ) );

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
    const int64_t *c_dtoff;  /* [cblknbr]   offset of the transposed diagonal
                                            block in the LL^t transposed array   */
    /* per off-diagonal row of every cblk (cblk k's rows start at r_base[k]):
     * x = global row, y = height of its block, (z, w) = coeftab offset of the
     * entry (row, first column) as low/high 32-bit halves */
    const int64_t *r_base;   /* [cblknbr]   first row descriptor of the cblk     */
    const int4    *r_info;   /* [sum of off-diagonal rows]                       */
    const int32_t *r_blk;    /* [sum of off-diagonal rows] block of the row      */
    const int64_t *c_tinv;   /* [cblknbr]   first inverted diagonal tile of the
                                            cblk (k_tile_inv, k_trsm_inv)        */
};

/* LDS row padding of the GEMM operand tiles: 64 bytes, so that the four
 * k-rows read by one MFMA operand load fall in different banks */
template <typename T> struct pa_gpad { static constexpr int value = 64 / (int)sizeof( T ); };

__device__ __forceinline__ int64_t pa_rinfo_off( const int4 &ri )
{
    return (int64_t)( (uint64_t)(uint32_t)ri.z | ( (uint64_t)(uint32_t)ri.w << 32 ) );
}

/* One output tile of the scatter-GEMM of a level. */
/* target-cblk blocks staged in LDS by the level GEMM (larger targets search
 * the block table in global memory) */
#define PA_GF_STAGE 512

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

/* wave-uniform broadcast of lane l's value (v_readlane, no LDS round trip) */
__device__ __forceinline__ float
pa_readlane( float v, int l )
{
    return __int_as_float( __builtin_amdgcn_readlane( __float_as_int( v ), l ) );
}
__device__ __forceinline__ double
pa_readlane( double v, int l )
{
    int lo = __builtin_amdgcn_readlane( __double2loint( v ), l );
    int hi = __builtin_amdgcn_readlane( __double2hiint( v ), l );
    return __hiloint2double( hi, lo );
}


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
/* Diagonal block factorization, one workgroup per cblk (1024 threads above  */
/* 64 columns, 64 below).                                                    */
/*                                                                           */
/* POTRF: Cholesky with static pivoting: a pivot below `crit` in magnitude   */
/* is replaced by `crit` and counted (reference kernels/core_zpotrfsp.c:     */
/* 68-107).  GETRF: LU without row exchange, a small pivot is replaced by    */
/* +-crit keeping its sign (reference kernels/core_zgetrfsp.c:65-185); the   */
/* diagonal block of U^T is first merged into L and the factored block is    */
/* transposed back into U^T (core_zgeadd / core_zgetmo, :253-264).           */
/*                                                                           */
/* Right-looking over panels of NB = 16 columns held in LDS: the NB x NB    */
/* top block is factored by one wave in registers with v_readlane broadcasts */
/* (no workgroup barrier per column, no branches), the rows below are solved */
/* by all threads against the pivots' reciprocals, and the trailing update   */
/* of the (L2-resident) diagonal block runs on MFMA 16x16x4 tiles.           */
/* ------------------------------------------------------------------------- */
template <typename T> struct pa_diag_nb;
template <> struct pa_diag_nb<double> { static constexpr int value = 16; };
template <> struct pa_diag_nb<float>  { static constexpr int value = 16; };
#define PA_DIAG_NB_MAX 32

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

/* Transposes between the two diagonal blocks of an LU cblk:
 * ADD: D(i, j) += Du(j, i) for i < j (the strict upper part of A, stored
 * transposed in the U^T block, merged into the block being factored);
 * ADD and MIRROR: D(i, j) = D(j, i) for i < j (LDL^t: the block is symmetric
 * and only its lower triangle holds the matrix and its updates);
 * otherwise Du = D^T (the factored block's U^T copy).  Wide blocks go through
 * LDS tiles of 32 x 32 (buf: at least 32 nc elements, the panel area of the
 * diagonal kernel) so that reads and writes are both coalesced; each thread
 * issues all of its loads of a batch of tiles before storing any.  In the
 * MIRROR case a strict upper entry is written from a strict lower one, which
 * no tile writes, so the order of the tiles does not matter. */
template <typename T, int TH, bool ADD, bool MIRROR = false>
__device__ __forceinline__ void
pa_lu_transpose( T *D, T *Du, int32_t nc, T *buf )
{
    const int t = threadIdx.x;
    constexpr int MB = 8, EPT = 1024 / TH; /* tiles per batch, elements per thread and tile */
    const int32_t cap = min( MB, ( 32 * nc ) / ( 32 * 33 ) );
    if constexpr ( TH >= 256 ) {
        if ( cap >= 2 ) {
            const int32_t ntb    = ( nc + 31 ) >> 5;
            const int32_t npairs = ADD ? ( ntb * ( ntb + 1 ) ) / 2 : ntb * ntb;
            /* tile pair pi -> (a, b): ADD enumerates a <= b, column of tiles b by b */
            auto pair = [&]( int32_t pi, int32_t &a, int32_t &b ) {
                if ( ADD ) {
                    b = (int32_t)( ( sqrtf( 8.f * (float)pi + 1.f ) - 1.f ) * 0.5f );
                    while ( ( b * ( b + 1 ) ) / 2 > pi ) b--;
                    while ( ( ( b + 1 ) * ( b + 2 ) ) / 2 <= pi ) b++;
                    a = pi - ( b * ( b + 1 ) ) / 2;
                }
                else {
                    a = pi % ntb;
                    b = pi / ntb;
                }
            };
            for ( int32_t base = 0; base < npairs; base += cap ) {
                const int32_t nb = min( cap, npairs - base );
                T v[MB][EPT];
                /* source tile S, rows r fastest: ADD S = Du(rows b, cols a), else D(rows a, cols b) */
#pragma unroll
                for ( int tl = 0; tl < MB; tl++ ) {
                    int32_t a = 0, b = 0;
                    if ( tl < nb ) pair( base + tl, a, b );
                    const T *src = ( ADD && !MIRROR ) ? Du : D;
#pragma unroll
                    for ( int u = 0; u < EPT; u++ ) {
                        const int32_t e = t + u * TH, r = e & 31, c = e >> 5;
                        const int32_t sr = ( ADD ? b : a ) * 32 + r, sc = ( ADD ? a : b ) * 32 + c;
                        v[tl][u] = ( tl < nb && sr < nc && sc < nc ) ? src[sr + (int64_t)sc * nc] : T( 0 );
                    }
                }
#pragma unroll
                for ( int tl = 0; tl < MB; tl++ )
#pragma unroll
                    for ( int u = 0; u < EPT; u++ ) {
                        const int32_t e = t + u * TH, r = e & 31, c = e >> 5;
                        if ( tl < nb ) buf[tl * 1056 + r * 33 + c] = v[tl][u];
                    }
                __syncthreads();
                /* destination, rows r fastest: ADD D(a*32 + r, b*32 + c) += S(c, r), else Du(b*32 + r, a*32 + c) = S(c, r) */
#pragma unroll
                for ( int tl = 0; tl < MB; tl++ ) {
                    int32_t a = 0, b = 0;
                    if ( tl < nb ) pair( base + tl, a, b );
#pragma unroll
                    for ( int u = 0; u < EPT; u++ ) {
                        const int32_t e = t + u * TH, r = e & 31, c = e >> 5;
                        const int32_t dr = ( ADD ? a : b ) * 32 + r, dc = ( ADD ? b : a ) * 32 + c;
                        if ( tl < nb ) {
                            const T x = buf[tl * 1056 + c * 33 + r];
                            if ( ADD && MIRROR ) {
                                if ( dr < dc && dc < nc ) D[dr + (int64_t)dc * nc] = x;
                            }
                            else if ( ADD ) {
                                if ( dr < dc && dc < nc ) D[dr + (int64_t)dc * nc] += x;
                            }
                            else if ( dr < nc && dc < nc ) Du[dr + (int64_t)dc * nc] = x;
                        }
                    }
                }
                __syncthreads();
            }
            return;
        }
    }
    for ( int32_t idx = t; idx < nc * nc; idx += TH ) {
        const int32_t i = idx % nc, j = idx / nc;
        if ( ADD && MIRROR ) {
            if ( i < j ) D[i + (int64_t)j * nc] = D[j + (int64_t)i * nc];
        }
        else if ( ADD ) {
            if ( i < j ) D[i + (int64_t)j * nc] += Du[j + (int64_t)i * nc];
        }
        else Du[j + (int64_t)i * nc] = D[i + (int64_t)j * nc];
    }
    __syncthreads();
}

/* TH threads per workgroup: 1024 for wide diagonal blocks, 64 for the
 * small blocks of the lower levels */
template <typename T, bool LU, int TH, bool SYM = false>
__device__ __forceinline__ void
pa_diag_factor( T *D, T *Du, int32_t nc, double crit, int32_t *npiv, int32_t *nneg, T *P )
{
    constexpr int NB = pa_diag_nb<T>::value;
    typedef typename pa_mfma<T>::acc_t acc_t;
    __shared__ T pa_rinv[NB]; /* reciprocals of the panel's pivots */
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int32_t   mypiv = 0, myneg = 0;
    const T   tcrit = (T)crit;

    if ( LU && SYM ) {
        /* LDL^t (reference core_zsytrfsp, kernels/core_zsytrfsp.c): the
         * lower triangle holds the block with all its updates; the no-pivot
         * LU of the symmetric block then leaves L below the diagonal, D on
         * it and D L^t above it, the layout of the reference's factored
         * diagonal block */
        pa_lu_transpose<T, TH, true, true>( D, Du, nc, P );
    }
    else if ( LU ) {
        /* L += U^T on the full square: the strict lower part of U^T holds the
         * strict upper part of A with the updates of the descendant cblks
         * (U-side scatter-GEMMs), L holds the rest */
        pa_lu_transpose<T, TH, true>( D, Du, nc, P );
    }

    for ( int32_t p = 0; p < nc; p += NB ) {
        const int32_t m = nc - p;
        const int32_t w = m < NB ? m : NB;
        const int32_t q = m - w;

        for ( int32_t i = t; i < m; i += TH ) {
            T v[NB];
#pragma unroll
            for ( int j = 0; j < NB; j++ ) v[j] = ( j < w ) ? D[( p + i ) + (int64_t)( p + j ) * nc] : T( 0 );
#pragma unroll
            for ( int j = 0; j < NB; j++ )
                if ( j < w ) P[j * m + i] = v[j];
        }
        __syncthreads();

        /* (A) w x w block by wave 0, lane i holding row i.  Branch-free:
         * broadcasts by v_readlane, updates masked by selects, and columns
         * past the last one of the matrix (last panel only) factored as an
         * identity without being counted as pivots.  As the reference kernel
         * the column is scaled by the reciprocal of the pivot, which lane k
         * keeps (rinv) for the rows below. */
        if ( wave == 0 ) {
            T a[NB];
            const int i = lane;
#pragma unroll
            for ( int j = 0; j < NB; j++ ) a[j] = ( i < w && j < w ) ? P[j * m + i] : T( 0 );
            T rinv = T( 1 );
#pragma unroll
            for ( int k = 0; k < NB; k++ ) {
                const bool pad = k >= w;
                T          d   = pa_readlane( a[k], k );
                const bool rep = !pad && pa_abs( d ) < tcrit;
                if ( LU ) d = pad ? T( 1 ) : ( rep ? ( d < T( 0 ) ? -tcrit : tcrit ) : d );
                else      d = pad ? T( 1 ) : ( rep ? tcrit : d );
                mypiv += rep ? 1 : 0;
                if ( !LU ) {
                    myneg += ( d < T( 0 ) ) ? 1 : 0;
                    d = sqrt( d );
                }
                const T alpha = T( 1 ) / d;
                rinv = ( i == k ) ? alpha : rinv;
                a[k] = ( i == k ) ? d : ( ( i > k ) ? a[k] * alpha : a[k] );
#pragma unroll
                for ( int j = 0; j < NB; j++ ) {
                    if ( j > k ) {
                        if ( LU ) {
                            const T ukj = pa_readlane( a[j], k );
                            a[j] = ( i > k ) ? a[j] - a[k] * ukj : a[j];
                        }
                        else {
                            const T ljk = pa_readlane( a[k], j );
                            a[j] = ( i >= j ) ? a[j] - a[k] * ljk : a[j];
                        }
                    }
                }
            }
            if ( i < w ) {
#pragma unroll
                for ( int j = 0; j < NB; j++ ) P[j * m + i] = a[j];
            }
            if ( i < NB ) pa_rinv[i] = rinv;
        }
        __syncthreads();

        /* (B) rows below: X L11^T = A21 (POTRF) or X U11 = A21 (LU), column
         * by column (x_c *= 1 / d_c, then x_j -= x_c t_cj for j > c, the
         * order of the reference kernel).  Only full panels have rows below. */
        for ( int32_t r = w + t; r < m; r += TH ) {
            T x[NB];
#pragma unroll
            for ( int j = 0; j < NB; j++ ) x[j] = P[j * m + r];
#pragma unroll
            for ( int c = 0; c < NB; c++ ) {
                x[c] *= pa_rinv[c];
#pragma unroll
                for ( int j = 0; j < NB; j++ )
                    if ( j > c ) x[j] -= x[c] * ( LU ? P[j * m + c] : P[c * m + j] );
                /* one column's coefficient loads at a time (register budget) */
                __builtin_amdgcn_sched_barrier( 0 );
            }
#pragma unroll
            for ( int j = 0; j < NB; j++ ) P[j * m + r] = x[j];
        }
        __syncthreads();

        for ( int32_t i = t; i < m; i += TH ) {
#pragma unroll
            for ( int j = 0; j < NB; j++ )
                if ( j < w ) D[( p + i ) + (int64_t)( p + j ) * nc] = P[j * m + i];
        }
        if ( LU && q > 0 ) {
            /* (C) U12 = L11^{-1} A12, unit lower, one column per thread
             * (u_j -= l_jc u_c for j > c, column by column); the row panel
             * is also kept in LDS (Ur[j * q + c]) for the trailing update's
             * MFMA operands.  q > 0: a full panel. */
            T *Ur = P + (int64_t)NB * nc;
            for ( int32_t c = p + w + t; c < nc; c += TH ) {
                T u[NB];
#pragma unroll
                for ( int j = 0; j < NB; j++ ) u[j] = D[( p + j ) + (int64_t)c * nc];
#pragma unroll
                for ( int l = 0; l < NB; l++ ) {
#pragma unroll
                    for ( int j = 0; j < NB; j++ )
                        if ( j > l ) u[j] -= P[l * m + j] * u[l];
                    __builtin_amdgcn_sched_barrier( 0 );
                }
#pragma unroll
                for ( int j = 0; j < NB; j++ ) {
                    D[( p + j ) + (int64_t)c * nc] = u[j];
                    Ur[j * q + ( c - p - w )] = u[j];
                }
            }
        }
        __syncthreads();

        /* (D) trailing update on 16x16 MFMA tiles, NU tiles per wave and
         * round (2 for FP64 with 1024 threads: the register budget); the
         * trailing-matrix tiles of the next round are loaded while the
         * current round runs its MFMAs.  The product is formed as (col
         * side) x (row side)^T so that the 16 lanes of an output row hold 16
         * consecutive rows of one column. */
        if ( q > 0 ) {
            constexpr int NU = ( sizeof( T ) == 8 && TH > 512 ) ? 2 : 4;
            const int32_t nt = ( q + 15 ) >> 4;
            const int32_t ntiles = LU ? nt * nt : ( nt * ( nt + 1 ) ) / 2;
            T *Dt = D + ( p + w ) + (int64_t)( p + w ) * nc; /* trailing matrix */
            /* tile -> row-side index of this lane and first column (row = q: none) */
            auto coords = [&]( int32_t tile0, int32_t ( &row )[NU], int32_t ( &tcol )[NU] ) {
#pragma unroll
                for ( int u = 0; u < NU; u++ ) {
                    const int32_t tile = tile0 + u;
                    int32_t ti = 0, tc = 0;
                    if ( tile < ntiles ) {
                        if ( LU ) {
                            ti = tile % nt;
                            tc = tile / nt;
                        }
                        else {
                            ti = (int32_t)( ( sqrtf( 8.f * (float)tile + 1.f ) - 1.f ) * 0.5f );
                            while ( ( ti * ( ti + 1 ) ) / 2 > tile ) ti--;
                            while ( ( ( ti + 1 ) * ( ti + 2 ) ) / 2 <= tile ) ti++;
                            tc = tile - ( ti * ( ti + 1 ) ) / 2;
                        }
                    }
                    row[u]  = ( tile < ntiles ) ? ti * 16 + ( lane & 15 ) : q;
                    tcol[u] = tc * 16;
                }
            };
            auto load = [&]( const int32_t ( &row )[NU], const int32_t ( &tcol )[NU], T ( &cold )[NU][4] ) {
#pragma unroll
                for ( int u = 0; u < NU; u++ )
#pragma unroll
                    for ( int r = 0; r < 4; r++ ) {
                        const int32_t c = tcol[u] + pa_mfma<T>::row( lane, r );
                        cold[u][r] = ( row[u] < q && c < q ) ? Dt[row[u] + (int64_t)c * nc] : T( 0 );
                    }
            };
            int32_t row[NU], tcol[NU];
            T       cold[NU][4];
            int32_t tile0 = wave * NU;
            coords( tile0, row, tcol );
            load( row, tcol, cold );
            for ( ; tile0 < ntiles; tile0 += ( TH / 64 ) * NU ) {
                int32_t rown[NU], tcoln[NU];
                T       coldn[NU][4];
                coords( tile0 + ( TH / 64 ) * NU, rown, tcoln );
                load( rown, tcoln, coldn );
                acc_t acc[NU];
#pragma unroll
                for ( int u = 0; u < NU; u++ )
#pragma unroll
                    for ( int r = 0; r < 4; r++ ) acc[u][r] = T( 0 );
#pragma unroll
                for ( int32_t k0 = 0; k0 < NB; k0 += 4 ) {
                    const int32_t kk = k0 + ( lane >> 4 );
#pragma unroll
                    for ( int u = 0; u < NU; u++ ) {
                        const int32_t col = tcol[u] + ( lane & 15 );
                        T ra = ( row[u] < q ) ? P[kk * m + w + row[u]] : T( 0 );
                        T ca;
                        if ( LU ) ca = ( col < q ) ? P[(int64_t)NB * nc + kk * q + col] : T( 0 );
                        else      ca = ( col < q ) ? P[kk * m + w + col] : T( 0 );
                        acc[u] = pa_mfma<T>::run( ca, ra, acc[u] );
                    }
                }
#pragma unroll
                for ( int u = 0; u < NU; u++ )
#pragma unroll
                    for ( int r = 0; r < 4; r++ ) {
                        const int32_t c = tcol[u] + pa_mfma<T>::row( lane, r );
                        if ( row[u] < q && c < q ) Dt[row[u] + (int64_t)c * nc] = cold[u][r] - acc[u][r];
                    }
#pragma unroll
                for ( int u = 0; u < NU; u++ ) {
                    row[u]  = rown[u];
                    tcol[u] = tcoln[u];
#pragma unroll
                    for ( int r = 0; r < 4; r++ ) cold[u][r] = coldn[u][r];
                }
            }
        }
        __syncthreads();
    }

    if ( LU ) pa_lu_transpose<T, TH, false>( D, Du, nc, P ); /* U^T copy of the factored block */
    if ( t == 0 ) {
        if ( mypiv ) atomicAdd( npiv, mypiv );
        if ( myneg ) atomicAdd( nneg, myneg );
    }
}

/* one workgroup per cblk of a level; the pivoting criterion is read from
 * device memory so that a recorded launch sequence stays valid when it
 * changes with the matrix values */
template <typename T, bool LU, int TH, bool SYM = false>
__global__ __launch_bounds__( TH ) void
k_diag( pa_dstruct S, const int32_t *cblks, T *Lc, T *Uc, const double *critp, int32_t *npiv,
        int32_t *nneg )
{
    extern __shared__ unsigned char pa_smem_raw[];
    int32_t k = cblks[blockIdx.x];
    pa_diag_factor<T, LU, TH, SYM>( Lc + S.c_coef[k], LU ? Uc + S.c_coef[k] : nullptr, S.c_ncol[k], *critp, npiv,
                                    nneg, reinterpret_cast<T *>( pa_smem_raw ) );
}

/* out[w] = sum of vals[src[k]] for k in [ptr[w], ptr[w+1]): the permuted
 * CSC gathered from the user's stored values through the value map */
__global__ __launch_bounds__( 256 ) void
k_gather_sum( int64_t nnz, const int64_t *ptr, const int64_t *src, const double *vals, double *out )
{
    for ( int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nnz; w += (int64_t)gridDim.x * blockDim.x ) {
        double s = 0.;
        for ( int64_t k = ptr[w]; k < ptr[w + 1]; k++ ) s += vals[src[k]];
        out[w] = s;
    }
}

__global__ void
k_set_f64( double *p, double v )
{
    if ( threadIdx.x == 0 && blockIdx.x == 0 ) *p = v;
}

/* single diagonal block given by pointers (reference-signature entry points) */
template <typename T, bool LU>
__global__ __launch_bounds__( 512 ) void
k_diag_one( T *D, T *Du, int32_t nc, double crit, int32_t *npiv, int32_t *nneg )
{
    extern __shared__ unsigned char pa_smem_raw[];
    pa_diag_factor<T, LU, 512>( D, Du, nc, crit, npiv, nneg, reinterpret_cast<T *>( pa_smem_raw ) );
}

/* ------------------------------------------------------------------------- */
/* Register-resident POTRF of an FP32 diagonal block of up to                */
/* 16 * PA_DREG_NT columns by one workgroup of 1024 threads.                 */
/*                                                                           */
/* The lower triangle lives in registers as 16 x 16 tiles in MFMA            */
/* accumulator layout: tile (i, j) (column-major lower index idx) is held by */
/* wave idx % 16 in slot idx / 16.  A register tile holds C(i, j)^T: element */
/* r of lane l is C[16 i + l % 16][16 j + 4 (l / 16) + r], so lanes 0..15    */
/* hold 16 consecutive rows of one column and global accesses coalesce.      */
/* Per 16-column panel p: the owners stage tile column p in LDS (double      */
/* buffered), wave 0 factors the 16 x 16 diagonal tile with a v_readlane     */
/* chain (static pivoting as the reference kernel), every thread solves one  */
/* panel row against it by substitution, the owners apply the panel to their */
/* trailing tiles with MFMA 16x16x4, and the idle waves store the panel      */
/* while wave 0 factors the next diagonal tile.  Three barriers per panel    */
/* and no global round trip inside the sweep.                               */
/* ------------------------------------------------------------------------- */
#define PA_DREG_NT 22 /* 352 columns */
#define PA_DREG_S  16 /* slots per wave: 16 * 16 >= 22 * 23 / 2 tiles */
#define PA_DREG_LD 17

__device__ __forceinline__ void
pa_diag_potrf_reg( float *D, int32_t nc, double crit, int32_t *npiv, int32_t *nneg )
{
    typedef pa_mfma<float>::acc_t acc_t;
    __shared__ float P[2][PA_DREG_NT * 16][PA_DREG_LD];
    __shared__ __align__( 16 ) float Lt[16][16]; /* Lt[c][j] = L(j, c), 1 / L(c, c) on the diagonal */
    const int     t = threadIdx.x, lane = t & 63;
    const int     wave = __builtin_amdgcn_readfirstlane( t >> 6 );
    const int32_t nt = ( nc + 15 ) >> 4, ntiles = ( nt * ( nt + 1 ) ) >> 1;
    const float   tcrit = (float)crit;
    int32_t       mypiv = 0, myneg = 0;

    int32_t ti[PA_DREG_S], tj[PA_DREG_S];
    acc_t   acc[PA_DREG_S];
#pragma unroll
    for ( int s = 0; s < PA_DREG_S; s++ ) {
        int32_t idx = s * 16 + wave, j = 0, rem = idx;
        while ( j < nt && rem >= nt - j ) {
            rem -= nt - j;
            j++;
        }
        const bool used = idx < ntiles;
        ti[s] = __builtin_amdgcn_readfirstlane( used ? j + rem : -1 );
        tj[s] = __builtin_amdgcn_readfirstlane( used ? j : -1 );
#pragma unroll
        for ( int r = 0; r < 4; r++ ) {
            /* unconditional load from a clamped address (no branch per load) */
            const int32_t row = 16 * ti[s] + ( lane & 15 ), col = 16 * tj[s] + 4 * ( lane >> 4 ) + r;
            const int32_t rc = max( 0, min( row, nc - 1 ) ), cc = max( 0, min( col, nc - 1 ) );
            const float   v  = D[rc + (int64_t)cc * nc];
            acc[s][r] = ( used && row < nc && col < nc ) ? v : 0.f;
        }
    }

    /* (4) store of a finished panel (lower part of its diagonal tile) by
     * threads t0..: panel p is stored while wave 0 factors the diagonal tile
     * of panel p + 1 (its LDS buffer is rewritten only by panel p + 2) */
    auto store_panel = [&]( int32_t pp, int32_t tt ) {
        const float( *Ps )[PA_DREG_LD] = P[pp & 1];
        const int32_t ms = nc - 16 * pp, ws = min( 16, ms );
        if ( tt >= 0 && tt < ms ) {
            float v[16];
#pragma unroll
            for ( int c = 0; c < 16; c++ ) v[c] = Ps[tt][c];
#pragma unroll
            for ( int c = 0; c < 16; c++ )
                if ( c < ws && tt >= c ) D[( 16 * pp + tt ) + (int64_t)( 16 * pp + c ) * nc] = v[c];
        }
    };

    for ( int32_t p = 0; p < nt; p++ ) {
        float( *Pp )[PA_DREG_LD] = P[p & 1];
        const int32_t w = min( 16, nc - 16 * p ); /* columns of the panel */
        const int32_t m = nc - 16 * p;            /* rows of the panel    */
        /* (1) stage tile column p: P[16 (i - p) + row][col] */
#pragma unroll
        for ( int s = 0; s < PA_DREG_S; s++ ) {
            if ( tj[s] == p ) {
#pragma unroll
                for ( int r = 0; r < 4; r++ ) Pp[16 * ( ti[s] - p ) + ( lane & 15 )][4 * ( lane >> 4 ) + r] = acc[s][r];
            }
        }
        __syncthreads();
        /* (2) 16 x 16 diagonal tile by wave 0, lane i (mod 16) holding row i.
         * Branch-free: every lane applies every column update (the entries
         * right of the diagonal of a row are scratch, never stored or read
         * back), and columns past the last one of the matrix (last panel
         * only) factor as an identity without being counted as pivots. */
        if ( wave == 0 ) {
            const int i = lane & 15;
            float     a[16];
#pragma unroll
            for ( int j = 0; j < 16; j++ ) a[j] = Pp[i][j];
            float rinv = 1.f;
#pragma unroll
            for ( int k = 0; k < 16; k++ ) {
                /* l_kk = sqrt(a_kk), column scaled by 1 / l_kk as the
                 * reference kernel (hardware sqrt / reciprocal: the FP32
                 * factor is refined in FP64 anyway) */
                float      d   = pa_readlane( a[k], k );
                const bool pad = k >= w;
                const bool rep = !pad && fabsf( d ) < tcrit;
                d = pad ? 1.f : ( rep ? tcrit : d );
                mypiv += rep ? 1 : 0;
                myneg += ( d < 0.f ) ? 1 : 0;
                const float l     = __builtin_amdgcn_sqrtf( d );
                const float alpha = __builtin_amdgcn_rcpf( l );
                a[k] = ( i == k ) ? l : a[k] * alpha;
                rinv = ( i == k ) ? alpha : rinv;
#pragma unroll
                for ( int j = 0; j < 16; j++ ) {
                    if ( j > k ) {
                        const float ljk = pa_readlane( a[k], j );
                        a[j] -= a[k] * ljk;
                    }
                }
            }
            if ( lane < 16 ) {
#pragma unroll
                for ( int j = 0; j < 16; j++ ) Pp[i][j] = a[j];
                /* column c of the tile, contiguous, 1 / l_cc on the diagonal */
#pragma unroll
                for ( int j = 0; j < 16; j++ )
                    if ( j <= i ) Lt[j][i] = ( j == i ) ? rinv : a[j];
            }
        }
        else if ( p > 0 ) store_panel( p - 1, t - 64 );
        __syncthreads();
        /* (3) panel rows below the diagonal tile: X L^T = A by substitution,
         * column by column (x_j -= x_c l_jc in increasing c, the order of
         * the reference kernel) so the dependent chain is 2 ops per column;
         * column c of the tile is read with 16-byte LDS loads.  Only full
         * panels have rows below the tile. */
        if ( t < m - 16 ) {
            float x[16];
#pragma unroll
            for ( int c = 0; c < 16; c++ ) x[c] = Pp[16 + t][c];
#pragma unroll
            for ( int c = 0; c < 16; c++ ) {
                float cur[16];
#pragma unroll
                for ( int q = c / 4; q < 4; q++ ) {
                    const float4 v = *reinterpret_cast<const float4 *>( &Lt[c][4 * q] );
                    cur[4 * q] = v.x; cur[4 * q + 1] = v.y; cur[4 * q + 2] = v.z; cur[4 * q + 3] = v.w;
                }
                x[c] *= cur[c];
#pragma unroll
                for ( int j = 0; j < 16; j++ )
                    if ( j > c ) x[j] -= x[c] * cur[j];
                __builtin_amdgcn_sched_barrier( 0 );
            }
#pragma unroll
            for ( int c = 0; c < 16; c++ ) Pp[16 + t][c] = x[c];
        }
        __syncthreads();
        /* (5) trailing update of the owned tiles right of the panel:
         * C(i, j)^T -= L(j, p) L(i, p)^T */
#pragma unroll
        for ( int s = 0; s < PA_DREG_S; s++ ) {
            if ( tj[s] > p ) {
                const int32_t ra = 16 * ( tj[s] - p ) + ( lane & 15 ), rb = 16 * ( ti[s] - p ) + ( lane & 15 );
#pragma unroll
                for ( int q = 0; q < 4; q++ ) {
                    const float a = -Pp[ra][4 * q + ( lane >> 4 )];
                    const float b = Pp[rb][4 * q + ( lane >> 4 )];
                    acc[s] = pa_mfma<float>::run( a, b, acc[s] );
                }
            }
            /* keep each slot's operand loads next to its MFMAs: hoisting
             * them over all slots exceeds the register budget */
            __builtin_amdgcn_sched_barrier( 0 );
        }
    }
    __syncthreads();
    store_panel( nt - 1, t );
    if ( t == 0 ) {
        if ( mypiv ) atomicAdd( npiv, mypiv );
        if ( myneg ) atomicAdd( nneg, myneg );
    }
}

__global__ __launch_bounds__( 1024 ) void
k_diag_potrf_reg( pa_dstruct S, const int32_t *cblks, float *Lc, const double *critp, int32_t *npiv,
                  int32_t *nneg )
{
    const int32_t k = cblks[blockIdx.x];
    pa_diag_potrf_reg( Lc + S.c_coef[k], S.c_ncol[k], *critp, npiv, nneg );
}

__global__ __launch_bounds__( 1024 ) void
k_diag_potrf_reg_one( float *D, int32_t nc, double crit, int32_t *npiv, int32_t *nneg )
{
    pa_diag_potrf_reg( D, nc, crit, npiv, nneg );
}

/* ------------------------------------------------------------------------- */
/* Narrow supernodes (LL^t): one wavefront factors a whole cblk of at most   */
/* PA_SM_NC columns and PA_SM_ROWS rows and applies its update, fusing the   */
/* work of core_zpotrfsp (kernels/core_zpotrfsp.c:68-107), the panel TRSM    */
/* (kernels/core_ztrsmsp.c:149-180) and the scatter-GEMM of                  */
/* core_zgemmsp_2d2d (kernels/core_zgemmsp.c:404-481) for the narrow cblks   */
/* that dominate the lower levels of the elimination tree (2D problems,      */
/* batched systems): one 256-thread workgroup per cblk there would leave     */
/* most lanes idle.                                                          */
/*                                                                           */
/* Lane l holds panel rows l, l+64, l+128 (diagonal rows first) in           */
/* registers; column j is eliminated right-looking over the whole panel:     */
/* static pivot check and sqrt on the diagonal, the rows below divided by    */
/* it, then the later columns updated with L[c][j] broadcast by readlane.    */
/* The update C[a][b] = sum_c X[a][c] X[b][c] (a >= b, off-diagonal rows)    */
/* is scattered into the facing cblks with atomic adds: the target row of    */
/* each lane is found by a binary search over the facing cblk's blocks.      */
/* ------------------------------------------------------------------------- */
#define PA_SM_NC   16
#define PA_SM_ROWS 192

template <typename T, int NS, bool PAIR = false>
__global__ __launch_bounds__( 256 ) void
k_small_llt( pa_dstruct S, const int32_t *cblks, int32_t ncb, T *Lc, const double *critp, int32_t *npiv,
             int32_t *nneg, int32_t *nerr )
{
    /* PAIR: the two half wavefronts take one cblk of at most 32 rows each
     * (hl: lane in the half, hb: first lane of the half); broadcasts stay
     * inside a half, and the update loops below run per half.  A half
     * without a cblk (odd count) has nc = stride = 0 and no blocks. */
    static_assert( !PAIR || NS == 1, "two cblks per wavefront: one row slot of 32" );
    const int     lane = threadIdx.x & 63;
    const int     hl = PAIR ? ( lane & 31 ) : lane, hb = PAIR ? ( lane & 32 ) : 0;
    const int32_t w0   = ( (int32_t)blockIdx.x * 4 + ( threadIdx.x >> 6 ) ) * ( PAIR ? 2 : 1 );
    if ( w0 >= ncb ) return; /* whole waves only; no workgroup barrier below */
    const int32_t idx = w0 + ( PAIR ? ( lane >> 5 ) : 0 );
    const bool    act = idx < ncb;
    const int32_t k = cblks[act ? idx : w0];
    const int32_t nc = act ? S.c_ncol[k] : 0, stride = act ? S.c_stride[k] : 0, R = stride - nc;
    const int64_t coef = S.c_coef[k], rb0 = S.r_base[k];
    const T       tcrit = (T)*critp;
    auto bcast = [&]( T v, int j ) -> T { return PAIR ? __shfl( v, hb + j, 64 ) : pa_readlane( v, j ); };
    /* column loop bound: nc, or with two cblks the larger width (the
     * columns past a cblk's own nc are scratch, cleared after the chain) */
    const int32_t ncl = PAIR ? __builtin_amdgcn_readfirstlane( max( nc, __shfl_xor( nc, 32, 64 ) ) ) : nc;

    T       P[NS][PA_SM_NC];
    int64_t rbase[NS];
    int32_t rld[NS], grow[NS];
    bool    valid[NS];
#pragma unroll
    for ( int q = 0; q < NS; q++ ) {
        const int32_t r = hl + 64 * q;
        valid[q] = r < stride;
        rbase[q] = coef + r;
        rld[q]   = nc;
        grow[q]  = 0;
        if ( valid[q] && r >= nc ) {
            const int4 ri = S.r_info[rb0 + ( r - nc )];
            rbase[q] = pa_rinfo_off( ri );
            rld[q]   = ri.y;
            grow[q]  = ri.x;
        }
#pragma unroll
        for ( int c = 0; c < PA_SM_NC; c++ )
            P[q][c] = ( valid[q] && c < nc ) ? Lc[rbase[q] + (int64_t)c * rld[q]] : T( 0 );
    }

    /* panel Cholesky, right-looking over the nc columns */
    int32_t mypiv = 0, myneg = 0;
#pragma unroll
    for ( int j = 0; j < PA_SM_NC; j++ ) {
        if ( j < ncl ) {
            const bool jv = j < nc; /* a column of this lane's cblk */
            T d = bcast( P[0][j], j );
            const bool rep = jv && pa_abs( d ) < tcrit;
            if ( rep ) d = tcrit;
            mypiv += rep ? 1 : 0;
            myneg += ( jv && d < T( 0 ) ) ? 1 : 0;
            if ( !jv ) d = T( 1 ); /* scratch column: kept finite */
            d = sqrt( d );
            /* column scaled by the pivot's reciprocal (the reference kernel's
             * order); selects instead of per-lane branches */
            const T alpha = T( 1 ) / d;
#pragma unroll
            for ( int q = 0; q < NS; q++ ) {
                const int32_t r = hl + 64 * q;
                P[q][j] = ( r == j ) ? d : ( ( r > j ) ? P[q][j] * alpha : P[q][j] );
            }
#pragma unroll
            for ( int c = 1; c < PA_SM_NC; c++ ) {
                if ( c > j && c < ncl ) {
                    const T f = bcast( P[0][j], c ); /* L[c][j] */
#pragma unroll
                    for ( int q = 0; q < NS; q++ )
                        P[q][c] = ( hl + 64 * q > j ) ? P[q][c] - P[q][j] * f : P[q][c];
                }
            }
        }
    }
    if ( PAIR ) {
#pragma unroll
        for ( int c = 0; c < PA_SM_NC; c++ )
            if ( c >= nc ) P[0][c] = T( 0 );
    }
#pragma unroll
    for ( int q = 0; q < NS; q++ ) {
        const int32_t r = hl + 64 * q;
        if ( valid[q] ) {
#pragma unroll
            for ( int c = 0; c < PA_SM_NC; c++ )
                if ( c < nc && ( r >= nc || c <= r ) ) Lc[rbase[q] + (int64_t)c * rld[q]] = P[q][c];
        }
    }
    if ( hl == 0 && act ) {
        if ( mypiv ) atomicAdd( npiv, mypiv );
        if ( myneg ) atomicAdd( nneg, myneg );
    }

    /* scatter update into the facing cblks, one group of consecutive blocks
     * facing the same cblk F at a time; the rows of the blocks from the
     * group on are written into the blocks of F that hold them. */
    int32_t bidx[NS];
#pragma unroll
    for ( int q = 0; q < NS; q++ ) {
        const int32_t r = hl + 64 * q;
        bidx[q] = ( valid[q] && r >= nc ) ? S.r_blk[rb0 + ( r - nc )] : -1;
    }
    const int32_t lblk = act ? S.c_fblok[k + 1] : S.c_fblok[k] + 1;
    int32_t       gb   = S.c_fblok[k] + 1;
    while ( gb < lblk ) {
        const int32_t F = S.b_fcblk[gb];
        int32_t       ge = gb;
        while ( ge + 1 < lblk && S.b_fcblk[ge + 1] == F ) ge++;
        const int32_t fcolF = S.c_fcol[F], ncF = S.c_ncol[F], flast = S.c_fblok[F + 1] - 1;
        const int64_t cF = S.c_coef[F];
        int64_t tbase[NS];
        int32_t tld[NS];
        bool    tok[NS];
#pragma unroll
        for ( int q = 0; q < NS; q++ ) {
            tbase[q] = 0;
            tld[q]   = 0;
            tok[q]   = false;
        }
        /* each lane finds the block of F holding its row by a binary
         * search over F's blocks (a few loads in parallel, not a walk of
         * one dependent scalar load per block) */
        const int32_t ffb = S.c_fblok[F];
#pragma unroll
        for ( int q = 0; q < NS; q++ ) {
            if ( bidx[q] >= gb ) {
                const int32_t f   = pa_find_globrow( S.b_frow, ffb, flast, grow[q] );
                const int32_t ffr = S.b_frow[f], flr = S.b_lrow[f];
                if ( grow[q] >= ffr && grow[q] <= flr ) {
                    tok[q]   = true;
                    tbase[q] = cF + (int64_t)S.b_rowoff[f] * ncF - ffr + grow[q];
                    tld[q]   = flr - ffr + 1;
                }
            }
        }
        /* a row of these blocks outside F's structure (missing fill) is
         * not written: counted for the host to report */
#pragma unroll
        for ( int q = 0; q < NS; q++ )
            if ( bidx[q] >= gb && !tok[q] ) atomicAdd( nerr, 1 );
        /* the columns of F updated by this group: the rows of blocks gb..ge */
        const int32_t b0 = S.b_rowoff[gb], b1 = S.b_rowoff[ge] + ( S.b_lrow[ge] - S.b_frow[ge] + 1 );
        for ( int32_t rbrow = b0; rbrow < b1; rbrow++ ) {
            const int32_t lb = rbrow & 63, qb = rbrow >> 6;
            /* row rbrow broadcast from its lane (uniform branch on its register slot) */
            T       xb[PA_SM_NC];
            int32_t growb;
            if ( PAIR ) {
                /* per half: rbrow differs between the halves */
#pragma unroll
                for ( int c = 0; c < PA_SM_NC; c++ ) xb[c] = __shfl( P[0][c], hb + lb, 64 );
                growb = __shfl( grow[0], hb + lb, 64 );
            }
            else if ( qb == 0 ) {
#pragma unroll
                for ( int c = 0; c < PA_SM_NC; c++ ) xb[c] = pa_readlane( P[0][c], lb );
                growb = __builtin_amdgcn_readlane( grow[0], lb );
            }
            else if ( NS > 1 && qb == 1 ) {
#pragma unroll
                for ( int c = 0; c < PA_SM_NC; c++ ) xb[c] = pa_readlane( P[NS > 1 ? 1 : 0][c], lb );
                growb = __builtin_amdgcn_readlane( grow[NS > 1 ? 1 : 0], lb );
            }
            else {
#pragma unroll
                for ( int c = 0; c < PA_SM_NC; c++ ) xb[c] = pa_readlane( P[NS - 1][c], lb );
                growb = __builtin_amdgcn_readlane( grow[NS - 1], lb );
            }
            const int32_t col = growb - fcolF;
#pragma unroll
            for ( int q = 0; q < NS; q++ ) {
                const int32_t r = hl + 64 * q;
                if ( tok[q] && r >= rbrow ) {
                    T acc = T( 0 );
#pragma unroll
                    for ( int c = 0; c < PA_SM_NC; c++ ) acc += P[q][c] * xb[c];
                    pa_atomic_add( Lc + tbase[q] + (int64_t)col * tld[q], -acc );
                }
            }
        }
        gb = ge + 1;
    }
}

/* ------------------------------------------------------------------------- */
/* Panel TRSM on the off-diagonal rows of a cblk (reference                   */
/* kernels/core_ztrsmsp.c:149-180):                                          */
/*   MODE 0 (POTRF):  X <- X L^{-T}, L lower non-unit diagonal tile          */
/*   MODE 1 (LU, L):  X <- X U^{-1}, U upper non-unit part of the L tile     */
/*   MODE 2 (LU, U):  X <- X V^{-1}, V upper unit part of the U^T tile       */
/* A task is (cblk, first below-diagonal row); TR rows per workgroup are     */
/* staged in LDS and solved with a column-blocked algorithm.  The level      */
/* factorization uses it above 384 columns only (and the reference-signature */
/* gpublok_*trsmsp calls); below, k_trsm_inv replaces it.                    */
/* ------------------------------------------------------------------------- */
/* rows per workgroup (the slab lives in LDS) */
#ifndef PA_TRSM_TR
#define PA_TRSM_TR 32
#endif
template <typename T> struct pa_trsm_rows_per_wg { static constexpr int value = PA_TRSM_TR; };

/*
 * Solve `rows` (<= SR) rows whose offsets/leading dimensions inside `panel`
 * are xoff/xld (in LDS); D is the nc x nc diagonal tile.  The slab X
 * (SR x nc) is staged in LDS; columns are processed in blocks of 16:
 *   X[:, jb:jb+16] -= X[:, 0:jb] C[0:jb, jb:jb+16]   (MFMA, K split among
 *                                                     waves when SR < 64)
 *   X[:, jb:jb+16] <- X[:, jb:jb+16] C[jb:jb+16, jb:jb+16]^{-1} (one row per thread)
 * where C(l, j) = L[j][l] (MODE 0) or D[l][j] (MODE 1, 2); the coefficients of
 * the block column are staged in LDS first.  LDS: pa_trsm_lds (padded strides).
 */
#define PA_TRSM_U 8
/* LDS strides of pa_trsm_rows, padded so that the MFMA operand reads and the
 * partial-sum writes hit distinct banks (the unpadded layout spent more
 * cycles in bank conflicts than in LDS instructions): FP32 reads 32 lanes at
 * a time (stride = 2 mod 32 words), FP64 16 lanes (stride = 1 mod 16 doubles) */
template <typename T> struct pa_trsm_pad;
template <> struct pa_trsm_pad<float> {
    static __host__ __device__ constexpr int32_t ld( int32_t n ) { return n + ( ( 2 - n ) & 31 ); }
    static constexpr int RS = PA_TRSM_TR + 4; /* 4 RS = 16 mod 32 */
};
template <> struct pa_trsm_pad<double> {
    static __host__ __device__ constexpr int32_t ld( int32_t n ) { return n + ( ( 1 - n ) & 15 ); }
    static constexpr int RS = PA_TRSM_TR;
};

template <typename T, int MODE>
__device__ void
pa_trsm_rows( T *panel, const T *D, int32_t nc, int32_t rows, const int64_t *xoff, const int32_t *xld, T *smem )
{
    typedef typename pa_mfma<T>::acc_t acc_t;
    constexpr int SR  = pa_trsm_rows_per_wg<T>::value;
    constexpr int NRT = SR / 16;   /* row tiles                  */
    constexpr int KS  = 4 / NRT;   /* K split among the waves    */
    constexpr int RS  = pa_trsm_pad<T>::RS;
    const int32_t XT  = pa_trsm_pad<T>::ld( nc );   /* row stride of Xs and Ls */
    T *Xs  = smem;                                  /* [SR][XT]: Xs[r*XT + j] = X(r, j)        */
    T *Ls  = Xs + (int64_t)SR * XT;                 /* [16][XT]: Ls[jj*XT + l] = C(l, jb+jj)   */
    T *red = Ls + 16 * XT;                          /* [KS][16][RS] partial sums               */
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;

    /* global <-> LDS staging loops issue PA_TRSM_U independent loads before
     * their stores, so the load latency is paid once per batch */
    for ( int32_t i0 = t; i0 < nc * SR; i0 += 256 * PA_TRSM_U ) {
        T v[PA_TRSM_U];
#pragma unroll
        for ( int u = 0; u < PA_TRSM_U; u++ ) {
            const int32_t idx = i0 + 256 * u, r = idx % SR, j = idx / SR;
            v[u] = ( idx < nc * SR && r < rows ) ? panel[xoff[r] + (int64_t)j * xld[r]] : T( 0 );
        }
#pragma unroll
        for ( int u = 0; u < PA_TRSM_U; u++ ) {
            const int32_t idx = i0 + 256 * u, r = idx % SR, j = idx / SR;
            if ( idx < nc * SR ) Xs[r * XT + j] = v[u];
        }
    }

    for ( int32_t jb = 0; jb < nc; jb += 16 ) {
        const int32_t cw = nc - jb < 16 ? nc - jb : 16;
        const int32_t kend = jb + cw; /* coefficients needed: l < jb + cw */
        /* stage C(l, jb+jj) for l < kend */
        for ( int32_t i0 = t; i0 < 16 * kend; i0 += 256 * PA_TRSM_U ) {
            T       v[PA_TRSM_U];
            int32_t dst[PA_TRSM_U];
#pragma unroll
            for ( int u = 0; u < PA_TRSM_U; u++ ) {
                const int32_t idx = i0 + 256 * u;
                int32_t jj, l;
                if ( MODE == 0 ) { jj = idx % 16; l = idx / 16; }  /* L[jb+jj][l]: contiguous in jj */
                else             { l = idx % kend; jj = idx / kend; } /* D[l][jb+jj]: contiguous in l  */
                const bool in = idx < 16 * kend;
                dst[u] = in ? jj * XT + l : -1;
                v[u]   = ( in && jj < cw ) ? ( MODE == 0 ? D[( jb + jj ) + (int64_t)l * nc] : D[l + (int64_t)( jb + jj ) * nc] )
                                           : T( 0 );
            }
#pragma unroll
            for ( int u = 0; u < PA_TRSM_U; u++ )
                if ( dst[u] >= 0 ) Ls[dst[u]] = v[u];
        }
        __syncthreads();
        if ( jb > 0 ) {
            /* partial products on MFMA: wave -> (row tile, K part) */
            const int rt = wave % NRT, ks = wave / NRT;
            const int32_t kchunk = ( ( jb / KS + 3 ) / 4 ) * 4;
            const int32_t kb = ks * kchunk, ke = min( jb, kb + kchunk );
            acc_t acc;
            for ( int r = 0; r < 4; r++ ) acc[r] = T( 0 );
            for ( int32_t k0 = kb; k0 < ke; k0 += 4 ) {
                const int32_t kk = k0 + ( lane >> 4 );
                T ca = ( kk < ke ) ? Ls[( lane & 15 ) * XT + kk] : T( 0 );           /* column side (j) */
                T ra = ( kk < ke ) ? Xs[( rt * 16 + ( lane & 15 ) ) * XT + kk] : T( 0 ); /* row side (r)    */
                acc = pa_mfma<T>::run( ca, ra, acc );
            }
            /* lane holds (row rt*16 + lane%16, column j = row(lane, r)) */
            for ( int r = 0; r < 4; r++ ) {
                int j = pa_mfma<T>::row( lane, r );
                red[( ks * 16 + j ) * RS + rt * 16 + ( lane & 15 )] = acc[r];
            }
        }
        __syncthreads();
        if ( t < SR ) {
            const int32_t r = t;
            T x[16];
#pragma unroll
            for ( int jj = 0; jj < 16; jj++ ) {
                T v = ( jj < cw ) ? Xs[r * XT + jb + jj] : T( 0 );
                if ( jb > 0 ) {
#pragma unroll
                    for ( int ks = 0; ks < KS; ks++ ) v -= red[( ks * 16 + jj ) * RS + r];
                }
                x[jj] = v;
            }
#pragma unroll
            for ( int jj = 0; jj < 16; jj++ ) {
                if ( jj < cw ) {
                    T v = x[jj];
#pragma unroll
                    for ( int l = 0; l < jj; l++ ) v -= x[l] * Ls[jj * XT + jb + l];
                    if ( MODE != 2 ) v /= Ls[jj * XT + jb + jj];
                    x[jj] = v;
                    Xs[r * XT + jb + jj] = v;
                }
            }
        }
        __syncthreads();
    }
    for ( int32_t idx = t; idx < nc * SR; idx += 256 ) {
        int32_t r = idx % SR, j = idx / SR;
        if ( r < rows ) panel[xoff[r] + (int64_t)j * xld[r]] = Xs[r * XT + j];
    }
}
#undef PA_TRSM_U

/* LDS bytes of pa_trsm_rows for nc columns */
template <typename T>
__host__ __device__ constexpr size_t
pa_trsm_lds( int32_t nc )
{
    return ( (size_t)pa_trsm_rows_per_wg<T>::value * pa_trsm_pad<T>::ld( nc ) + 16 * (size_t)pa_trsm_pad<T>::ld( nc ) +
             4 * 16 * (size_t)pa_trsm_pad<T>::RS ) * sizeof( T );
}

template <typename T, int MODE>
__global__ __launch_bounds__( 256 ) void
k_trsm( pa_dstruct S, const int2 *tasks, T *Lc, T *Uc )
{
    constexpr int SR = pa_trsm_rows_per_wg<T>::value;
    extern __shared__ unsigned char pa_smem_raw[];
    __shared__ int64_t xoff[SR];
    __shared__ int32_t xld[SR];

    int2    task = tasks[blockIdx.x];
    int32_t k = task.x, r0 = task.y;
    int32_t nc = S.c_ncol[k];
    int32_t R  = S.c_stride[k] - nc;
    int32_t rows = R - r0 < SR ? R - r0 : SR;
    T      *panel = ( MODE == 2 ? Uc : Lc ) + S.c_coef[k];
    int t = threadIdx.x;

    if ( t < SR ) {
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

/* ------------------------------------------------------------------------- */
/* TRSM by inverted diagonal tiles (the scheme of the rocBLAS and MAGMA       */
/* TRSMs).  The rows X below the diagonal block of a cblk solve X T = B with  */
/* T upper triangular, T(l, c) = Q[c + l nc] for the diagonal block Q that    */
/* holds it contiguously in c: L^T of LL^t (Q = L), U of LU (Q = the U^T      */
/* copy) and the unit L^T of LU (Q = L).  k_tile_inv inverts the 16 x 16      */
/* diagonal tiles of T once per level (one wave per tile); k_trsm_inv keeps   */
/* 32 rows of X in MFMA accumulators and sweeps the column tiles left to      */
/* right: X_j = X_j T_jj^{-1} by the wave that owns column tile j, then       */
/* X_c -= X_j T_jc for the later column tiles c by their owners.  The        */
/* accumulators hold X^T tiles, so the 16 lanes of an output row are 16      */
/* consecutive rows of X (coalesced loads and stores), and an accumulator    */
/* tile is already the B operand of the next 16x16x4 MFMA when the K index   */
/* of lane group q, step s is the accumulator row row(l, s).                 */
/* ------------------------------------------------------------------------- */
#define PA_TI_NJ 6 /* column tiles per wave: nc <= 4 * 16 * PA_TI_NJ */
#define PA_TI_MAXNC ( 64 * PA_TI_NJ )

template <typename T, int MODE>
__global__ __launch_bounds__( 256 ) void
k_tile_inv( pa_dstruct S, const int2 *tasks, int32_t ntask, const T *Lc, const T *Uc, T *tinv, int64_t tinv2 )
{
    /* MODE 3: both LU modes in one launch, mode 1 + blockIdx.y; mode 2's
     * inverses start tinv2 elements further.  MODE 4 (LDL^t): the tiles of
     * DL^t, stored as in mode 1 */
    const int md = ( MODE == 3 ) ? 1 + (int)blockIdx.y : ( MODE == 4 ? 1 : MODE );
    if ( md == 2 ) tinv += tinv2;
    __shared__ __align__( 16 ) T sT[4][16][16]; /* per wave: sT[c][l] = T(l, c) of the tile */
    const int     lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int32_t task  = blockIdx.x * 4 + wave;
    const bool    valid = task < ntask;
    int32_t k = 0, tl = 0, nc = 16;
    if ( valid ) {
        const int2 tk = tasks[task];
        k  = tk.x;
        tl = tk.y;
        nc = S.c_ncol[k];
    }
    const int32_t c0 = 16 * tl, w = min( 16, nc - c0 );
    const T      *Q  = ( md == 1 ? Uc : Lc ) + S.c_coef[k];
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        const int idx = lane + 64 * u, c = idx >> 4, l = idx & 15;
        /* columns past the cblk: identity */
        T v = ( c == l ) ? T( 1 ) : T( 0 );
        if ( valid && c < w && l <= c && !( md == 2 && l == c ) ) v = Q[( c0 + c ) + (int64_t)( c0 + l ) * nc];
        sT[wave][c][l] = v;
    }
    __syncthreads();
    if ( valid && lane < 16 ) {
        /* row r of M = T^{-1}: sum_{l <= c} M(r, l) T(l, c) = delta_rc */
        const int r = lane;
        T m[16];
#pragma unroll
        for ( int c = 0; c < 16; c++ ) {
            T v = ( r == c ) ? T( 1 ) : T( 0 );
#pragma unroll
            for ( int l = 0; l < 16; l++ )
                if ( l < c ) v -= m[l] * sT[wave][c][l];
            m[c] = ( md == 2 ) ? v : v / sT[wave][c][c];
        }
        T *out = tinv + ( S.c_tinv[k] + tl ) * 256 + r * 16;
#pragma unroll
        for ( int c = 0; c < 16; c++ ) out[c] = m[c];
    }
}

template <typename T, int MODE, int NJ>
__global__ __launch_bounds__( 256 ) void
k_trsm_inv( pa_dstruct S, const int2 *tasks, T *Lc, T *Uc, const T *tinv, int64_t tinv2, T *dinv, int64_t dinv2 )
{
    /* NJ: column tiles per wave, at least the level's widest cblk / 64
     * (the accumulators are sized by it: a narrow level's launch keeps
     * fewer registers per lane and more waves per SIMD) */
    /* MODE 3: both LU modes in one launch, mode 1 + blockIdx.y.  MODE 4
     * (LDL^t, reference cpucblk_zsytrfsp1d_panel, kernels/core_zsytrfsp.c:
     * 296-318): X <- X (DL^t)^{-1} as mode 1, the upper part of the factored
     * diagonal block being DL^t, and the copy X D (cpucblk_zscalo) stored in
     * the second array for the updates */
    const int md = ( MODE == 3 ) ? 1 + (int)blockIdx.y : ( MODE == 4 ? 1 : MODE );
    if ( md == 2 ) tinv += tinv2;
    typedef typename pa_mfma<T>::acc_t acc_t;
    constexpr int SR  = PA_TRSM_TR;
    constexpr int NRT = SR / 16;
    __shared__ int64_t xoff[SR];
    __shared__ int32_t xld[SR];
    __shared__ T       xb[2][NRT][4][64]; /* solved column tile, B-operand layout */
    const int t = threadIdx.x, lane = t & 63, li = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane( t >> 6 );

    /* a task (k, -1 - r0) solves identity rows r0.. instead of panel rows:
     * row r of T^{-1} is stored as column r of the cblk's inverse block, so
     * that block is the column-major inverse of the lower triangle whose
     * transpose is T (mode 1: the U^T array's, into the second region) */
    const int2    task  = tasks[blockIdx.x];
    const bool    ident = task.y < 0;
    const int32_t k = task.x, r0 = ident ? -1 - task.y : task.y;
    const int32_t nc = S.c_ncol[k], R = S.c_stride[k] - nc;
    const int32_t rows = ident ? min( nc - r0, SR ) : ( R - r0 < SR ? R - r0 : SR );
    const int32_t nt = ( nc + 15 ) >> 4;
    T            *panel = ident ? dinv + ( md == 1 ? dinv2 : 0 ) + S.c_dtoff[k] : ( md == 2 ? Uc : Lc ) + S.c_coef[k];
    const T      *Q     = ( md == 1 ? Uc : Lc ) + S.c_coef[k];
    const T      *Mi    = tinv + S.c_tinv[k] * 256;

    if ( t < SR && ident ) {
        xoff[t] = (int64_t)( r0 + min( t, rows - 1 ) ) * nc;
        xld[t]  = 1;
    }
    else if ( t < SR ) {
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

    /* this lane's rows rt * 16 + li; empty rows read X(0, 0) and are zeroed */
    int64_t ro[NRT];
    int32_t rl[NRT];
    bool    rv[NRT];
#pragma unroll
    for ( int rt = 0; rt < NRT; rt++ ) {
        const int row = rt * 16 + li;
        rv[rt] = row < rows;
        ro[rt] = xoff[row];
        rl[rt] = xld[row];
    }
    /* accumulator tile (rt, jj): X^T of rows rt*16.., column tile wave + 4 jj */
    acc_t acc[NRT][NJ];
#pragma unroll
    for ( int jj = 0; jj < NJ; jj++ ) {
        const int32_t j = wave + 4 * jj;
#pragma unroll
        for ( int s = 0; s < 4; s++ ) {
            const int32_t col = 16 * j + pa_mfma<T>::row( lane, s );
            const int32_t cc  = min( col, nc - 1 );
#pragma unroll
            for ( int rt = 0; rt < NRT; rt++ ) {
                if ( ident ) {
                    acc[rt][jj][s] = ( rv[rt] && r0 + rt * 16 + li == col ) ? T( 1 ) : T( 0 );
                }
                else {
                    const T v = panel[ro[rt] + (int64_t)cc * rl[rt]];
                    acc[rt][jj][s] = ( rv[rt] && col < nc ) ? v : T( 0 );
                }
            }
        }
    }
    /* operands of the first step: T_0c tiles of the owned columns, and the
     * inverse of the first owned diagonal tile */
    T a[NJ][4], ai[4];
#pragma unroll
    for ( int jj = 0; jj < NJ; jj++ ) {
        const int32_t j = wave + 4 * jj, col = min( 16 * j + li, nc - 1 );
#pragma unroll
        for ( int s = 0; s < 4; s++ ) a[jj][s] = -Q[col + (int64_t)min( pa_mfma<T>::row( lane, s ), nc - 1 ) * nc];
    }
#pragma unroll
    for ( int s = 0; s < 4; s++ ) ai[s] = Mi[min( wave, nt - 1 ) * 256 + pa_mfma<T>::row( lane, s ) * 16 + li];

    for ( int32_t jb = 0; jb < nt; jb++ ) {
        T( *xs )[4][64] = xb[jb & 1];
        if ( wave == ( jb & 3 ) ) {
            const int jo = jb >> 2;
#pragma unroll
            for ( int jj = 0; jj < NJ; jj++ ) {
                if ( jj == jo ) {
#pragma unroll
                    for ( int rt = 0; rt < NRT; rt++ ) {
                        acc_t x;
                        for ( int r = 0; r < 4; r++ ) x[r] = T( 0 );
#pragma unroll
                        for ( int s = 0; s < 4; s++ ) x = pa_mfma<T>::run( ai[s], acc[rt][jj][s], x );
                        acc[rt][jj] = x;
#pragma unroll
                        for ( int s = 0; s < 4; s++ ) xs[rt][s][lane] = x[s];
                    }
                }
            }
            /* inverse of the next owned diagonal tile */
            const int32_t jn = min( jb + 4, nt - 1 );
#pragma unroll
            for ( int s = 0; s < 4; s++ ) ai[s] = Mi[jn * 256 + pa_mfma<T>::row( lane, s ) * 16 + li];
        }
        __syncthreads();
        T bx[NRT][4];
#pragma unroll
        for ( int rt = 0; rt < NRT; rt++ )
#pragma unroll
            for ( int s = 0; s < 4; s++ ) bx[rt][s] = xs[rt][s][lane];
        const int32_t kn = min( 16 * ( jb + 1 ), nc - 1 ); /* first row of the next step's T tiles */
#pragma unroll
        for ( int jj = 0; jj < NJ; jj++ ) {
            const int32_t j = wave + 4 * jj;
            if ( j > jb && j < nt ) {
#pragma unroll
                for ( int rt = 0; rt < NRT; rt++ )
#pragma unroll
                    for ( int s = 0; s < 4; s++ ) acc[rt][jj] = pa_mfma<T>::run( a[jj][s], bx[rt][s], acc[rt][jj] );
                /* T_(jb+1) c for the next step */
                const int32_t col = min( 16 * j + li, nc - 1 );
#pragma unroll
                for ( int s = 0; s < 4; s++ )
                    a[jj][s] = -Q[col + (int64_t)min( kn + pa_mfma<T>::row( lane, s ), nc - 1 ) * nc];
            }
        }
    }
#pragma unroll
    for ( int jj = 0; jj < NJ; jj++ ) {
        const int32_t j = wave + 4 * jj;
        if ( j < nt ) {
#pragma unroll
            for ( int s = 0; s < 4; s++ ) {
                const int32_t col = 16 * j + pa_mfma<T>::row( lane, s );
                /* D(col) for the LDL^t copy: the diagonal of the factored block */
                const T d = ( MODE == 4 ) ? Q[(int64_t)min( col, nc - 1 ) * ( nc + 1 )] : T( 0 );
#pragma unroll
                for ( int rt = 0; rt < NRT; rt++ )
                    if ( rv[rt] && col < nc ) {
                        panel[ro[rt] + (int64_t)col * rl[rt]] = acc[rt][jj][s];
                        if ( MODE == 4 && !ident ) Uc[S.c_coef[k] + ro[rt] + (int64_t)col * rl[rt]] = acc[rt][jj][s] * d;
                    }
            }
        }
    }
}

/* LDL^t panel copy X D of the rows below the diagonal block (reference
 * cpucblk_zscalo, kernels/core_zscalo.c), after the substitution TRSM of
 * the widest cblks: one TRSM slab (cblk, 32 rows) per workgroup, D the
 * diagonal of the factored diagonal block of Lc. */
template <typename T>
__global__ __launch_bounds__( 256 ) void
k_ldlt_scalo( pa_dstruct S, const int2 *tasks, const T *Lc, T *Uc )
{
    const int2    task = tasks[blockIdx.x];
    const int32_t k = task.x, r0 = task.y;
    const int32_t nc = S.c_ncol[k], R = S.c_stride[k] - nc;
    const int32_t rows = min( R - r0, (int32_t)PA_TRSM_TR );
    const int64_t cb = S.c_coef[k], rb = S.r_base[k];
    for ( int32_t e = threadIdx.x; e < rows * nc; e += 256 ) {
        const int32_t r = e % rows, c = e / rows;
        const int4    ri = S.r_info[rb + r0 + r];
        const int64_t o  = pa_rinfo_off( ri ) + (int64_t)c * ri.y;
        Uc[o] = Lc[o] * Lc[cb + (int64_t)c * ( nc + 1 )];
    }
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

/* TRSM over the rows of a run of blocks given by value; one workgroup per SR rows. */
template <typename T, int MODE>
__global__ __launch_bounds__( 256 ) void
k_trsm_call( T *C, const T *D, int32_t nc, pa_call_rows R )
{
    constexpr int SR = pa_trsm_rows_per_wg<T>::value;
    extern __shared__ unsigned char pa_smem_raw[];
    __shared__ int64_t xoff[SR];
    __shared__ int32_t xld[SR];
    int32_t total = R.pref[R.nblk];
    int32_t r0 = blockIdx.x * SR;
    int32_t rows = total - r0 < SR ? total - r0 : SR;
    int t = threadIdx.x;
    if ( t < SR ) {
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
/* product runs on MFMA 16x16x4 (FP32 32 x 32 wave blocks: 32x32x2) and the */
/* epilogue writes into F.  The product is computed transposed (sigma x rho) */
/* so that the lanes of an MFMA output row hold 16 (32x32x2: 32) consecutive */
/* rows of one target column: the scatter store is coalesced; 64 x 64 FP32   */
/* tiles go through an LDS copy of the output so that an instruction covers  */
/* 64 consecutive rows.                                                      */
/* ------------------------------------------------------------------------- */
#define PA_GT  64
#ifndef PA_GKC
#define PA_GKC 16
#endif


/* Raw buffer loads at a 32-bit byte offset from a wave-uniform base (the
 * resource lives in scalar registers, so a lane keeps one 32-bit offset per
 * access instead of a 64-bit address). */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t
pa_buf_rsrc( const void *base )
{
    return __builtin_amdgcn_make_buffer_rsrc( const_cast<void *>( base ), 0, -1, 0x00020000 );
}
__device__ __forceinline__ float
pa_buf_load( __amdgpu_buffer_rsrc_t r, uint32_t off, float * )
{
    return __builtin_bit_cast( float, __builtin_amdgcn_raw_buffer_load_b32( r, off, 0, 0 ) );
}
__device__ __forceinline__ double
pa_buf_load( __amdgpu_buffer_rsrc_t r, uint32_t off, double * )
{
    return __builtin_bit_cast( double, __builtin_amdgcn_raw_buffer_load_b64( r, off, 0, 0 ) );
}

/*
 * Tile core shared by the level-batched kernel and the reference-signature
 * entry points.  TM x TN output (rho x sigma: 128 x 128, 128 x 64, 128 x 32,
 * 64 x 64, 64 x 32), 4 waves in a 2 x 2 arrangement, K in chunks of PA_GKC staged in LDS, the next chunk prefetched
 * into registers while the current one feeds the MFMAs (deeper chunks cost
 * more registers and LDS than their longer MFMA runs win back).  The loads are
 * unconditional from clamped in-range addresses (empty rows are zeroed when
 * the chunk is stored to LDS), so no per-load branch holds the prefetch back.
 * Row maps (in LDS)
 * must be filled by the caller:
 *   a_off/a_ld: source of the TM rho rows inside Asrc (a_ld = 0: empty row)
 *   b_off/b_ld: source of the TN sigma rows inside Bsrc
 *   c_off/c_ld: target row base and leading dimension for each rho row
 *   c_col:      target column of each sigma row
 *   a_blk/b_blk: block index of each row for the i >= j (+shift) mask
 */
template <typename T, int TM, int TN, bool BUF, bool XP = false>
__device__ __forceinline__ void
pa_gemm_tile( const T *__restrict__ Asrc, const T *__restrict__ Bsrc, T *C, int32_t K,
              const int64_t *a_off, const int32_t *a_ld, const int64_t *b_off, const int32_t *b_ld,
              const int64_t *c_off, const int32_t *c_ld, const int32_t *c_col,
              const int32_t *a_blk, const int32_t *b_blk, int32_t shift, bool atomic,
              T ( *As )[TM + pa_gpad<T>::value], T ( *Bs )[TN + pa_gpad<T>::value], T *Cs = nullptr )
{
    typedef typename pa_mfma<T>::acc_t acc_t;
    constexpr int NBM = TM / 32;          /* MFMA blocks per wave, rho          */
    constexpr int NBN = TN / 32;          /* MFMA blocks per wave, sigma        */
    constexpr int KLA = 256 / TM;         /* k lanes of the A loader            */
    constexpr int KLB = 256 / TN;         /* k lanes of the B loader            */
    constexpr int NQA = PA_GKC / KLA;     /* A loads per thread per chunk       */
    constexpr int NQB = PA_GKC / KLB;     /* B loads per thread per chunk       */
    const int t    = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm   = wave & 1;  /* rho half   */
    const int wn   = wave >> 1; /* sigma half */
    /* FP32 with wave blocks of at least 32 x 32: 32x32x2 MFMAs, whose output
     * puts 32 consecutive rho rows (consecutive target addresses) in the 32
     * lanes of a half wave, so each epilogue instruction covers two target
     * columns of 128 bytes instead of four of 64 */
    constexpr bool M32 = sizeof( T ) == 4 && TM >= 64 && TN >= 64;
    constexpr int NBM32 = M32 ? TM / 64 : 1, NBN32 = M32 ? TN / 64 : 1;
    acc_t acc[M32 ? 1 : NBN][M32 ? 1 : NBM];
    pa_f16 acc32[NBN32][NBM32];
    if constexpr ( M32 ) {
#pragma unroll
        for ( int i = 0; i < NBN32; i++ )
#pragma unroll
            for ( int j = 0; j < NBM32; j++ )
#pragma unroll
                for ( int r = 0; r < 16; r++ ) acc32[i][j][r] = 0.f;
    }
    else {
#pragma unroll
        for ( int i = 0; i < NBN; i++ )
#pragma unroll
            for ( int j = 0; j < NBM; j++ )
#pragma unroll
                for ( int r = 0; r < 4; r++ ) acc[i][j][r] = T( 0 );
    }

    const int arow = t % TM, ak = t / TM, brow = t % TN, bk = t / TN;
    const int32_t ald = a_ld[arow], bld = b_ld[brow];
    /* an empty row reads row 0 of its operand (in range) */
    const int64_t aof = ald > 0 ? a_off[arow] : 0, bof = bld > 0 ? b_off[brow] : 0;
    const int32_t alds = ald > 0 ? ald : 1, blds = bld > 0 ? bld : 1;
    T pa[NQA], pb[NQB];
    /* Full K chunks read through per-thread column pointers advanced once per
     * chunk (no clamp, index product or K mask per load); the one partial
     * chunk at the end of K, if any, uses clamped in-range addresses and is
     * masked when stored. */
    const int32_t nfull = K / PA_GKC;
    /* BUF: both operands lie in one source cblk, less than 2^32 bytes past
     * Asrc / Bsrc (checked when the schedule is built), and are read with
     * raw buffer loads at 32-bit byte offsets: one register per load
     * instead of a 64-bit pointer */
    const T *pA[BUF ? 1 : NQA], *pB[BUF ? 1 : NQB];
    uint32_t oA[BUF ? NQA : 1], oB[BUF ? NQB : 1];
    const __amdgpu_buffer_rsrc_t rA = pa_buf_rsrc( Asrc ), rB = pa_buf_rsrc( Bsrc );
    if ( BUF ) {
#pragma unroll
        for ( int q = 0; q < NQA; q++ ) oA[q] = (uint32_t)( ( aof + (int64_t)( ak + KLA * q ) * alds ) * (int64_t)sizeof( T ) );
#pragma unroll
        for ( int q = 0; q < NQB; q++ ) oB[q] = (uint32_t)( ( bof + (int64_t)( bk + KLB * q ) * blds ) * (int64_t)sizeof( T ) );
    }
    else {
#pragma unroll
        for ( int q = 0; q < NQA; q++ ) pA[BUF ? 0 : q] = Asrc + aof + (int64_t)( ak + KLA * q ) * alds;
#pragma unroll
        for ( int q = 0; q < NQB; q++ ) pB[BUF ? 0 : q] = Bsrc + bof + (int64_t)( bk + KLB * q ) * blds;
    }
    const int64_t ca = (int64_t)PA_GKC * alds, cb = (int64_t)PA_GKC * blds;
    const uint32_t cab = (uint32_t)( ca * (int64_t)sizeof( T ) ), cbb = (uint32_t)( cb * (int64_t)sizeof( T ) );
    auto load_full = [&]() {
        if ( BUF ) {
#pragma unroll
            for ( int q = 0; q < NQA; q++ ) pa[q] = pa_buf_load( rA, oA[BUF ? q : 0], (T *)nullptr );
#pragma unroll
            for ( int q = 0; q < NQB; q++ ) pb[q] = pa_buf_load( rB, oB[BUF ? q : 0], (T *)nullptr );
        }
        else {
#pragma unroll
            for ( int q = 0; q < NQA; q++ ) pa[q] = *pA[BUF ? 0 : q];
#pragma unroll
            for ( int q = 0; q < NQB; q++ ) pb[q] = *pB[BUF ? 0 : q];
        }
    };
    auto load_tail = [&]( int32_t k0 ) {
        if ( BUF ) {
#pragma unroll
            for ( int q = 0; q < NQA; q++ )
                pa[q] = pa_buf_load( rA, (uint32_t)( ( aof + (int64_t)min( k0 + ak + KLA * q, K - 1 ) * alds ) * (int64_t)sizeof( T ) ),
                                     (T *)nullptr );
#pragma unroll
            for ( int q = 0; q < NQB; q++ )
                pb[q] = pa_buf_load( rB, (uint32_t)( ( bof + (int64_t)min( k0 + bk + KLB * q, K - 1 ) * blds ) * (int64_t)sizeof( T ) ),
                                     (T *)nullptr );
        }
        else {
#pragma unroll
            for ( int q = 0; q < NQA; q++ ) pa[q] = Asrc[aof + (int64_t)min( k0 + ak + KLA * q, K - 1 ) * alds];
#pragma unroll
            for ( int q = 0; q < NQB; q++ ) pb[q] = Bsrc[bof + (int64_t)min( k0 + bk + KLB * q, K - 1 ) * blds];
        }
    };
    if ( nfull > 0 ) load_full();
    else             load_tail( 0 );
    for ( int32_t c = 0, k0 = 0; k0 < K; c++, k0 += PA_GKC ) {
        if ( c < nfull ) {
            /* an empty row (ld 0) holds finite values of row 0 of its
             * operand here: its products land in output rows / columns the
             * epilogue skips, so no mask is needed */
#pragma unroll
            for ( int q = 0; q < NQA; q++ ) As[ak + KLA * q][arow] = pa[q];
#pragma unroll
            for ( int q = 0; q < NQB; q++ ) Bs[bk + KLB * q][brow] = pb[q];
        }
        else {
#pragma unroll
            for ( int q = 0; q < NQA; q++ ) As[ak + KLA * q][arow] = ( ald > 0 && k0 + ak + KLA * q < K ) ? pa[q] : T( 0 );
#pragma unroll
            for ( int q = 0; q < NQB; q++ ) Bs[bk + KLB * q][brow] = ( bld > 0 && k0 + bk + KLB * q < K ) ? pb[q] : T( 0 );
        }
        __syncthreads();
        if ( k0 + PA_GKC < K ) {
            if ( BUF ) {
#pragma unroll
                for ( int q = 0; q < NQA; q++ ) oA[BUF ? q : 0] += cab;
#pragma unroll
                for ( int q = 0; q < NQB; q++ ) oB[BUF ? q : 0] += cbb;
            }
            else {
#pragma unroll
                for ( int q = 0; q < NQA; q++ ) pA[BUF ? 0 : q] += ca;
#pragma unroll
                for ( int q = 0; q < NQB; q++ ) pB[BUF ? 0 : q] += cb;
            }
            if ( c + 1 < nfull ) load_full();
            else                 load_tail( k0 + PA_GKC );
        }
        if constexpr ( M32 ) {
#pragma unroll
            for ( int ks = 0; ks < PA_GKC / 2; ks++ ) {
                int kk = ks * 2 + ( lane >> 5 );
                float xa[NBN32], yb[NBM32];
#pragma unroll
                for ( int mi = 0; mi < NBN32; mi++ ) xa[mi] = Bs[kk][wn * ( TN / 2 ) + mi * 32 + ( lane & 31 )];
#pragma unroll
                for ( int mj = 0; mj < NBM32; mj++ ) yb[mj] = As[kk][wm * ( TM / 2 ) + mj * 32 + ( lane & 31 )];
#pragma unroll
                for ( int mi = 0; mi < NBN32; mi++ )
#pragma unroll
                    for ( int mj = 0; mj < NBM32; mj++ )
                        acc32[mi][mj] = __builtin_amdgcn_mfma_f32_32x32x2f32( xa[mi], yb[mj], acc32[mi][mj], 0, 0, 0 );
            }
        }
        else {
#pragma unroll
            for ( int ks = 0; ks < PA_GKC / 4; ks++ ) {
                int kk = ks * 4 + ( lane >> 4 );
                T xa[NBN], yb[NBM];
#pragma unroll
                for ( int mi = 0; mi < NBN; mi++ ) xa[mi] = Bs[kk][wn * ( TN / 2 ) + mi * 16 + ( lane & 15 )];
#pragma unroll
                for ( int mj = 0; mj < NBM; mj++ ) yb[mj] = As[kk][wm * ( TM / 2 ) + mj * 16 + ( lane & 15 )];
#pragma unroll
                for ( int mi = 0; mi < NBN; mi++ )
#pragma unroll
                    for ( int mj = 0; mj < NBM; mj++ ) acc[mi][mj] = pa_mfma<T>::run( xa[mi], yb[mj], acc[mi][mj] );
            }
        }
        __syncthreads();
    }

    if constexpr ( XP ) {
        /* 64 x 64 FP32 tile: the output goes through LDS (Cs[sigma][rho],
         * the operand area, free after the K loop's last barrier), then each
         * wave updates 16 target columns, one column of 64 rho rows per
         * instruction (lane = rho row) */
        static_assert( M32 && TM == 64 && TN == 64, "output staging is for 64 x 64 FP32 tiles" );
#pragma unroll
        for ( int r = 0; r < 16; r++ )
            Cs[( wn * 32 + 8 * ( r >> 2 ) + 4 * ( lane >> 5 ) + ( r & 3 ) ) * TM + wm * 32 + ( lane & 31 )] =
                acc32[0][0][r];
        __syncthreads();
        if ( a_ld[lane] == 0 ) return;
        const int64_t cbase = c_off[lane];
        const int32_t cld = c_ld[lane], ablk = a_blk[lane];
#pragma unroll
        for ( int j = 0; j < 16; j++ ) {
            const int sl = wave * 16 + j;
            if ( b_ld[sl] == 0 || ablk < b_blk[sl] + shift ) continue;
            T *dst = C + cbase + (int64_t)c_col[sl] * cld;
            const T v = Cs[sl * TM + lane];
            if ( atomic ) pa_atomic_add( dst, -v );
            else          *dst -= v;
        }
        return;
    }
    if constexpr ( M32 ) {
#pragma unroll
        for ( int mj = 0; mj < NBM32; mj++ ) {
            const int rl = wm * ( TM / 2 ) + mj * 32 + ( lane & 31 );
            if ( a_ld[rl] == 0 ) continue;
            const int64_t cbase = c_off[rl];
            const int32_t cld = c_ld[rl], ablk = a_blk[rl];
#pragma unroll
            for ( int mi = 0; mi < NBN32; mi++ ) {
#pragma unroll
                for ( int r = 0; r < 16; r++ ) {
                    /* output element r of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4 */
                    const int sl = wn * ( TN / 2 ) + mi * 32 + 8 * ( r >> 2 ) + 4 * ( lane >> 5 ) + ( r & 3 );
                    if ( b_ld[sl] == 0 || ablk < b_blk[sl] + shift ) continue;
                    T *dst = C + cbase + (int64_t)c_col[sl] * cld;
                    const T v = acc32[mi][mj][r];
                    if ( atomic ) pa_atomic_add( dst, -v );
                    else          *dst -= v;
                }
            }
        }
        return;
    }
#pragma unroll
    for ( int mj = 0; mj < NBM; mj++ ) {
        int rl = wm * ( TM / 2 ) + mj * 16 + ( lane & 15 );
        if ( a_ld[rl] == 0 ) continue;
        int64_t cbase = c_off[rl];
        int32_t cld   = c_ld[rl];
        int32_t ablk  = a_blk[rl];
#pragma unroll
        for ( int mi = 0; mi < NBN; mi++ ) {
#pragma unroll
            for ( int r = 0; r < 4; r++ ) {
                int sl = wn * ( TN / 2 ) + mi * 16 + pa_mfma<T>::row( lane, r );
                if ( b_ld[sl] == 0 || ablk < b_blk[sl] + shift ) continue;
                T *dst = C + cbase + (int64_t)c_col[sl] * cld;
                T  v   = acc[mi][mj][r];
                if ( atomic ) pa_atomic_add( dst, -v );
                else          *dst -= v;
            }
        }
    }
}

/* Level-batched scatter-GEMM: one workgroup per pa_gemm_task (TM x TN tile).
 * An update row or column that the target cblk's block structure does not
 * hold (a symbolic factorization that misses fill) is not written and is
 * counted in *nerr, which the host turns into an error. */
template <typename T, int TM, int TN>
__global__ __launch_bounds__( 256 ) void
k_gemm_level( pa_dstruct S, const pa_gemm_task *tasks, T *Lc, T *Uc, int32_t *nerr )
{
    constexpr int PAD = pa_gpad<T>::value;
    /* FP32 64 x 64 tiles stage their output through LDS for the epilogue */
    constexpr bool XP = sizeof( T ) == 4 && TM == 64 && TN == 64;
    /* one LDS area, used in turn by the target-block staging (prologue),
     * the operand chunks (K loop) and the output tile (epilogue) */
    constexpr size_t AB_BYTES = sizeof( T ) * PA_GKC * ( TM + PAD + TN + PAD );
    constexpr size_t FB_BYTES = sizeof( int32_t ) * 3 * PA_GF_STAGE;
    constexpr size_t CS_BYTES = XP ? sizeof( T ) * TM * TN : 0;
    constexpr size_t SH_BYTES = AB_BYTES > FB_BYTES ? ( AB_BYTES > CS_BYTES ? AB_BYTES : CS_BYTES )
                                                    : ( FB_BYTES > CS_BYTES ? FB_BYTES : CS_BYTES );
    __shared__ __attribute__( ( aligned( 16 ) ) ) unsigned char smem[SH_BYTES];
    T ( *As )[TM + PAD] = reinterpret_cast<T ( * )[TM + PAD]>( smem );
    T ( *Bs )[TN + PAD] = reinterpret_cast<T ( * )[TN + PAD]>( smem + sizeof( T ) * PA_GKC * ( TM + PAD ) );
    __shared__ int64_t a_off[TM], b_off[TN], c_off[TM];
    __shared__ int32_t a_ld[TM], b_ld[TN], c_ld[TM], c_col[TN], a_blk[TM], b_blk[TN], grow_s[TM];
    /* first rows / last rows / row offsets of the target cblk's blocks */
    int32_t *f_frow = reinterpret_cast<int32_t *>( smem );
    int32_t *f_lrow = f_frow + PA_GF_STAGE, *f_roff = f_lrow + PA_GF_STAGE;

    const pa_gemm_task tk = tasks[blockIdx.x];
    const int32_t k = tk.k, F = tk.F, nc = S.c_ncol[k];
    const int32_t R = S.c_stride[k] - nc;
    const bool uside = ( tk.flags & PA_TASK_USIDE ) != 0;
    const int t = threadIdx.x;
    const int64_t cbase = S.c_coef[k], rb = S.r_base[k];
    const int32_t fb0 = S.c_fblok[F], nfb = S.c_fblok[F + 1] - fb0;
    const bool staged = nfb <= PA_GF_STAGE;

    /* row descriptors from S.r_info (rho rows: u < TM, sigma rows: u >= TM) */
    for ( int u = t; u < TM + TN; u += 256 ) {
        const bool rho = u < TM;
        const int32_t tt = rho ? u : u - TM;
        const int32_t v  = ( rho ? tk.r0 : tk.s0 ) + tt;
        const bool ok = ( v < R ) && ( rho || tt < tk.slen );
        int4 ri = make_int4( 0, 0, 0, 0 );
        int32_t blk = 0;
        if ( ok ) {
            ri  = S.r_info[rb + v];
            blk = S.r_blk[rb + v];
        }
        const int64_t off = ok ? pa_rinfo_off( ri ) - cbase : 0;
        if ( rho ) {
            a_off[tt] = off; a_ld[tt] = ri.y; a_blk[tt] = blk; grow_s[tt] = ok ? ri.x : -1;
        }
        else {
            const int32_t col = ri.x - S.c_fcol[F];
            const bool    in  = col >= 0 && col < S.c_ncol[F];
            b_off[tt] = off; b_blk[tt] = blk; c_col[tt] = col;
            /* a column outside the target cblk is never written */
            b_ld[tt] = ( ok && in ) ? ri.y : 0;
            if ( ok && !in ) atomicAdd( nerr, 1 );
        }
    }
    if ( staged ) {
        for ( int i = t; i < nfb; i += 256 ) {
            f_frow[i] = S.b_frow[fb0 + i];
            f_lrow[i] = S.b_lrow[fb0 + i];
            f_roff[i] = S.b_rowoff[fb0 + i];
        }
    }
    __syncthreads();
    /* target block of each rho row in F */
    const int32_t ncF = S.c_ncol[F];
    const int64_t coefF = S.c_coef[F];
    for ( int tt = t; tt < TM; tt += 256 ) {
        const int32_t grow = grow_s[tt];
        if ( grow < 0 ) {
            a_ld[tt] = 0; c_off[tt] = 0; c_ld[tt] = 0;
            continue;
        }
        int32_t fr, lr, ro;
        if ( staged ) {
            int32_t lo = 0, hi = nfb - 1;
            while ( lo < hi ) {
                int32_t mid = ( lo + hi + 1 ) >> 1;
                if ( f_frow[mid] <= grow ) lo = mid;
                else                       hi = mid - 1;
            }
            fr = f_frow[lo]; lr = f_lrow[lo]; ro = f_roff[lo];
        }
        else {
            int32_t f = pa_find_globrow( S.b_frow, fb0, fb0 + nfb - 1, grow );
            fr = S.b_frow[f]; lr = S.b_lrow[f]; ro = S.b_rowoff[f];
        }
        c_off[tt] = coefF + (int64_t)ro * ncF + ( grow - fr );
        c_ld[tt]  = lr - fr + 1;
        /* a row outside the target structure is never written */
        if ( grow < fr || grow > lr ) {
            a_ld[tt] = 0;
            atomicAdd( nerr, 1 );
        }
    }
    __syncthreads();

    const T *Asrc = ( uside ? Uc : Lc ) + cbase;
    const T *Bsrc = ( ( Uc != nullptr && !uside ) ? Uc : Lc ) + cbase;
    T *C = uside ? Uc : Lc;
    pa_gemm_tile<T, TM, TN, true, XP>( Asrc, Bsrc, C, nc, a_off, a_ld, b_off, b_ld, c_off, c_ld, c_col, a_blk, b_blk,
                                     uside ? 1 : 0, ( tk.flags & PA_TASK_ATOMIC ) != 0, As, Bs,
                                     reinterpret_cast<T *>( smem ) );
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
    __shared__ T As[PA_GKC][PA_GT + pa_gpad<T>::value];
    __shared__ T Bs[PA_GKC][PA_GT + pa_gpad<T>::value];
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
    pa_gemm_tile<T, PA_GT, PA_GT, false>( A, B, C, K, a_off, a_ld, b_off, b_ld, c_off, c_ld, c_col, a_blk, b_blk, 0, false,
                                        As, Bs );
}

/* ------------------------------------------------------------------------- */
/* Supernodal triangular solve (reference kernels/solve_ztrsmsp.c:302-661).  */
/* x is an n x nrhs column-major array in the factorization numbering.       */
/* ------------------------------------------------------------------------- */

/* x_k <- D_kk^{-1} x_k (lower, forward) or D_kk^{-T} x_k (backward), one
 * workgroup per cblk of the level, x_k read from x, staged in LDS and
 * written to xo.  Blocked by 16
 * columns: the 16 x 16 triangle is loaded into registers (one row, or one
 * column for the transposed sweep, per lane) and solved by wave 0 with
 * shuffles; the rest of the block column is applied by all threads. */
template <typename T, bool BACKWARD, bool UNIT>
__global__ __launch_bounds__( 256 ) void
k_solve_diag( pa_dstruct S, const int32_t *cblks, const T *Coef, const T *x, int32_t nrhs, int64_t ldx, T *xo,
              int64_t ldo )
{
    extern __shared__ unsigned char pa_smem_raw[];
    T *xs = reinterpret_cast<T *>( pa_smem_raw ); /* [nrhs][nc] */
    __shared__ T part[16][17];
    const int32_t k = cblks[blockIdx.x], nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const T *D = Coef + S.c_coef[k];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for ( int32_t idx = t; idx < nc * nrhs; idx += 256 ) {
        int32_t i = idx % nc, r = idx / nc;
        xs[r * nc + i] = x[fcol + i + r * ldx];
    }
    __syncthreads();
    const int32_t nblk = ( nc + 15 ) / 16;
    for ( int32_t bi = 0; bi < nblk; bi++ ) {
        const int32_t jb = BACKWARD ? ( nblk - 1 - bi ) * 16 : bi * 16;
        const int32_t w  = nc - jb < 16 ? nc - jb : 16;
        const int32_t rest0 = jb + w, nrest = nc - rest0;
        if ( BACKWARD && nrest > 0 ) {
            /* x[jb..jb+w) -= D[rest, jb..jb+w)^T x[rest] */
            for ( int32_t r = 0; r < nrhs; r++ ) {
                const int jj = t & 15, g = t >> 4; /* 16 columns x 16 row groups */
                T acc = T( 0 );
                if ( jj < w ) {
                    const T *dc = D + (int64_t)( jb + jj ) * nc;
#pragma unroll 4
                    for ( int32_t i = rest0 + g; i < nc; i += 16 ) acc += dc[i] * xs[r * nc + i];
                }
                part[g][jj] = acc;
                __syncthreads();
                if ( t < w ) {
                    T s2 = T( 0 );
#pragma unroll
                    for ( int gg = 0; gg < 16; gg++ ) s2 += part[gg][t];
                    xs[r * nc + jb + t] -= s2;
                }
                __syncthreads();
            }
        }
        if ( wave == 0 ) {
            const int i = lane & 15;
            /* forward: lane i holds row i of the triangle; backward: column i */
            T tri[16];
#pragma unroll
            for ( int j = 0; j < 16; j++ ) {
                bool in = ( i < w ) && ( j < w );
                tri[j] = in ? ( BACKWARD ? D[( jb + j ) + (int64_t)( jb + i ) * nc] : D[( jb + i ) + (int64_t)( jb + j ) * nc] )
                            : T( 1 );
            }
            for ( int32_t r0 = 0; r0 < nrhs; r0 += 4 ) {
                const int32_t r = r0 + ( lane >> 4 );
                const bool act = ( r < nrhs ) && ( i < w );
                T xi = act ? xs[r * nc + jb + i] : T( 0 );
                if ( !BACKWARD ) {
#pragma unroll
                    for ( int j = 0; j < 16; j++ ) {
                        T xj = __shfl( xi, ( lane & 48 ) + j );
                        if ( !UNIT ) xj /= __shfl( tri[j], ( lane & 48 ) + j );
                        if ( i == j ) xi = xj;
                        else if ( i > j ) xi -= tri[j] * xj;
                    }
                }
                else {
#pragma unroll
                    for ( int j = 15; j >= 0; j-- ) {
                        T xj = __shfl( xi, ( lane & 48 ) + j );
                        if ( !UNIT ) xj /= __shfl( tri[j], ( lane & 48 ) + j );
                        if ( i == j ) xi = xj;
                        else if ( i < j ) xi -= tri[j] * xj;
                    }
                }
                if ( act ) xs[r * nc + jb + i] = xi;
            }
        }
        __syncthreads();
        if ( !BACKWARD && nrest > 0 ) {
            /* x[rest] -= D[rest, jb..jb+w) x[jb..jb+w) */
            for ( int32_t idx = t; idx < nrest * nrhs; idx += 256 ) {
                int32_t i = rest0 + idx % nrest, r = idx / nrest;
                T acc = T( 0 );
#pragma unroll
                for ( int32_t j = 0; j < 16; j++ )
                    if ( j < w ) acc += D[i + (int64_t)( jb + j ) * nc] * xs[r * nc + jb + j];
                xs[r * nc + i] -= acc;
            }
            __syncthreads();
        }
    }
    for ( int32_t idx = t; idx < nc * nrhs; idx += 256 ) {
        int32_t i = idx % nc, r = idx / nc;
        xo[fcol + i + r * ldo] = xs[r * nc + i];
    }
}

/* Sum of v[0..15] over the 64 lanes of a wave by a transposing butterfly:
 * on return lane l holds the wave total of column (l >> 2) & 15 (the four
 * lanes of a quad hold the same value); 17 exchanges instead of 16 x 6. */
template <typename T>
__device__ __forceinline__ T
pa_wave_sum16( T v[16], int lane )
{
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
    for ( int k = 0; k < 8; k++ ) {
        T send = b5 ? v[k] : v[k + 8];
        T keep = b5 ? v[k + 8] : v[k];
        v[k]   = keep + __shfl_xor( send, 32 );
    }
#pragma unroll
    for ( int k = 0; k < 4; k++ ) {
        T send = b4 ? v[k] : v[k + 4];
        T keep = b4 ? v[k + 4] : v[k];
        v[k]   = keep + __shfl_xor( send, 16 );
    }
#pragma unroll
    for ( int k = 0; k < 2; k++ ) {
        T send = b3 ? v[k] : v[k + 2];
        T keep = b3 ? v[k + 2] : v[k];
        v[k]   = keep + __shfl_xor( send, 8 );
    }
    T send = b2 ? v[0] : v[1];
    T s    = ( b2 ? v[1] : v[0] ) + __shfl_xor( send, 4 );
    s += __shfl_xor( s, 2 );
    s += __shfl_xor( s, 1 );
    return s;
}

/* The same over each half wavefront (32 lanes): column c of the half at
 * lanes hb + 2c and hb + 2c + 1. */
template <typename T>
__device__ __forceinline__ T
pa_half_sum16( T v[16], int lane )
{
    const bool b4 = lane & 16, b3 = lane & 8, b2 = lane & 4, b1 = lane & 2;
#pragma unroll
    for ( int k = 0; k < 8; k++ ) {
        T send = b4 ? v[k] : v[k + 8];
        T keep = b4 ? v[k + 8] : v[k];
        v[k]   = keep + __shfl_xor( send, 16 );
    }
#pragma unroll
    for ( int k = 0; k < 4; k++ ) {
        T send = b3 ? v[k] : v[k + 4];
        T keep = b3 ? v[k + 4] : v[k];
        v[k]   = keep + __shfl_xor( send, 8 );
    }
#pragma unroll
    for ( int k = 0; k < 2; k++ ) {
        T send = b2 ? v[k] : v[k + 2];
        T keep = b2 ? v[k + 2] : v[k];
        v[k]   = keep + __shfl_xor( send, 4 );
    }
    T send = b1 ? v[0] : v[1];
    T s    = ( b1 ? v[1] : v[0] ) + __shfl_xor( send, 2 );
    s += __shfl_xor( s, 1 );
    return s;
}


/*
 * Diagonal solves for cblks of at most TH columns, one workgroup per (cblk,
 * RHS chunk), x_k read from x, staged in LDS and written to xo.  Thread t owns row t (forward) or column t
 * (backward) of the diagonal block and keeps the 16 entries of a 16-column
 * block in registers.  A ring of PA_SD_RING blocks is kept in flight, so the
 * loads of block b + PA_SD_RING - 1 overlap the work on blocks b..: the
 * sequential sweep over the blocks does not wait on global memory.  Per
 * block: the 16 lanes owning the block's rows (columns) solve the 16 x 16
 * triangle with v_readlane broadcasts, a barrier publishes x_block, then
 * every thread applies it to its own entry: one barrier per block.
 *   forward:  x_k <- D^{-1} x_k, lane t holds D[t, jb..jb+15]
 *   backward: x_k <- D^{-T} x_k, lane t holds D[jb..jb+15, t], read from a
 *             transposed copy of D (the other factor's diagonal block for
 *             LU, a separate array for LL^t) so the loads are coalesced
 * Loads use clamped in-range addresses, so the loops carry no per-element
 * branches.
 */
#define PA_SD_RING 4

/* The ring loads are unconditional with clamped in-range addresses: a load
 * that sits under a lane-dependent branch makes the compiler wait for every
 * outstanding load at the next use of any ring entry, which would serialise
 * the prefetch.  The entries fetched for rows/columns outside the block, past
 * nc or above the diagonal are never combined with a live value (lanes of
 * other rows never write, lanes past nc carry zeros through the triangle,
 * and the rest updates of a block only exist when the block is full). */
template <typename T>
__device__ __forceinline__ void
pa_sd_load_rows( T (&v)[16], const T *D, int32_t nc, int t, int32_t jb )
{
    /* D[t, jb + j] */
    const T *Dt = D + min( t, nc - 1 );
#pragma unroll
    for ( int j = 0; j < 16; j++ ) v[j] = Dt[(int64_t)min( jb + j, nc - 1 ) * nc];
}

template <typename T>
__device__ __forceinline__ void
pa_sd_load_cols( T (&v)[16], const T *DT, int32_t nc, int t, int32_t jb )
{
    /* D[jb + i, t] = DT[t, jb + i], read from the transposed block so that
     * the lanes of a wave read consecutive addresses */
    pa_sd_load_rows( v, DT, nc, t, jb );
}

template <typename T, bool UNIT, int TH>
__global__ __launch_bounds__( TH ) void
k_solve_diag_fwd( pa_dstruct S, const int32_t *cblks, const T *Coef, const T *x, int32_t nrhs, int32_t rchunk,
                  int64_t ldx, T *xo, int64_t ldo )
{
    extern __shared__ unsigned char pa_smem_raw[];
    T *xs = reinterpret_cast<T *>( pa_smem_raw ); /* [nr][nc] */
    const int32_t k = cblks[blockIdx.x], nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const int32_t r0 = blockIdx.y * rchunk, nr = min( rchunk, nrhs - r0 );
    const T *D = Coef + S.c_coef[k];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for ( int32_t idx = t; idx < nc * nr; idx += TH ) {
        int32_t i = idx % nc, r = idx / nc;
        xs[r * nc + i] = x[fcol + i + ( r0 + r ) * ldx];
    }
    T ring[PA_SD_RING][16] = {};
#pragma unroll
    for ( int u = 0; u < PA_SD_RING - 1; u++ ) pa_sd_load_rows( ring[u], D, nc, t, u * 16 );
    __syncthreads();
    const int32_t nblk = ( nc + 15 ) >> 4;
    for ( int32_t b0 = 0; b0 < nblk; b0 += PA_SD_RING ) {
#pragma unroll
        for ( int u = 0; u < PA_SD_RING; u++ ) {
            const int32_t b = b0 + u;
            if ( b >= nblk ) break;
            pa_sd_load_rows( ring[( u + PA_SD_RING - 1 ) % PA_SD_RING], D, nc, t, ( b + PA_SD_RING - 1 ) * 16 );
            const T(&cur)[16] = ring[u];
            const int32_t jb = b * 16, jn = jb + 16;
            if ( wave == ( jb >> 6 ) ) {
                /* rows jb..jb+15 live in lanes base..base+15 of this wave.
                 * cs[j] = L[i][j] / L[j][j] below the diagonal of the real
                 * rows and 0 elsewhere, so lane j's running value can be
                 * broadcast unscaled and the chain is readlane + fma. */
                const int  base = jb & 63, i = lane - base;
                const bool in = ( i >= 0 ) && ( i < 16 ) && ( jb + i < nc );
                T d = T( 1 );
#pragma unroll
                for ( int j = 0; j < 16; j++ )
                    if ( i == j ) d = cur[j];
                const T dinv = ( UNIT || !in ) ? T( 1 ) : T( 1 ) / d;
                T cs[16];
#pragma unroll
                for ( int j = 0; j < 16; j++ ) {
                    const T rj = UNIT ? T( 1 ) : pa_readlane( dinv, base + j );
                    cs[j] = ( in && i > j ) ? cur[j] * rj : T( 0 );
                }
                for ( int32_t r = 0; r < nr; r++ ) {
                    T xi = in ? xs[r * nc + jb + i] : T( 0 );
#pragma unroll
                    for ( int j = 0; j < 16; j++ ) xi -= cs[j] * pa_readlane( xi, base + j );
                    if ( in ) xs[r * nc + jb + i] = xi * dinv;
                }
            }
            __syncthreads();
            if ( t < nc && t >= jn ) {
                for ( int32_t r = 0; r < nr; r++ ) {
                    T s = T( 0 );
#pragma unroll
                    for ( int j = 0; j < 16; j++ ) s += cur[j] * xs[r * nc + jb + j];
                    xs[r * nc + t] -= s;
                }
            }
        }
    }
    __syncthreads();
    for ( int32_t idx = t; idx < nc * nr; idx += TH ) {
        int32_t i = idx % nc, r = idx / nc;
        xo[fcol + i + ( r0 + r ) * ldo] = xs[r * nc + i];
    }
}

template <typename T, bool UNIT, int TH>
__global__ __launch_bounds__( TH ) void
k_solve_diag_bwd( pa_dstruct S, const int32_t *cblks, const T *DTab, const int64_t *dt_off, const T *x, int32_t nrhs,
                  int32_t rchunk, int64_t ldx, T *xo, int64_t ldo )
{
    extern __shared__ unsigned char pa_smem_raw[];
    T *xs = reinterpret_cast<T *>( pa_smem_raw ); /* [nr][nc] */
    const int32_t k = cblks[blockIdx.x], nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const int32_t r0 = blockIdx.y * rchunk, nr = min( rchunk, nrhs - r0 );
    const T *D = DTab + dt_off[k]; /* the transposed diagonal block */
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for ( int32_t idx = t; idx < nc * nr; idx += TH ) {
        int32_t i = idx % nc, r = idx / nc;
        xs[r * nc + i] = x[fcol + i + ( r0 + r ) * ldx];
    }
    const int32_t nblk = ( nc + 15 ) >> 4;
    /* step u of the sweep processes block nblk - 1 - u */
    T ring[PA_SD_RING][16] = {};
#pragma unroll
    for ( int u = 0; u < PA_SD_RING - 1; u++ ) {
        /* unconditional (see above): past the first block, block 0 is reloaded */
        pa_sd_load_cols( ring[u], D, nc, t, max( nblk - 1 - u, 0 ) * 16 );
    }
    __syncthreads();
    for ( int32_t s0 = 0; s0 < nblk; s0 += PA_SD_RING ) {
#pragma unroll
        for ( int u = 0; u < PA_SD_RING; u++ ) {
            const int32_t st = s0 + u;
            if ( st >= nblk ) break;
            const int32_t b  = nblk - 1 - st;
            const int32_t bp = b - ( PA_SD_RING - 1 );
            pa_sd_load_cols( ring[( u + PA_SD_RING - 1 ) % PA_SD_RING], D, nc, t, max( bp, 0 ) * 16 );
            const T(&cur)[16] = ring[u];
            const int32_t jb = b * 16;
            if ( wave == ( jb >> 6 ) ) {
                /* rows jb..jb+15 of U = D^T live in lanes base..base+15;
                 * same scaled-coefficient chain as the forward sweep */
                const int  base = jb & 63, j = lane - base;
                const bool in = ( j >= 0 ) && ( j < 16 ) && ( jb + j < nc );
                T d = T( 1 );
#pragma unroll
                for ( int i = 0; i < 16; i++ )
                    if ( j == i ) d = cur[i];
                const T dinv = ( UNIT || !in ) ? T( 1 ) : T( 1 ) / d;
                T cs[16];
#pragma unroll
                for ( int i = 0; i < 16; i++ ) {
                    const T ri = UNIT ? T( 1 ) : pa_readlane( dinv, base + i );
                    cs[i] = ( in && j < i ) ? cur[i] * ri : T( 0 );
                }
                for ( int32_t r = 0; r < nr; r++ ) {
                    T xj = in ? xs[r * nc + jb + j] : T( 0 );
#pragma unroll
                    for ( int i = 15; i >= 0; i-- ) xj -= cs[i] * pa_readlane( xj, base + i );
                    if ( in ) xs[r * nc + jb + j] = xj * dinv;
                }
            }
            __syncthreads();
            if ( t < nc && t < jb ) {
                /* the first block processed may be partial (rows past nc) */
                for ( int32_t r = 0; r < nr; r++ ) {
                    T sum = T( 0 );
#pragma unroll
                    for ( int i = 0; i < 16; i++ ) {
                        const int32_t row = min( jb + i, nc - 1 );
                        sum += ( jb + i < nc ? cur[i] : T( 0 ) ) * xs[r * nc + row];
                    }
                    xs[r * nc + t] -= sum;
                }
            }
        }
    }
    __syncthreads();
    for ( int32_t idx = t; idx < nc * nr; idx += TH ) {
        int32_t i = idx % nc, r = idx / nc;
        xo[fcol + i + ( r0 + r ) * ldo] = xs[r * nc + i];
    }
}

/* Diagonal-block solves by products with the inverted blocks that the TRSM
 * launches write (M_k = inverse of the lower triangle of the swept factor's
 * diagonal block, column major, c_dtoff offsets, zeros above the diagonal):
 * forward x_k <- M_k x_k, backward x_k <- M_k^T x_k.  A task is a block of
 * 64 output entries of one cblk (k, b), so that a wide cblk spreads over
 * several workgroups; the inputs come from xin and the products go to xout,
 * a different array (the blocks of one cblk run concurrently).  One
 * workgroup of TH threads per (task, chunk of rchunk right-hand sides); the
 * x entries the block needs are staged in LDS.  Each wave takes tasks whose
 * loads are all independent, so that many loads are in flight at once:
 *   forward:  rows 64b.. (lane = row) by chunks of 32 columns left of the
 *             block's end, 32 coalesced loads; the waves' partial row sums
 *             are added in LDS;
 *   backward: columns 64b.. by groups of 16, rows from the group's first
 *             column down in chunks of 128 (lane = row, two rows per lane),
 *             32 coalesced loads; the 16 column sums are reduced by the
 *             transposing butterfly and added in LDS. */
template <typename T, bool BWD, int TH>
__global__ __launch_bounds__( TH ) void
k_solve_dinv( pa_dstruct S, const int2 *tasks, const T *Dinv, const T *xin, int64_t ldi, T *xout, int64_t ldo,
              int32_t nrhs, int32_t rchunk )
{
    extern __shared__ unsigned char pa_smem_raw[];
    constexpr int NW = TH / 64;
    const int2    tk = tasks[blockIdx.x];
    const int32_t k = tk.x, nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const int32_t b0 = 64 * tk.y, nb = min( 64, nc - b0 );
    const int32_t r0 = blockIdx.y * rchunk, nr = min( rchunk, nrhs - r0 );
    const T      *M = Dinv + S.c_dtoff[k];
    const int     t = threadIdx.x, lane = t & 63, wave = t >> 6;
    /* x entries of the block: forward columns [0, b0 + nb), backward rows
     * [b0, nc) */
    const int32_t x0 = BWD ? b0 : 0, xn = BWD ? nc - b0 : b0 + nb;
    T *xs = reinterpret_cast<T *>( pa_smem_raw ); /* [nr][xn]                          */
    T *ps = xs + (int64_t)nr * xn;                /* forward [NW][nr][64], backward [nr][64] */
    for ( int32_t idx = t; idx < xn * nr; idx += TH ) {
        const int32_t i = idx % xn, r = idx / xn;
        xs[idx] = xin[fcol + x0 + i + ( r0 + r ) * ldi];
    }
    for ( int32_t idx = t; idx < ( BWD ? 1 : NW ) * nr * 64; idx += TH ) ps[idx] = T( 0 );
    __syncthreads();
    if ( !BWD ) {
        const int32_t i = b0 + lane, ic = min( i, nc - 1 ), nch = ( xn + 31 ) >> 5;
        T *pw = ps + (int64_t)wave * nr * 64;
        for ( int32_t ch = wave; ch < nch; ch += NW ) {
            const int32_t c0 = 32 * ch, cnt = min( 32, xn - c0 );
            T m[32];
#pragma unroll
            for ( int j = 0; j < 32; j++ ) m[j] = M[ic + (int64_t)( c0 + min( j, cnt - 1 ) ) * nc];
            for ( int32_t r = 0; r < nr; r++ ) {
                const T *xr  = xs + r * xn + c0;
                T        acc = T( 0 );
#pragma unroll
                for ( int j = 0; j < 32; j++ ) acc += ( j < cnt ? m[j] : T( 0 ) ) * xr[min( j, cnt - 1 )];
                pw[r * 64 + lane] += acc; /* this wave's own slot */
            }
        }
        __syncthreads();
        for ( int32_t idx = t; idx < nr * 64; idx += TH ) {
            const int32_t r = idx >> 6, l = idx & 63;
            if ( l >= nb ) continue;
            T s = T( 0 );
#pragma unroll
            for ( int w = 0; w < NW; w++ ) s += ps[( (int64_t)w * nr + r ) * 64 + l];
            xout[fcol + b0 + l + ( r0 + r ) * ldo] = s;
        }
    }
    else {
        const int     col = lane >> 2;
        const int32_t ng  = ( nb + 15 ) >> 4;
        int32_t       task = 0;
        for ( int32_t g = 0; g < ng; g++ ) {
            const int32_t jg = b0 + 16 * g, ncols = min( 16, nc - jg ), nq = ( nc - jg + 127 ) >> 7;
            for ( int32_t q = 0; q < nq; q++, task++ ) {
                if ( task % NW != wave ) continue;
                const int32_t ia = jg + 128 * q + lane, ib = ia + 64;
                const int32_t ica = min( ia, nc - 1 ), icb = min( ib, nc - 1 );
                T ma[16], mb[16];
#pragma unroll
                for ( int cc = 0; cc < 16; cc++ ) {
                    const int64_t cofs = (int64_t)min( jg + cc, nc - 1 ) * nc;
                    ma[cc] = M[ica + cofs];
                    mb[cc] = M[icb + cofs];
                }
                for ( int32_t r = 0; r < nr; r++ ) {
                    const T ya = ( ia < nc ) ? xs[r * xn + ica - x0] : T( 0 );
                    const T yb = ( ib < nc ) ? xs[r * xn + icb - x0] : T( 0 );
                    T       v[16];
#pragma unroll
                    for ( int cc = 0; cc < 16; cc++ ) v[cc] = ( cc < ncols ) ? ma[cc] * ya + mb[cc] * yb : T( 0 );
                    const T sum = pa_wave_sum16<T>( v, lane ); /* lane 4c: column jg + c */
                    if ( ( lane & 3 ) == 0 && col < ncols ) atomicAdd( ps + r * 64 + 16 * g + col, sum );
                }
            }
        }
        __syncthreads();
        for ( int32_t idx = t; idx < nr * 64; idx += TH ) {
            const int32_t r = idx >> 6, l = idx & 63;
            if ( l < nb ) xout[fcol + b0 + l + ( r0 + r ) * ldo] = ps[idx];
        }
    }
}

/* The same products in place (x_k read and overwritten), one workgroup of
 * TH threads per (cblk, chunk of rchunk right-hand sides): the sweeps of
 * several right-hand sides work in place on one array, and the whole cblk's
 * inputs are staged in LDS (xs) before any product is written back; the
 * products accumulate in LDS (ys).
 * The work is cut into wave tasks whose loads are all independent, so that
 * each wave has many loads in flight at once instead of a chain:
 *   forward:  (64 rows, 32 columns) tiles of the lower triangle; lane = row,
 *             32 coalesced column loads, partial row sums added to ys;
 *   backward: (16 columns, 64 rows) tiles from the group's first column
 *             down; lane = row, 16 coalesced loads, the 16 column sums
 *             reduced by the transposing butterfly and added to ys. */
template <typename T, bool BWD, int TH>
__global__ __launch_bounds__( TH ) void
k_solve_dinv_cblk( pa_dstruct S, const int32_t *cblks, const T *Dinv, T *x, int32_t nrhs, int32_t rchunk, int64_t ldx )
{
    extern __shared__ unsigned char pa_smem_raw[];
    const int32_t k = cblks[blockIdx.x], nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const int32_t r0 = blockIdx.y * rchunk, nr = min( rchunk, nrhs - r0 );
    T *xs = reinterpret_cast<T *>( pa_smem_raw ); /* [nr][nc] inputs   */
    T *ys = xs + (int64_t)rchunk * nc;            /* [nr][nc] products */
    const T      *M = Dinv + S.c_dtoff[k];
    const int     t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int NW = TH / 64;
    for ( int32_t idx = t; idx < nc * nr; idx += TH ) {
        const int32_t i = idx % nc, r = idx / nc;
        xs[r * nc + i] = x[fcol + i + ( r0 + r ) * ldx];
        ys[r * nc + i] = T( 0 );
    }
    __syncthreads();
    if ( !BWD ) {
        const int32_t nrb = ( nc + 63 ) >> 6, ncb = ( nc + 31 ) >> 5;
        for ( int32_t task = wave; task < nrb * ncb; task += NW ) {
            const int32_t rb = task / ncb, cb = task % ncb;
            const int32_t c0 = 32 * cb;
            if ( c0 > 64 * rb + 63 ) continue; /* above the diagonal: zero */
            const int32_t i = 64 * rb + lane, ic = min( i, nc - 1 ), cnt = min( 32, nc - c0 );
            T m[32];
#pragma unroll
            for ( int j = 0; j < 32; j++ ) m[j] = M[ic + (int64_t)( c0 + min( j, cnt - 1 ) ) * nc];
            for ( int32_t r = 0; r < nr; r++ ) {
                const T *xr  = xs + r * nc + c0;
                T        acc = T( 0 );
#pragma unroll
                for ( int j = 0; j < 32; j++ ) acc += ( j < cnt ? m[j] : T( 0 ) ) * xr[min( j, cnt - 1 )];
                if ( i < nc ) atomicAdd( ys + r * nc + i, acc );
            }
        }
    }
    else {
        const int32_t ng = ( nc + 15 ) >> 4;
        const int     col = lane >> 2;
        /* tasks: column group g, row chunk q from the group's first column */
        int32_t task = 0;
        for ( int32_t g = 0; g < ng; g++ ) {
            const int32_t c0 = 16 * g, ncols = min( 16, nc - c0 ), nq = ( nc - c0 + 63 ) >> 6;
            for ( int32_t q = 0; q < nq; q++, task++ ) {
                if ( task % NW != wave ) continue;
                const int32_t i = c0 + 64 * q + lane, ic = min( i, nc - 1 );
                T m[16];
#pragma unroll
                for ( int cc = 0; cc < 16; cc++ ) m[cc] = M[ic + (int64_t)min( c0 + cc, nc - 1 ) * nc];
                for ( int32_t r = 0; r < nr; r++ ) {
                    const T yi = ( i < nc ) ? xs[r * nc + ic] : T( 0 );
                    T       v[16];
#pragma unroll
                    for ( int cc = 0; cc < 16; cc++ ) v[cc] = ( cc < ncols ? m[cc] : T( 0 ) ) * yi;
                    const T sum = pa_wave_sum16<T>( v, lane ); /* lane 4c: column c0 + c */
                    if ( ( lane & 3 ) == 0 && col < ncols ) atomicAdd( ys + r * nc + c0 + col, sum );
                }
            }
        }
    }
    __syncthreads();
    for ( int32_t idx = t; idx < nc * nr; idx += TH ) {
        const int32_t i = idx % nc, r = idx / nc;
        x[fcol + i + ( r0 + r ) * ldx] = ys[r * nc + i];
    }
}

/* DT_k = D_k^T for the lower triangle of every diagonal block (LL^t), through
 * 32 x 32 LDS tiles so that reads and writes are coalesced; one workgroup
 * per cblk.  The upper triangle of DT_k is left as it is (zero). */
template <typename T>
__global__ __launch_bounds__( 256 ) void
k_diag_transpose( pa_dstruct S, const T *Coef, T *DTab )
{
    __shared__ T tile[32][33];
    const int32_t k = blockIdx.x, nc = S.c_ncol[k];
    const T *D  = Coef + S.c_coef[k];
    T       *DT = DTab + S.c_dtoff[k];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; /* 32 x 8 */
    const int32_t nt = ( nc + 31 ) >> 5;
    for ( int32_t tj = 0; tj < nt; tj++ )
        for ( int32_t ti = tj; ti < nt; ti++ ) {
            /* tile (ti, tj) of D: rows 32 ti.., columns 32 tj.. */
            for ( int yy = ty; yy < 32; yy += 8 ) {
                int32_t i = ti * 32 + tx, j = tj * 32 + yy;
                tile[yy][tx] = ( i < nc && j < nc ) ? D[i + (int64_t)j * nc] : T( 0 );
            }
            __syncthreads();
            for ( int yy = ty; yy < 32; yy += 8 ) {
                int32_t i = ti * 32 + yy, j = tj * 32 + tx; /* DT[j, i] = D[i, j] */
                if ( i < nc && j < nc && i >= j ) DT[j + (int64_t)i * nc] = tile[tx][yy];
            }
            __syncthreads();
        }
}

/*
 * Off-diagonal part of the solve, with the row descriptors of S.r_info (no
 * search).  A task is (cblk k, first off-diagonal row r0, first column c0,
 * descriptor index of row r0).
 * Forward:  x[rows] -= A[rows, cols] x_k[cols]: 256 rows x 64 columns, one
 *           row per lane, one atomic per row; x_k is read from xk (the
 *           array the diagonal step wrote its results to).
 * Backward: x_k[cols] -= A[rows, cols]^T x[rows]: each wave takes 16 columns
 *           and 256 rows (4 groups of 64) and reduces over its lanes with a
 *           transposing butterfly, one atomic per column; the solved rows
 *           x[rows] are read from xa (the array the diagonal steps wrote
 *           their results to).  A task has PA_SB_NCG(nc) column groups and
 *           4 / PA_SB_NCG(nc) row parts.
 */
#define PA_SF_ROWS 256
#define PA_SF_COLS 64
#define PA_SB_NCG( nc ) ( ( nc ) <= 16 ? 1 : ( ( nc ) <= 32 ? 2 : 4 ) )

template <typename T, int RC>
__global__ __launch_bounds__( 256 ) void
k_solve_fwd_update( pa_dstruct S, const int4 *tasks, const T *Coef, T *x, int32_t nrhs, int64_t ldx, const T *xk,
                    int64_t ldk )
{
    __shared__ T xks[RC][PA_SF_COLS];
    const int4    tk = tasks[blockIdx.x];
    const int32_t k = tk.x, nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const int32_t R = S.c_stride[k] - nc, c0 = tk.z, ncols = min( PA_SF_COLS, nc - c0 );
    const int t = threadIdx.x;
    const int32_t v = tk.y + t;
    const bool    valid = v < R;
    int4 ri = make_int4( 0, 0, 0, 0 ); /* an empty row reads Coef[0] */
    if ( valid ) ri = S.r_info[S.r_base[k] + v];
    const T *A = Coef + pa_rinfo_off( ri ) + (int64_t)c0 * ri.y;
    /* One FP32 right-hand side: the row's PA_SF_COLS coefficients, all loads
     * issued at once from in-range addresses (columns past the cblk repeat
     * its last one and are masked below), one memory latency per task
     * instead of a chain.  FP64 or several right-hand sides: that many
     * registers would leave one or two waves per SIMD (512 VGPRs for four
     * FP64 right-hand sides), so the coefficients are read NG columns at a
     * time, again for each chunk of RC right-hand sides; the staged x_k
     * entries past the cblk are zero, so those products vanish. */
    constexpr int NG = ( RC == 1 && sizeof( T ) == 4 ) ? PA_SF_COLS : ( RC == 1 ? 32 : 16 );
    if ( NG != PA_SF_COLS ) {
        for ( int32_t r0 = 0; r0 < nrhs; r0 += RC ) {
            const int32_t nr = min( RC, nrhs - r0 );
            for ( int32_t idx = t; idx < PA_SF_COLS * nr; idx += 256 ) {
                const int32_t c = idx % PA_SF_COLS, u = idx / PA_SF_COLS;
                xks[u][c] = ( c < ncols ) ? xk[fcol + c0 + c + (int64_t)( r0 + u ) * ldk] : T( 0 );
            }
            __syncthreads();
            if ( valid ) {
                T acc[RC];
#pragma unroll
                for ( int u = 0; u < RC; u++ ) acc[u] = T( 0 );
                for ( int32_t cg = 0; cg < ncols; cg += NG ) {
                    T a[NG];
#pragma unroll
                    for ( int w = 0; w < NG; w++ ) a[w] = A[(int64_t)min( cg + w, ncols - 1 ) * ri.y];
#pragma unroll
                    for ( int w = 0; w < NG; w++ )
#pragma unroll
                        for ( int u = 0; u < RC; u++ ) acc[u] += a[w] * xks[u][cg + w];
                }
#pragma unroll
                for ( int u = 0; u < RC; u++ )
                    if ( u < nr ) pa_atomic_add( x + ri.x + (int64_t)( r0 + u ) * ldx, -acc[u] );
            }
            __syncthreads();
        }
        return;
    }
    T a[PA_SF_COLS];
#pragma unroll
    for ( int w = 0; w < PA_SF_COLS; w++ ) a[w] = A[(int64_t)min( w, ncols - 1 ) * ri.y];
    /* right-hand sides RC at a time */
    for ( int32_t r0 = 0; r0 < nrhs; r0 += RC ) {
        const int32_t nr = min( RC, nrhs - r0 );
        for ( int32_t idx = t; idx < PA_SF_COLS * nr; idx += 256 ) {
            const int32_t c = idx % PA_SF_COLS, u = idx / PA_SF_COLS;
            xks[u][c] = ( c < ncols ) ? xk[fcol + c0 + c + (int64_t)( r0 + u ) * ldk] : T( 0 );
        }
        __syncthreads();
        if ( valid ) {
            T acc[RC];
#pragma unroll
            for ( int u = 0; u < RC; u++ ) acc[u] = T( 0 );
#pragma unroll
            for ( int w = 0; w < PA_SF_COLS; w++ )
#pragma unroll
                for ( int u = 0; u < RC; u++ ) acc[u] += a[w] * xks[u][w];
#pragma unroll
            for ( int u = 0; u < RC; u++ )
                if ( u < nr ) pa_atomic_add( x + ri.x + (int64_t)( r0 + u ) * ldx, -acc[u] );
        }
        __syncthreads();
    }
}

template <typename T, bool MR>
__global__ __launch_bounds__( 256 ) void
k_solve_bwd_update( pa_dstruct S, const int4 *tasks, const T *Coef, T *x, int32_t nrhs, int64_t ldx, const T *xa,
                    int64_t lda )
{
    const int4    tk = tasks[blockIdx.x];
    const int32_t k = tk.x, nc = S.c_ncol[k], fcol = S.c_fcol[k];
    const int32_t R = S.c_stride[k] - nc;
    const int t = threadIdx.x, lane = t & 63, g = t >> 6;
    const int     ncg = PA_SB_NCG( nc );
    const int32_t cb = tk.z + ( g % ncg ) * 16, ncols = max( 0, min( 16, nc - cb ) );
    const int32_t rw = tk.y + ( g / ncg ) * 256; /* first row of this wave */
    const int4   *info = S.r_info + S.r_base[k];
    int4 ri[4];
    bool ok[4];
#pragma unroll
    for ( int q = 0; q < 4; q++ ) {
        const int32_t row = rw + q * 64 + lane;
        ok[q] = ( row < R ) && ( ncols > 0 );
        ri[q] = ok[q] ? info[row] : make_int4( 0, 1, 0, 0 );
    }
    const int col = lane >> 2;
    if ( !MR ) { /* one right-hand side */
        /* every load of the wave is unconditional, from an in-range address
         * (an empty row reads x[0] and the panel's first entries; a column
         * past the cblk repeats its last one), and the contributions are
         * masked: the loads of the four row groups are all in flight at once */
        T xv[4];
#pragma unroll
        for ( int q = 0; q < 4; q++ ) {
            const T v = xa[ok[q] ? ri[q].x : 0];
            xv[q] = ok[q] ? v : T( 0 );
        }
        T v[16];
#pragma unroll
        for ( int cc = 0; cc < 16; cc++ ) v[cc] = T( 0 );
#pragma unroll
        for ( int q = 0; q < 4; q++ ) {
            const T *A = Coef + ( ok[q] ? pa_rinfo_off( ri[q] ) + (int64_t)cb * ri[q].y : 0 );
#pragma unroll
            for ( int cc = 0; cc < 16; cc++ ) {
                const T a = A[(int64_t)min( cc, max( ncols - 1, 0 ) ) * ri[q].y];
                v[cc] += ( cc < ncols ? a : T( 0 ) ) * xv[q];
            }
        }
        T s = pa_wave_sum16<T>( v, lane );
        if ( ( lane & 3 ) == 0 && col < ncols ) pa_atomic_add( x + fcol + cb + col, -s );
        return;
    }
    /* several right-hand sides: the tile is read into registers once, NQ
     * row groups at a time (all four in FP32; two in FP64, whose 64 values
     * per lane would leave two waves per SIMD), and reused for every RHS */
    constexpr int NQ = sizeof( T ) == 8 ? 2 : 4;
#pragma unroll
    for ( int q0 = 0; q0 < 4; q0 += NQ ) {
        T a[NQ][16];
#pragma unroll
        for ( int q = 0; q < NQ; q++ ) {
            const T *A = Coef + pa_rinfo_off( ri[q0 + q] ) + (int64_t)cb * ri[q0 + q].y;
#pragma unroll
            for ( int cc = 0; cc < 16; cc++ )
                a[q][cc] = ( ok[q0 + q] && cc < ncols ) ? A[(int64_t)cc * ri[q0 + q].y] : T( 0 );
        }
        for ( int32_t r = 0; r < nrhs; r++ ) {
            T xv[NQ];
#pragma unroll
            for ( int q = 0; q < NQ; q++ ) xv[q] = ok[q0 + q] ? xa[ri[q0 + q].x + r * lda] : T( 0 );
            T v[16];
#pragma unroll
            for ( int cc = 0; cc < 16; cc++ ) {
                v[cc] = a[0][cc] * xv[0];
#pragma unroll
                for ( int q = 1; q < NQ; q++ ) v[cc] += a[q][cc] * xv[q];
            }
            T s = pa_wave_sum16<T>( v, lane );
            if ( ( lane & 3 ) == 0 && col < ncols ) pa_atomic_add( x + fcol + cb + col + r * ldx, -s );
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Supernodal solve of the narrow LL^t cblks (those of k_small_llt), one     */
/* wavefront per cblk for all right-hand sides: the panel is read once into  */
/* registers (lane l: rows l, l+64, l+128; lanes < nc also the columns of    */
/* the diagonal block) and reused for every RHS.  Forward (reference          */
/* kernels/solve_ztrsmsp.c:302-466): x_k = L_kk^{-1} x_k, written to xo,    */
/* then the rows below are updated in x with atomic adds.  Backward          */
/* (:468-661): x_k -= L_bk^T x_b gathered from the final rows below (in xo,  */
/* transposing wave reduction), then x_k = L_kk^{-T} x_k, written to xo.     */
/* ------------------------------------------------------------------------- */
#define PA_SM_RC 4

template <typename T, bool BWD, int NS, bool INPLACE, bool PAIR = false>
__global__ __launch_bounds__( 256 ) void
k_small_solve( pa_dstruct S, const int32_t *cblks, int32_t ncb, const T *Lc, T *x, int32_t nrhs, int64_t ldx, T *xo_arg,
               int64_t ldo_arg )
{
    /* PAIR: the two half wavefronts take one cblk of at most 32 rows each
     * (hl: lane in the half, hb: first lane of the half); broadcasts stay
     * inside a half.  A half without a cblk (odd count) has nc = stride = 0,
     * so its loads, stores and atomics are all masked off. */
    static_assert( !PAIR || NS == 1, "two cblks per wavefront: one row slot of 32" );
    /* in place (xo == x): one base for every access, which keeps the
     * backward variant at three waves per SIMD in FP64 */
    T *const      xo  = INPLACE ? x : xo_arg;
    const int64_t ldo = INPLACE ? ldx : ldo_arg;
    const int     lane = threadIdx.x & 63;
    const int     hl = PAIR ? ( lane & 31 ) : lane, hb = PAIR ? ( lane & 32 ) : 0;
    const int32_t w0   = ( (int32_t)blockIdx.x * 4 + ( threadIdx.x >> 6 ) ) * ( PAIR ? 2 : 1 );
    if ( w0 >= ncb ) return; /* whole waves only; no workgroup barrier below */
    const int32_t idx = w0 + ( PAIR ? ( lane >> 5 ) : 0 );
    const bool    act = idx < ncb;
    const int32_t k = cblks[act ? idx : w0];
    const int32_t nc = act ? S.c_ncol[k] : 0, stride = act ? S.c_stride[k] : 0, fcol = S.c_fcol[k];
    const int64_t coef = S.c_coef[k], rb0 = S.r_base[k];
    /* value of lane j of this lane's cblk */
    auto bcast = [&]( T v, int j ) -> T { return PAIR ? __shfl( v, hb + j, 64 ) : pa_readlane( v, j ); };
    /* column loops: the wave-uniform bound nc, or with two cblks the larger
     * of their widths (columns past a cblk's nc are no-ops for its half:
     * zero reciprocals and panel entries) */
    const int32_t ncl = PAIR ? __builtin_amdgcn_readfirstlane( max( nc, __shfl_xor( nc, 32, 64 ) ) ) : nc;

    /* off-diagonal rows: slots q with row lane + 64 q in [nc, stride) */
    T       P[NS][PA_SM_NC];
    int32_t grow[NS];
    bool    offd[NS];
#pragma unroll
    for ( int q = 0; q < NS; q++ ) {
        const int32_t r = hl + 64 * q;
        offd[q]          = r >= nc && r < stride;
        int64_t base     = 0;
        int32_t ld       = 1;
        grow[q]          = 0;
        if ( offd[q] ) {
            const int4 ri = S.r_info[rb0 + ( r - nc )];
            base    = pa_rinfo_off( ri );
            ld      = ri.y;
            grow[q] = ri.x;
        }
#pragma unroll
        for ( int c = 0; c < PA_SM_NC; c++ ) P[q][c] = ( offd[q] && c < nc ) ? Lc[base + (int64_t)c * ld] : T( 0 );
    }
    /* diagonal block: row `lane` (forward) or column `lane` (backward), in
     * the first slot of the lanes below nc, which hold no off-diagonal row
     * (one register array for both: the chains below only read it on lanes
     * below nc, the update parts only use the off-diagonal lanes) */
    T (&D)[PA_SM_NC] = P[0];
#pragma unroll
    for ( int c = 0; c < PA_SM_NC; c++ )
        if ( hl < nc )
            D[c] = ( c < nc ) ? ( BWD ? Lc[coef + (int64_t)hl * nc + c] : Lc[coef + (int64_t)c * nc + hl] ) : T( 0 );

    /* reciprocals of the diagonal, uniform across the wave */
    T dinv[PA_SM_NC];
#pragma unroll
    for ( int j = 0; j < PA_SM_NC; j++ ) {
        const T dj = bcast( D[j], j );
        dinv[j]    = ( j < nc ) ? T( 1 ) / dj : T( 0 );
    }

    /* right-hand sides PA_SM_RC at a time: independent chains interleave */
    for ( int32_t r0 = 0; r0 < nrhs; r0 += PA_SM_RC ) {
        T    xv[PA_SM_RC];
        bool rok[PA_SM_RC];
#pragma unroll
        for ( int u = 0; u < PA_SM_RC; u++ ) {
            rok[u] = r0 + u < nrhs;
            xv[u]  = ( rok[u] && hl < nc ) ? x[(int64_t)( r0 + u ) * ldx + fcol + hl] : T( 0 );
        }
        if ( !BWD ) {
#pragma unroll
            for ( int j = 0; j < PA_SM_NC; j++ ) {
                if ( j < ncl ) {
#pragma unroll
                    for ( int u = 0; u < PA_SM_RC; u++ ) {
                        const T xj = bcast( xv[u], j ) * dinv[j];
                        if ( hl == j ) xv[u] = xj;
                        else if ( hl > j && hl < nc ) xv[u] -= D[j] * xj; /* L[hl][j] x_j */
                    }
                }
            }
#pragma unroll
            for ( int u = 0; u < PA_SM_RC; u++ ) {
                if ( !rok[u] ) continue;
                T *xr = x + (int64_t)( r0 + u ) * ldx;
                if ( hl < nc ) xo[(int64_t)( r0 + u ) * ldo + fcol + hl] = xv[u];
                T acc[NS];
#pragma unroll
                for ( int q = 0; q < NS; q++ ) acc[q] = T( 0 );
#pragma unroll
                for ( int c = 0; c < PA_SM_NC; c++ ) {
                    if ( c < ncl ) {
                        const T xc = bcast( xv[u], c );
#pragma unroll
                        for ( int q = 0; q < NS; q++ ) acc[q] += P[q][c] * xc;
                    }
                }
#pragma unroll
                for ( int q = 0; q < NS; q++ )
                    if ( offd[q] ) pa_atomic_add( xr + grow[q], -acc[q] );
            }
        }
        else {
            T y[PA_SM_RC][NS];
#pragma unroll
            for ( int u = 0; u < PA_SM_RC; u++ )
#pragma unroll
                for ( int q = 0; q < NS; q++ )
                    y[u][q] = ( rok[u] && offd[q] ) ? xo[(int64_t)( r0 + u ) * ldo + grow[q]] : T( 0 );
#pragma unroll
            for ( int u = 0; u < PA_SM_RC; u++ ) {
                T v[PA_SM_NC];
#pragma unroll
                for ( int c = 0; c < PA_SM_NC; c++ ) {
                    v[c] = T( 0 );
#pragma unroll
                    for ( int q = 0; q < NS; q++ ) v[c] += P[q][c] * y[u][q];
                }
                if ( PAIR ) {
                    const T s = pa_half_sum16<T>( v, lane ); /* lane hb + 2c: column c */
                    xv[u] -= __shfl( s, hb + ( ( hl & 15 ) << 1 ) );
                }
                else {
                    const T s = pa_wave_sum16<T>( v, lane ); /* lane 4c: column c */
                    xv[u] -= __shfl( s, ( lane & 15 ) << 2 );
                }
            }
            /* x_k = L_kk^{-T} x_k: lane c holds column c of L_kk in D */
#pragma unroll
            for ( int j = PA_SM_NC - 1; j >= 0; j-- ) {
                if ( j < ncl ) {
#pragma unroll
                    for ( int u = 0; u < PA_SM_RC; u++ ) {
                        const T xj = bcast( xv[u], j ) * dinv[j];
                        if ( hl == j ) xv[u] = xj;
                        else if ( hl < j ) xv[u] -= D[j] * xj; /* L[j][hl] x_j */
                    }
                }
            }
#pragma unroll
            for ( int u = 0; u < PA_SM_RC; u++ )
                if ( rok[u] && hl < nc ) xo[(int64_t)( r0 + u ) * ldo + fcol + hl] = xv[u];
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

/* dst[i] = src[perm[i]] (gather) for nrhs columns; blockIdx.y walks the
 * columns (no 64-bit division per element) */
template <typename TD, typename TS>
__global__ __launch_bounds__( 256 ) void
k_gather( int64_t n, int32_t nrhs, const int32_t *perm, const TS *src, int64_t lds, TD *dst, int64_t ldd )
{
    for ( int32_t r = blockIdx.y; r < nrhs; r += gridDim.y ) {
        const TS *s = src + (int64_t)r * lds;
        TD       *d = dst + (int64_t)r * ldd;
        for ( int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256 ) d[j] = (TD)s[perm[j]];
    }
}
/* Sampled outer product on a CSC pattern (the adjoint of a sparse solve):
 * out[e] = alpha * sum_r u[row(e), r] v[col(e), r] for every stored entry e
 * of column col(e); one thread per column. */
__global__ __launch_bounds__( 256 ) void
k_sampled_outer( int64_t n, const int64_t *colptr, const int64_t *rows, int32_t nrhs, const double *u, int64_t ldu,
                 const double *v, int64_t ldv, double alpha, double *out )
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( c >= n ) return;
    for ( int64_t e = colptr[c]; e < colptr[c + 1]; e++ ) {
        const int64_t r = rows[e];
        double        s = 0.;
        for ( int32_t q = 0; q < nrhs; q++ ) s += u[r + q * ldu] * v[c + q * ldv];
        out[e] = alpha * s;
    }
}

/* x[j] /= D_j for every column j of the factorization (nrhs columns,
 * blockIdx.y walks them), D_j the diagonal entry of the L diagonal block of
 * j's cblk (reference solve_cblk_zdiag, kernels/solve_ztrsmsp.c:664-716) */
template <typename T>
__global__ __launch_bounds__( 256 ) void
k_solve_dscale( pa_dstruct S, int32_t n, const int32_t *col2cblk, const T *Lc, T *x, int32_t nrhs, int64_t ldx )
{
    for ( int32_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256 ) {
        const int32_t k  = col2cblk[j];
        const int32_t nc = S.c_ncol[k];
        const int32_t jl = j - S.c_fcol[k];
        const T       d  = Lc[S.c_coef[k] + (int64_t)jl * ( nc + 1 )];
        for ( int32_t r = blockIdx.y; r < nrhs; r += gridDim.y ) x[j + r * ldx] /= d;
    }
}

template <typename TD, typename TS>
__global__ void
k_convert( int64_t n, const TS *src, TD *dst )
{
    for ( int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 ) dst[i] = (TD)src[i];
}

} /* anonymous namespace */
