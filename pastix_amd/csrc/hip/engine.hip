/*
 * engine.hip -- device state, structure upload, level schedule and the
 * drivers of the factorization, the solve and the FP64 vector operations.
 *
 * The factorization replaces the reference's sequential/static/dynamic
 * schedulers (reference sopalin/sequential_zpotrf.c:51-325) and its
 * StarPU/PaRSEC runtimes by a level schedule: cblks are grouped by their
 * height in the supernodal elimination tree; a level runs as three batched
 * launches (diagonal factorization, panel TRSM, scatter-GEMM).  Every cblk of
 * a level only updates cblks of higher levels, so the launches of a level are
 * dependency-free; updates of cblks of one level that meet in the same target
 * cblk are accumulated with hardware FP atomics.
 */
#include <hip/hip_runtime.h>
#include <vector>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <mutex>
#include "kernels.hpp"

extern "C" {
#include "../host/internal.h"
#include "pa_device.h"
double pa_flops_potrf( double n );
double pa_flops_getrf( double n );
double pa_flops_trsm( double m, double n );
double pa_flops_gemm( double m, double n, double k );
double pa_flops_syrk( double k, double n );
double pa_flops_sytrf( double n );
}

#define PA_HIP( call )                                                                       \
    do {                                                                                     \
        hipError_t _e = ( call );                                                            \
        if ( _e != hipSuccess ) {                                                            \
            fprintf( stderr, "pastix_amd: HIP error %s at %s:%d: %s\n", hipGetErrorName( _e ), \
                     __FILE__, __LINE__, #call );                                             \
            return PASTIX_ERR_INTERNAL;                                                      \
        }                                                                                    \
    } while ( 0 )

/* Scratch of the allocation-free solve (pa_dev_solve_noalloc): the
 * right-hand sides in the factorization numbering (FP64, n*nrhs, padded to
 * 256 bytes), then the second array of the sweeps (FP64 factors), or the
 * FP32 copy (padded to 256 bytes) and the second array in FP32.  One copy
 * per device, held until the process ends and shared by every handle on that
 * device; its address on each device is resolved once (pa_scratch_resolve)
 * and only that address is passed to kernels and copies. */
#define PA_SCRATCH_BYTES ( (size_t)128 << 20 )
#define PA_SCRATCH_MAXDEV 64
static __device__ __align__( 256 ) unsigned char scratch[PA_SCRATCH_BYTES];
static void *pa_scratch_addr[PA_SCRATCH_MAXDEV];

/* Scatter-GEMM tile classes (rho rows x sigma columns of the output tile):
 * a column group narrower than the tile would pad the MFMA work, so the
 * sigma extent follows the group width */
#define PA_NGC 5
/* flops of the current level's GEMM launched ahead of the next level's
 * TRSM in the look-ahead (about 0.3 ms of scatter-GEMM in FP32, twice the
 * next diagonal factorization; measured on the 100^3 FP32 step: 0 117.3,
 * 5e9 113.2, 1e10 111.0, 2e10 111.0, 4e10 111.0 ms per step) */
#define PA_LA_SPLIT_FLOP 2.0e10
static const int pa_gc_tm[PA_NGC] = { 64, 64, 128, 128, 128 };
static const int pa_gc_tn[PA_NGC] = { 64, 32, 128, 64, 32 };

struct pa_level_t {
    int64_t cblk_beg, cblk_end;   /* into d_lev_cblks                       */
    int64_t fac_end;              /* [cblk_beg, fac_end): diag/TRSM/GEMM    */
                                  /* kernels; [fac_end, cblk_end): narrow   */
                                  /* cblks of the fused one-wave kernels,   */
    int64_t sm_end[3];            /* by row slots: [fac_end, sm_end[0]) of  */
                                  /* <= 64 rows, then <= 128, <= 192        */
    int64_t sm_pair_end;          /* [fac_end, sm_pair_end): the <= 64-row  */
                                  /* ones of at most 32 rows, solved two    */
                                  /* per wavefront                          */
    int64_t trsm_beg, trsm_end;   /* into d_trsm_tasks (32-row slabs)       */
    int64_t trsm_mid;             /* [trsm_mid, trsm_end): identity slabs of
                                     the diagonal-block inverses            */
    int64_t tinv_beg, tinv_end;   /* into d_tinv_tasks (diagonal tiles)     */
    bool    trsm_inv;             /* TRSM by inverted tiles (k_trsm_inv)    */
    bool    dinv;                 /* the TRSM launch also inverts the level's
                                     diagonal blocks (solve by products)    */
    int64_t gc_beg[PA_NGC], gc_end[PA_NGC]; /* into d_gemm_tasks, per tile class */
    int64_t gc_split[PA_NGC];     /* [gc_mid, gc_split): the part of the other
                                     tasks launched before the next level's
                                     TRSM is ready (look-ahead)            */
    int64_t gc_mid[PA_NGC];       /* [gc_beg, gc_mid): tasks whose target is a
                                     cblk of the next level (look-ahead)   */
    int64_t solve_beg, solve_end; /* into d_solve_tasks (forward updates)   */
    int64_t dvt_beg, dvt_end;     /* into d_dinv_tasks: 64-entry blocks of
                                     the products with the inverted
                                     diagonal blocks (dinv levels)          */
    int64_t bsolve_beg, bsolve_end; /* into d_bsolve_tasks (backward)       */
    int32_t ncmax;                /* widest cblk of the level               */
    int32_t ncmax_fac;            /* widest cblk of [cblk_beg, fac_end)     */
    double  flops[3];             /* diag/trsm/gemm flops (reference model)  */
};

/* A captured launch sequence and the arguments it was captured with. */
struct pa_graph_entry {
    hipGraph_t     graph = nullptr;
    hipGraphExec_t exec  = nullptr;
    const void    *key[5] = { nullptr, nullptr, nullptr, nullptr, nullptr };
    int64_t        ikey[4] = { 0, 0, 0, 0 };
    uint64_t       stamp = 0;
};
#define PA_SOLVE_GRAPHS 6

/* Tunable choices of the launch schedule (pastix_amd_tune): 128-row GEMM
 * tiles for updates of at least tm_min rows (FP32, and FP64 when tall64:
 * there at most 64 columns wide, their accumulators would not fit a
 * 128 x 128 tile), sigma width following the column-group width (narrow),
 * the fused narrow-cblk kernels (small) */
struct pa_tune_t {
    int32_t tm_min;
    int32_t narrow;
    int32_t small;
    int32_t tall64;
};
static const pa_tune_t pa_tune_default = { 1 << 30, 1, 1, 0 };

struct pa_wblock {
    void  *p;
    size_t bytes;
    int    used;
};

struct pa_device_s {
    int         device;
    hipStream_t stream;
    hipStream_t stream2;          /* look-ahead panels of the next level     */
    int         lds_limit;
    /* structure */
    int32_t *d_c_fcol, *d_c_ncol, *d_c_stride, *d_c_fblok;
    int64_t *d_c_coef;
    int32_t *d_b_frow, *d_b_lrow, *d_b_fcblk, *d_b_rowoff;
    int32_t *d_col2cblk, *d_perm, *d_invp;
    int64_t *d_c_dtoff, *d_r_base;
    int4    *d_r_info;
    int32_t *d_r_blk;
    int64_t  dtnbr;      /* entries of the transposed diagonal blocks   */
    pa_dstruct S;
    int64_t  n, coefnbr;
    int32_t  colmax;
    /* schedule */
    std::vector<pa_level_t> levels;
    int32_t      *d_lev_cblks;
    int2         *d_trsm_tasks;
    pa_gemm_task *d_gemm_tasks;
    int4         *d_solve_tasks, *d_bsolve_tasks;
    int2         *d_dinv_tasks = nullptr;
    /* TRSM by inverted diagonal tiles: first tile of each cblk (-1: none),
     * the (cblk, tile) list of every level and the inverses (one region per
     * TRSM mode: two for LU) */
    int64_t      *d_c_tinv = nullptr;
    int2         *d_tinv_tasks = nullptr;
    void         *d_tinv = nullptr;
    int64_t       tinv_tiles = 0;
    /* inverses of the lower triangles of the diagonal blocks (offsets
     * c_dtoff, column major), one region per factor array: L, and the U^T
     * array of LU; written by the identity tasks of the TRSM launches */
    void         *d_dinv = nullptr;
    bool          all_dinv = false;   /* every level inverts its diagonal blocks */
    /* matrix */
    int64_t *d_colptr;
    int32_t *d_rowtab;
    double  *d_Lv, *d_Uv;
    /* value map of the permuted CSC (device-resident value updates) */
    int64_t *d_vm_Lptr, *d_vm_Lsrc, *d_vm_Uptr, *d_vm_Usrc;
    /* factors */
    int   prec, lu;   /* lu: two factor arrays (LU: U^T; LDL^t: the D L^t copy) */
    int   ldlt;
    void *d_L, *d_U;
    void *d_DT;          /* LL^t: transposed diagonal blocks (solve)     */
    /* hipGraphs of the factorization and of recent solves (the launch
     * sequences are long: ~5 launches per level) */
    int            use_graphs = 1;
    int            lookahead  = 1;     /* PASTIX_AMD_LOOKAHEAD=0 disables      */
    std::vector<hipEvent_t> la_ev;     /* fork / join events, 2 per level    */
    pa_graph_entry fact_graph;
    pa_graph_entry solve_graphs[PA_SOLVE_GRAPHS];
    uint64_t       graph_clock = 0;
    int32_t *d_counters;
    double  *d_crit;     /* static pivoting criterion of the next factorization */
    /* work */
    double *d_partial;
    double *h_partial;
    float  *d_xs;      /* FP32 solve buffer of the mixed path */
    int64_t xs_len;
    void   *d_alt = nullptr;   /* second array of the solve sweeps       */
    size_t  alt_bytes = 0;
    double *d_xtmp;    /* permuted rhs buffer */
    int64_t xtmp_len;
    void   *d_scratch = nullptr;   /* this device's `scratch` array */
    std::vector<pa_wblock> wpool;   /* pa_dev_wget / pa_dev_wput */
    /* host copy of the structure the schedule is built from */
    std::vector<int32_t> h_c_ncol, h_c_stride, h_c_fblok, h_b_frow, h_b_lrow, h_b_fcblk, h_b_rowoff;
    pa_tune_t tune = pa_tune_default;
    /* factorization kind and arithmetic the task lists were built for: the
     * schedule is rebuilt when the parameters change after the analysis */
    int sched_fact = -1, sched_big = -1;
    int fact_kind = -1;   /* factorization kind of the factors in HBM        */
    hipEvent_t ev0, ev1;
    double  last_fact_event_s;
    /* per kernel class profiling (diag, trsm, gemm) with HIP events */
    int     profiling;
    std::vector<hipEvent_t> prof_ev;   /* 6 per level */
    double  prof_ms[3];
    int64_t prof_launch[3];
    double  class_flops[3];            /* algorithmic flops of a full factorization */
    double  real_flops = 0.;           /* flops the kernels perform: the GEMM tiles
                                          padded to their class, diagonal and TRSM
                                          as the model */
};

static inline pa_device_s *DEV( const pastix_data_t *d ) { return (pa_device_s *)d->dev; }

/* LDL^t in real arithmetic: PastixFactLDLT and PastixFactLDLH */
static inline bool pa_is_ldlt( int fact ) { return fact == PastixFactLDLT || fact == PastixFactLDLH; }

/* FP32 factors (mixed precision or a single precision matrix) */
static inline int pa_sched_big( const pastix_data_t *pd )
{
    return ( pd->iparm[IPARM_MIXED] != 0 ) || ( pd->iparm[IPARM_FLOAT] == SpmFloat );
}

template <typename T>
static int
pa_upload( T **dptr, const std::vector<T> &h )
{
    size_t bytes = std::max<size_t>( h.size(), 1 ) * sizeof( T );
    PA_HIP( hipMalloc( (void **)dptr, bytes ) );
    if ( !h.empty() ) PA_HIP( hipMemcpy( *dptr, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice ) );
    return PASTIX_SUCCESS;
}

/* allow the kernels that stage a slab in LDS to use up to the device limit */
template <typename K>
static void
pa_lds_attr( K kernel, int bytes )
{
    /* the dynamic limit excludes the kernel's static LDS; a refused request
     * must not leave a sticky error behind for the next runtime call of the
     * process (the HIP runtime is shared with the caller's framework) */
    hipFuncAttributes fa;
    int               stat = 0;
    if ( hipFuncGetAttributes( &fa, reinterpret_cast<const void *>( kernel ) ) == hipSuccess ) stat = (int)fa.sharedSizeBytes;
    (void)hipGetLastError();
    if ( hipFuncSetAttribute( reinterpret_cast<const void *>( kernel ), hipFuncAttributeMaxDynamicSharedMemorySize,
                              bytes - stat ) != hipSuccess ) {
        (void)hipGetLastError();
    }
}

template <typename T, bool UNIT>
static void
pa_lds_attr_sd( int bytes )
{
    if ( !UNIT ) {
        pa_lds_attr( k_solve_dinv<T, false, 256>, bytes ); pa_lds_attr( k_solve_dinv<T, true, 256>, bytes );
        pa_lds_attr( k_solve_dinv_cblk<T, false, 64>, bytes );  pa_lds_attr( k_solve_dinv_cblk<T, true, 64>, bytes );
        pa_lds_attr( k_solve_dinv_cblk<T, false, 128>, bytes ); pa_lds_attr( k_solve_dinv_cblk<T, true, 128>, bytes );
        pa_lds_attr( k_solve_dinv_cblk<T, false, 256>, bytes ); pa_lds_attr( k_solve_dinv_cblk<T, true, 256>, bytes );
        pa_lds_attr( k_solve_dinv_cblk<T, false, 512>, bytes ); pa_lds_attr( k_solve_dinv_cblk<T, true, 512>, bytes );
        pa_lds_attr( k_solve_dinv<T, false, 512>, bytes ); pa_lds_attr( k_solve_dinv<T, true, 512>, bytes );
    }
    pa_lds_attr( k_solve_diag_fwd<T, UNIT, 64>, bytes );  pa_lds_attr( k_solve_diag_bwd<T, UNIT, 64>, bytes );
    pa_lds_attr( k_solve_diag_fwd<T, UNIT, 128>, bytes ); pa_lds_attr( k_solve_diag_bwd<T, UNIT, 128>, bytes );
    pa_lds_attr( k_solve_diag_fwd<T, UNIT, 256>, bytes ); pa_lds_attr( k_solve_diag_bwd<T, UNIT, 256>, bytes );
    pa_lds_attr( k_solve_diag_fwd<T, UNIT, 512>, bytes ); pa_lds_attr( k_solve_diag_bwd<T, UNIT, 512>, bytes );
}

static void
pa_set_lds_limits( int bytes )
{
    pa_lds_attr( k_trsm<double, 0>, bytes ); pa_lds_attr( k_trsm<double, 1>, bytes ); pa_lds_attr( k_trsm<double, 2>, bytes );
    pa_lds_attr( k_trsm<float, 0>, bytes );  pa_lds_attr( k_trsm<float, 1>, bytes );  pa_lds_attr( k_trsm<float, 2>, bytes );
    pa_lds_attr_sd<double, false>( bytes ); pa_lds_attr_sd<double, true>( bytes );
    pa_lds_attr_sd<float, false>( bytes );  pa_lds_attr_sd<float, true>( bytes );
    pa_lds_attr( k_solve_diag<double, false, false>, bytes ); pa_lds_attr( k_solve_diag<double, false, true>, bytes );
    pa_lds_attr( k_solve_diag<double, true, false>, bytes );  pa_lds_attr( k_solve_diag<double, true, true>, bytes );
    pa_lds_attr( k_solve_diag<float, false, false>, bytes );  pa_lds_attr( k_solve_diag<float, false, true>, bytes );
    pa_lds_attr( k_solve_diag<float, true, false>, bytes );   pa_lds_attr( k_solve_diag<float, true, true>, bytes );
    pa_lds_attr( k_diag<double, true, 64>, bytes );  pa_lds_attr( k_diag<double, false, 64>, bytes );
    pa_lds_attr( k_diag<float, true, 1024>, bytes );  pa_lds_attr( k_diag<float, false, 1024>, bytes );
    pa_lds_attr( k_diag<double, true, 1024>, bytes ); pa_lds_attr( k_diag<double, false, 1024>, bytes );
    pa_lds_attr( k_diag<float, true, 64>, bytes );   pa_lds_attr( k_diag<float, false, 64>, bytes );
    pa_lds_attr( k_diag<double, true, 64, true>, bytes );   pa_lds_attr( k_diag<double, true, 1024, true>, bytes );
    pa_lds_attr( k_diag<float, true, 64, true>, bytes );    pa_lds_attr( k_diag<float, true, 1024, true>, bytes );
}

/* Device address of `scratch` on `device` (the current device), looked up
 * once per device and kept: each device holds its own copy of the array, and
 * the host-side name is not a device address. */
static int
pa_scratch_resolve( int device, void **addr )
{
    static std::mutex lock;
    if ( device < 0 || device >= PA_SCRATCH_MAXDEV ) {
        fprintf( stderr, "pastix_amd: device %d is beyond the %d devices the solve scratch serves\n", device,
                 PA_SCRATCH_MAXDEV );
        return PASTIX_ERR_BADPARAMETER;
    }
    std::lock_guard<std::mutex> g( lock );
    if ( !pa_scratch_addr[device] ) {
        void      *p = nullptr;
        hipError_t e = hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) );
        if ( e != hipSuccess || p == nullptr ) {
            fprintf( stderr, "pastix_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
                     hipGetErrorName( e ) );
            (void)hipGetLastError();
            return PASTIX_ERR_INTERNAL;
        }
        pa_scratch_addr[device] = p;
    }
    *addr = pa_scratch_addr[device];
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_open( pastix_data_t *pd )
{
    int ndev = 0;
    if ( pd->dev != NULL ) return PASTIX_SUCCESS;
    if ( hipGetDeviceCount( &ndev ) != hipSuccess || ndev < 1 ) {
        fprintf( stderr, "pastix_amd: no HIP device available; this engine has no CPU fallback\n" );
        return PASTIX_ERR_INTERNAL;
    }
    pa_device_s *dev = new pa_device_s();
    memset( (void *)&dev->S, 0, sizeof( dev->S ) );
    dev->device = 0;
    const char *env = getenv( "PASTIX_AMD_DEVICE" );
    const char *genv = getenv( "PASTIX_AMD_GRAPHS" );
    if ( genv && atoi( genv ) == 0 ) dev->use_graphs = 0;
    if ( env ) dev->device = atoi( env );
    PA_HIP( hipSetDevice( dev->device ) );
    if ( pa_scratch_resolve( dev->device, &dev->d_scratch ) != PASTIX_SUCCESS ) {
        delete dev;
        return PASTIX_ERR_INTERNAL;
    }
    PA_HIP( hipStreamCreate( &dev->stream ) );
    PA_HIP( hipStreamCreate( &dev->stream2 ) );
    {
        const char *la = getenv( "PASTIX_AMD_LOOKAHEAD" );
        if ( la && atoi( la ) == 0 ) dev->lookahead = 0;
    }
    PA_HIP( hipEventCreate( &dev->ev0 ) );
    PA_HIP( hipEventCreate( &dev->ev1 ) );
    int lds = 65536;
    hipDeviceGetAttribute( &lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev->device );
    dev->lds_limit = lds;
    pa_set_lds_limits( lds );
    PA_HIP( hipMalloc( (void **)&dev->d_counters, 16 * sizeof( int32_t ) ) );
    PA_HIP( hipMalloc( (void **)&dev->d_crit, sizeof( double ) ) );
    PA_HIP( hipMalloc( (void **)&dev->d_partial, 4096 * sizeof( double ) ) );
    PA_HIP( hipHostMalloc( (void **)&dev->h_partial, 4096 * sizeof( double ) ) );
    pd->dev = dev;
    return PASTIX_SUCCESS;
}

static void
pa_graph_drop( pa_graph_entry &g )
{
    if ( g.exec ) hipGraphExecDestroy( g.exec );
    if ( g.graph ) hipGraphDestroy( g.graph );
    g = pa_graph_entry();
}

static void
pa_graphs_drop( pa_device_s *dev )
{
    pa_graph_drop( dev->fact_graph );
    for ( auto &g : dev->solve_graphs ) pa_graph_drop( g );
}

/* Run `launch()` on the library stream through a cached hipGraph captured
 * from the same launch sequence (capture on the first call with a new key). */
template <typename F>
static int
pa_graph_run( pa_device_s *dev, pa_graph_entry &g, bool hit, F launch )
{
    if ( !hit ) {
        pa_graph_drop( g );
        hipGraph_t graph = nullptr;
        PA_HIP( hipStreamBeginCapture( dev->stream, hipStreamCaptureModeRelaxed ) );
        int rc = launch();
        hipError_t ec = hipStreamEndCapture( dev->stream, &graph );
        if ( rc ) {
            if ( graph ) hipGraphDestroy( graph );
            return rc;
        }
        PA_HIP( ec );
        hipError_t ei = hipGraphInstantiate( &g.exec, graph, nullptr, nullptr, 0 );
        if ( ei != hipSuccess ) {
            hipGraphDestroy( graph );
            g.exec = nullptr;
            PA_HIP( ei );
        }
        g.graph = graph;
    }
    PA_HIP( hipGraphLaunch( g.exec, dev->stream ) );
    return PASTIX_SUCCESS;
}

static void
pa_free_matrix( pa_device_s *dev )
{
    hipFree( dev->d_colptr ); hipFree( dev->d_rowtab ); hipFree( dev->d_Lv ); hipFree( dev->d_Uv );
    dev->d_colptr = NULL; dev->d_rowtab = NULL; dev->d_Lv = NULL; dev->d_Uv = NULL;
    hipFree( dev->d_vm_Lptr ); hipFree( dev->d_vm_Lsrc ); hipFree( dev->d_vm_Uptr ); hipFree( dev->d_vm_Usrc );
    dev->d_vm_Lptr = dev->d_vm_Lsrc = dev->d_vm_Uptr = dev->d_vm_Usrc = NULL;
}

static void
pa_free_factors( pa_device_s *dev )
{
    pa_graphs_drop( dev );
    hipFree( dev->d_L ); hipFree( dev->d_U ); hipFree( dev->d_DT );
    dev->d_L = dev->d_U = dev->d_DT = NULL;
}

static void
pa_free_structure( pa_device_s *dev )
{
    pa_graphs_drop( dev );
    void *ptrs[] = { dev->d_c_fcol, dev->d_c_ncol, dev->d_c_stride, dev->d_c_fblok, dev->d_c_coef,
                     dev->d_b_frow, dev->d_b_lrow, dev->d_b_fcblk, dev->d_b_rowoff, dev->d_col2cblk,
                     dev->d_perm, dev->d_invp, dev->d_lev_cblks, dev->d_trsm_tasks, dev->d_gemm_tasks,
                     dev->d_solve_tasks, dev->d_bsolve_tasks, dev->d_c_dtoff, dev->d_r_base, dev->d_r_info,
                     dev->d_r_blk };
    for ( void *p : ptrs ) hipFree( p );
    hipFree( dev->d_c_tinv ); hipFree( dev->d_tinv_tasks ); hipFree( dev->d_tinv ); hipFree( dev->d_dinv );
    dev->d_c_tinv = NULL; dev->d_tinv_tasks = NULL; dev->d_tinv = NULL; dev->d_dinv = NULL;
    hipFree( dev->d_dinv_tasks );
    dev->d_dinv_tasks = NULL;
    dev->tinv_tiles = 0;
    dev->S.c_tinv = NULL;
    dev->d_c_fcol = dev->d_c_ncol = dev->d_c_stride = dev->d_c_fblok = NULL;
    dev->d_c_coef = NULL;
    dev->d_b_frow = dev->d_b_lrow = dev->d_b_fcblk = dev->d_b_rowoff = NULL;
    dev->d_col2cblk = dev->d_perm = dev->d_invp = NULL;
    dev->d_lev_cblks = NULL; dev->d_trsm_tasks = NULL; dev->d_gemm_tasks = NULL; dev->d_solve_tasks = NULL;
    dev->d_bsolve_tasks = NULL;
    dev->d_c_dtoff = dev->d_r_base = NULL;
    dev->d_r_info = NULL;
    dev->d_r_blk  = NULL;
    dev->levels.clear();
}

extern "C" void
pa_dev_close( pastix_data_t *pd )
{
    pa_device_s *dev = DEV( pd );
    if ( !dev ) return;
    hipSetDevice( dev->device );
    hipStreamSynchronize( dev->stream );
    pa_graphs_drop( dev );
    pa_free_matrix( dev );
    pa_free_factors( dev );
    pa_free_structure( dev );
    for ( const pa_wblock &w : dev->wpool ) hipFree( w.p );
    dev->wpool.clear();
    hipFree( dev->d_counters );
    hipFree( dev->d_crit );
    hipFree( dev->d_partial );
    hipHostFree( dev->h_partial );
    hipFree( dev->d_xs );
    hipFree( dev->d_alt );
    hipFree( dev->d_xtmp );
    hipEventDestroy( dev->ev0 );
    hipEventDestroy( dev->ev1 );
    hipStreamDestroy( dev->stream );
    hipStreamDestroy( dev->stream2 );
    for ( hipEvent_t e : dev->la_ev ) hipEventDestroy( e );
    delete dev;
    pd->dev = NULL;
}

/* ------------------------------------------------------------------------- */
/* structure + schedule                                                      */
/* ------------------------------------------------------------------------- */
static int pa_build_schedule( pastix_data_t *pd );

extern "C" int
pa_dev_setup_structure( pastix_data_t *pd )
{
    pa_device_s        *dev  = DEV( pd );
    const SolverMatrix *solv = pd->solvmatr;
    const pastix_int_t  cblknbr = solv->cblknbr, bloknbr = solv->bloknbr, n = solv->nodenbr;
    int rc;

    if ( !dev ) return PASTIX_ERR_BADPARAMETER;
    if ( n >= ( 1LL << 31 ) || bloknbr >= ( 1LL << 31 ) ) {
        fprintf( stderr, "pastix_amd: problem too large for 32-bit device row indices\n" );
        return PASTIX_ERR_INTEGER_TYPE;
    }
    PA_HIP( hipSetDevice( dev->device ) );
    pa_free_structure( dev );
    pa_free_factors( dev );
    dev->prec = 0;

    std::vector<int32_t> c_fcol( cblknbr + 1 ), c_ncol( cblknbr + 1 ), c_stride( cblknbr + 1 ), c_fblok( cblknbr + 1 );
    std::vector<int64_t> c_coef( cblknbr + 1 );
    std::vector<int32_t> b_frow( bloknbr + 1 ), b_lrow( bloknbr + 1 ), b_fcblk( bloknbr + 1 ), b_rowoff( bloknbr + 1 );
    std::vector<int32_t> col2cblk( n + 1 ), perm( n + 1 ), invp( n + 1 );
    std::vector<int64_t> c_dtoff( cblknbr + 1 ), r_base( cblknbr + 1 );
    std::vector<int4>    r_info;
    std::vector<int32_t> r_blk;
    int64_t coef = 0, dtoff = 0, rbase = 0;
    int32_t colmax = 0;
    for ( pastix_int_t k = 0; k < cblknbr; k++ ) {
        const SolverCblk *c = solv->cblktab + k;
        c_fcol[k]   = (int32_t)c->fcolnum;
        c_ncol[k]   = (int32_t)cblk_colnbr( c );
        c_stride[k] = (int32_t)c->stride;
        c_fblok[k]  = (int32_t)( c->fblokptr - solv->bloktab );
        c_coef[k]   = coef;
        c_dtoff[k]  = dtoff;
        r_base[k]   = rbase;
        dtoff += (int64_t)c_ncol[k] * c_ncol[k];
        rbase += c->stride - c_ncol[k];
        colmax = std::max( colmax, c_ncol[k] );
        for ( pastix_int_t j = c->fcolnum; j <= c->lcolnum; j++ ) col2cblk[j] = (int32_t)k;
        int32_t rowoff = 0;
        for ( const SolverBlok *b = c->fblokptr; b < c[1].fblokptr; b++ ) {
            pastix_int_t bi = b - solv->bloktab;
            b_frow[bi]   = (int32_t)b->frownum;
            b_lrow[bi]   = (int32_t)b->lrownum;
            b_fcblk[bi]  = (int32_t)b->fcblknm;
            b_rowoff[bi] = rowoff;
            if ( b > c->fblokptr ) {
                /* descriptors of the off-diagonal rows, in panel order */
                const int32_t M = (int32_t)blok_rownbr( b );
                for ( int32_t loc = 0; loc < M; loc++ ) {
                    int64_t off = coef + (int64_t)rowoff * c_ncol[k] + loc;
                    r_info.push_back( make_int4( (int32_t)b->frownum + loc, M, (int32_t)( off & 0xffffffffLL ),
                                                 (int32_t)( off >> 32 ) ) );
                    r_blk.push_back( (int32_t)bi );
                }
            }
            rowoff += (int32_t)blok_rownbr( b );
        }
        coef += (int64_t)c->stride * c_ncol[k];
    }
    c_fcol[cblknbr]  = (int32_t)n;
    c_fblok[cblknbr] = (int32_t)bloknbr;
    c_coef[cblknbr]  = coef;
    c_dtoff[cblknbr] = dtoff;
    r_base[cblknbr]  = rbase;
    for ( pastix_int_t i = 0; i < n; i++ ) {
        perm[i] = (int32_t)pd->ordemesh->permtab[i];
        invp[i] = (int32_t)pd->ordemesh->peritab[i];
    }
    dev->n       = n;
    dev->coefnbr = coef;
    dev->dtnbr   = dtoff;
    dev->colmax  = colmax;

    if ( ( rc = pa_upload( &dev->d_c_fcol, c_fcol ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_c_ncol, c_ncol ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_c_stride, c_stride ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_c_fblok, c_fblok ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_c_coef, c_coef ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_b_frow, b_frow ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_b_lrow, b_lrow ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_b_fcblk, b_fcblk ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_b_rowoff, b_rowoff ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_col2cblk, col2cblk ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_perm, perm ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_invp, invp ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_c_dtoff, c_dtoff ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_r_base, r_base ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_r_info, r_info ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_r_blk, r_blk ) ) ) return rc;

    pa_dstruct &S = dev->S;
    S.cblknbr  = (int32_t)cblknbr;
    S.bloknbr  = (int32_t)bloknbr;
    S.c_fcol   = dev->d_c_fcol;
    S.c_ncol   = dev->d_c_ncol;
    S.c_stride = dev->d_c_stride;
    S.c_fblok  = dev->d_c_fblok;
    S.c_coef   = dev->d_c_coef;
    S.b_frow   = dev->d_b_frow;
    S.b_lrow   = dev->d_b_lrow;
    S.b_fcblk  = dev->d_b_fcblk;
    S.b_rowoff = dev->d_b_rowoff;
    S.c_dtoff  = dev->d_c_dtoff;
    S.r_base   = dev->d_r_base;
    S.r_info   = dev->d_r_info;
    S.r_blk    = dev->d_r_blk;

    dev->h_c_ncol   = std::move( c_ncol );
    dev->h_c_stride = std::move( c_stride );
    dev->h_c_fblok  = std::move( c_fblok );
    dev->h_b_frow   = std::move( b_frow );
    dev->h_b_lrow   = std::move( b_lrow );
    dev->h_b_fcblk  = std::move( b_fcblk );
    dev->h_b_rowoff = std::move( b_rowoff );
    return pa_build_schedule( pd );
}

/* The level schedule and the task lists of every kernel class, from the
 * host copy of the structure and the tuning choices in dev->tune. */
static int
pa_build_schedule( pastix_data_t *pd )
{
    pa_device_s        *dev  = DEV( pd );
    const SolverMatrix *solv = pd->solvmatr;
    const pastix_int_t  cblknbr = solv->cblknbr;
    const int32_t       colmax  = dev->colmax;
    int rc;
    const std::vector<int32_t> &c_ncol = dev->h_c_ncol, &c_stride = dev->h_c_stride, &c_fblok = dev->h_c_fblok;
    const std::vector<int32_t> &b_frow = dev->h_b_frow, &b_lrow = dev->h_b_lrow, &b_fcblk = dev->h_b_fcblk;
    const std::vector<int32_t> &b_rowoff = dev->h_b_rowoff;
    pa_graphs_drop( dev );
    {
        void *ptrs[] = { dev->d_lev_cblks, dev->d_trsm_tasks, dev->d_gemm_tasks, dev->d_solve_tasks, dev->d_bsolve_tasks };
        for ( void *p : ptrs ) hipFree( p );
        dev->d_lev_cblks = NULL; dev->d_trsm_tasks = NULL; dev->d_gemm_tasks = NULL; dev->d_solve_tasks = NULL;
        dev->d_bsolve_tasks = NULL;
        hipFree( dev->d_c_tinv ); hipFree( dev->d_tinv_tasks ); hipFree( dev->d_tinv ); hipFree( dev->d_dinv );
        dev->d_c_tinv = NULL; dev->d_tinv_tasks = NULL; dev->d_tinv = NULL; dev->d_dinv = NULL;
        hipFree( dev->d_dinv_tasks );
        dev->d_dinv_tasks = NULL;
        dev->tinv_tiles = 0;
        dev->levels.clear();
    }

    /* ---- level schedule ---- */
    std::vector<pastix_int_t> level( cblknbr + 1 );
    pastix_int_t nlev = cblknbr ? pa_solver_levels( solv, level.data() ) : 0;
    std::vector<std::vector<int32_t>> bylev( nlev );
    for ( pastix_int_t k = 0; k < cblknbr; k++ ) bylev[level[k]].push_back( (int32_t)k );

    const int fact = (int)pd->iparm[IPARM_FACTORIZATION];
    const int ldlt = pa_is_ldlt( fact );
    /* lu: the U-side updates of an LU factorization; two: a second factor
     * array (U^T for LU, the D L^t copy of the panels for LDL^t) */
    const int lu   = ( fact == PastixFactLU );
    const int two  = lu || ldlt;
    /* 128 x 128 GEMM tiles pay off in FP32; in FP64 their accumulators
     * (64 doubles per lane) leave one wave per SIMD, so 64 x 64 is kept */
    const int big_tiles = pa_sched_big( pd );
    dev->sched_fact = (int)pd->iparm[IPARM_FACTORIZATION];
    dev->sched_big  = big_tiles;
    std::vector<int32_t>      lev_cblks;
    std::vector<int2>         trsm_tasks, tinv_tasks;
    std::vector<int64_t>      c_tinv( std::max<pastix_int_t>( cblknbr, 1 ), -1 );
    int64_t                   tinv_tiles = 0;
    /* PASTIX_AMD_TRSM=0 keeps the substitution TRSM (k_trsm) on every level */
    const char *tr_env  = getenv( "PASTIX_AMD_TRSM" );
    const bool  tr_inv  = !( tr_env && tr_env[0] == '0' );
    /* PASTIX_AMD_DINV=0 keeps the substitution diagonal solves */
    const char *di_env  = getenv( "PASTIX_AMD_DINV" );
    const bool  dinv_on = !( di_env && di_env[0] == '0' );
    /* PASTIX_AMD_LA_SPLIT: flops of the look-ahead's first GEMM part */
    const char  *ls_env   = getenv( "PASTIX_AMD_LA_SPLIT" );
    const double la_split = ( ls_env && atof( ls_env ) >= 0. ) ? atof( ls_env ) : PA_LA_SPLIT_FLOP;
    bool        any_dinv = false, all_dinv = !ldlt && dinv_on;
    std::vector<int2> ident_tasks, dinv_tasks;
    std::vector<int4>         solve_tasks, bsolve_tasks;
    std::vector<pa_gemm_task> gtasks[PA_NGC];
    std::vector<int32_t>      srccount( cblknbr + 1, 0 ), lastsrc( cblknbr + 1, -1 );
    dev->levels.resize( nlev );

    /* narrow cblks (LL^t) are factored and applied by one wavefront each
     * (k_small_llt); PASTIX_AMD_SMALL=0 sends them through the general
     * kernels (same results up to rounding) */
    const char *sm_env = getenv( "PASTIX_AMD_SMALL" );
    const bool  use_small = !two && dev->tune.small && !( sm_env && sm_env[0] == '0' );
    auto is_small = [&]( pastix_int_t k ) {
        return use_small && c_ncol[k] <= PA_SM_NC && c_stride[k] <= PA_SM_ROWS;
    };
    /* reference flop model (symbol/symbol_cost_flops.c, common/flops.h);
     * the narrow cblks' flops all count in the diagonal class, whose
     * launch slot runs their fused kernel */
    auto cblk_flops = [&]( pastix_int_t k, double *f ) {
        double N = c_ncol[k], M = c_stride[k] - c_ncol[k];
        if ( lu ) {
            f[0] = pa_flops_getrf( N );
            f[1] = 2. * pa_flops_trsm( M, N );
            f[2] = pa_flops_gemm( M, M, N );
        }
        else if ( ldlt ) {
            f[0] = pa_flops_sytrf( N );
            f[1] = pa_flops_trsm( M, N ) + N * M;
            f[2] = pa_flops_syrk( N, M ) + M * M;
        }
        else {
            f[0] = pa_flops_potrf( N );
            f[1] = pa_flops_trsm( M, N );
            f[2] = pa_flops_syrk( N, M );
        }
        if ( is_small( k ) ) {
            f[0] += f[1] + f[2];
            f[1] = f[2] = 0.;
        }
    };
    dev->class_flops[0] = dev->class_flops[1] = dev->class_flops[2] = 0.;
    double gemm_real = 0.;
    for ( pastix_int_t k = 0; k < cblknbr; k++ ) {
        double f[3];
        cblk_flops( k, f );
        for ( int c = 0; c < 3; c++ ) dev->class_flops[c] += f[c];
    }
    for ( pastix_int_t l = 0; l < nlev; l++ ) {
        pa_level_t &L = dev->levels[l];
        L.cblk_beg  = lev_cblks.size();
        L.trsm_beg  = trsm_tasks.size();
        L.tinv_beg  = tinv_tasks.size();
        for ( int c = 0; c < PA_NGC; c++ ) L.gc_beg[c] = gtasks[c].size();
        L.solve_beg = solve_tasks.size();
        L.bsolve_beg = bsolve_tasks.size();
        L.dvt_beg = dinv_tasks.size();
        L.ncmax     = 0;
        L.ncmax_fac = 0;
        L.flops[0] = L.flops[1] = L.flops[2] = 0.;
        /* the level's cblks: general ones first, then the narrow ones */
        for ( int32_t k : bylev[l] )
            if ( !is_small( k ) ) {
                lev_cblks.push_back( k );
                L.ncmax_fac = std::max( L.ncmax_fac, c_ncol[k] );
            }
        L.fac_end = lev_cblks.size();
        /* narrow cblks by row slots; those of at most 32 rows first, by
         * width, so that the two cblks of a wavefront in the solve have
         * about the same number of columns */
        {
            const size_t p0 = lev_cblks.size();
            for ( int32_t k : bylev[l] )
                if ( is_small( k ) && c_stride[k] <= 32 ) lev_cblks.push_back( k );
            std::stable_sort( lev_cblks.begin() + p0, lev_cblks.end(),
                              [&]( int32_t a, int32_t b ) { return c_ncol[a] < c_ncol[b]; } );
        }
        L.sm_pair_end = lev_cblks.size();
        for ( int ns = 1; ns <= 3; ns++ ) {
            for ( int32_t k : bylev[l] )
                if ( is_small( k ) && ( c_stride[k] + 63 ) / 64 == ns && ( ns > 1 || c_stride[k] > 32 ) )
                    lev_cblks.push_back( k );
            L.sm_end[ns - 1] = lev_cblks.size();
        }
        /* number of distinct source cblks of this level per target cblk */
        std::vector<int32_t> touched;
        for ( int32_t k : bylev[l] ) {
            L.ncmax = std::max( L.ncmax, c_ncol[k] );
            {
                double f[3];
                cblk_flops( k, f );
                for ( int c = 0; c < 3; c++ ) L.flops[c] += f[c];
            }
            for ( int32_t b = c_fblok[k] + 1; b < c_fblok[k + 1]; b++ ) {
                int32_t F = b_fcblk[b];
                if ( lastsrc[F] != k ) {
                    lastsrc[F] = k;
                    if ( srccount[F]++ == 0 ) touched.push_back( F );
                }
            }
        }
        /* TRSM by inverted tiles for the level; the same launch then also
         * inverts the diagonal blocks for the solve (LL^t and LU; LDL^t
         * solves by substitution) */
        const bool lev_inv  = tr_inv && L.ncmax_fac <= PA_TI_MAXNC;
        const bool lev_dinv = lev_inv && !ldlt && dinv_on;
        for ( int32_t k : bylev[l] ) {
            int32_t nc = c_ncol[k], R = c_stride[k] - nc;
            if ( is_small( k ) ) continue; /* k_small_llt / k_small_solve */
            /* solve update tasks: forward 256 rows x 64 columns; backward
             * 16 ncg columns x 1024 / ncg rows (4 waves of 16 x 256) */
            for ( int32_t r = 0; r < R; r += PA_SF_ROWS )
                for ( int32_t c = 0; c < nc; c += PA_SF_COLS ) solve_tasks.push_back( make_int4( k, r, c, 0 ) );
            const int32_t ncg = PA_SB_NCG( nc );
            for ( int32_t r = 0; r < R; r += 1024 / ncg )
                for ( int32_t c = 0; c < nc; c += 16 * ncg ) bsolve_tasks.push_back( make_int4( k, r, c, 0 ) );
            for ( int32_t r = 0; r < R; r += PA_TRSM_TR ) trsm_tasks.push_back( make_int2( k, r ) );
            /* identity rows r0.. of the diagonal block, as (k, -1 - r0),
             * after the level's panel slabs */
            if ( lev_dinv ) {
                for ( int32_t r = 0; r < nc; r += PA_TRSM_TR ) ident_tasks.push_back( make_int2( k, -1 - r ) );
                for ( int32_t b = 0; 64 * b < nc; b++ ) dinv_tasks.push_back( make_int2( k, b ) );
            }
            if ( R > 0 || lev_dinv ) {
                c_tinv[k] = tinv_tiles;
                for ( int32_t tl = 0; tl < ( nc + 15 ) / 16; tl++ ) tinv_tasks.push_back( make_int2( k, tl ) );
                tinv_tiles += ( nc + 15 ) / 16;
            }
            /* column groups: consecutive blocks facing the same cblk */
            int32_t b = c_fblok[k] + 1, lb = c_fblok[k + 1];
            while ( b < lb ) {
                int32_t F = b_fcblk[b], e = b;
                while ( e + 1 < lb && b_fcblk[e + 1] == F ) e++;
                int32_t g0 = b_rowoff[b] - nc;
                int32_t g1 = b_rowoff[e] - nc + ( b_lrow[e] - b_frow[e] + 1 );
                int32_t flags = ( srccount[F] > 1 ) ? PA_TASK_ATOMIC : 0;
                /* tall updates use 128-row output tiles (FP32 only: in FP64
                 * their accumulators leave one wave per SIMD), the others 64;
                 * the tile width follows the remaining columns of the group */
                const bool    tall = big_tiles || dev->tune.tall64;
                const int32_t TM   = ( tall && R - g0 >= dev->tune.tm_min ) ? 128 : 64;
                for ( int side = 0; side < ( lu ? 2 : 1 ); side++ ) {
                    /* U side needs rows of strictly later blocks only */
                    if ( side == 1 && e + 1 >= lb && b == e ) continue; /* single last block */
                    for ( int32_t s = g0; s < g1; ) {
                        const int32_t rem = g1 - s;
                        int32_t TN, cls;
                        if ( TM == 128 && !big_tiles ) {
                            TN  = ( rem > 32 || !dev->tune.narrow ) ? 64 : 32;
                            cls = TN == 64 ? 3 : 4;
                        }
                        else if ( TM == 128 ) {
                            TN  = ( rem > 64 || !dev->tune.narrow ) ? 128 : ( rem > 32 ? 64 : 32 );
                            cls = TN == 128 ? 2 : ( TN == 64 ? 3 : 4 );
                        }
                        else {
                            TN  = ( rem > 32 || !dev->tune.narrow ) ? 64 : 32;
                            cls = TN == 64 ? 0 : 1;
                        }
                        const int32_t slen = std::min( TN, rem );
                        gemm_real += 2. * TM * TN * nc * (double)( ( R - g0 + TM - 1 ) / TM );
                        for ( int32_t r = g0; r < R; r += TM ) {
                            pa_gemm_task t;
                            t.k = k; t.F = F; t.s0 = s; t.r0 = r; t.slen = slen;
                            t.flags = flags | ( side ? PA_TASK_USIDE : 0 );
                            t.pad0 = t.pad1 = 0;
                            gtasks[cls].push_back( t );
                        }
                        s += TN;
                    }
                }
                b = e + 1;
            }
        }
        for ( int32_t F : touched ) { srccount[F] = 0; lastsrc[F] = -1; }
        L.cblk_end  = lev_cblks.size();
        L.tinv_end  = tinv_tasks.size();
        L.trsm_mid  = trsm_tasks.size();
        trsm_tasks.insert( trsm_tasks.end(), ident_tasks.begin(), ident_tasks.end() );
        ident_tasks.clear();
        L.trsm_end  = trsm_tasks.size();
        L.trsm_inv  = lev_inv;
        L.dinv      = lev_dinv;
        any_dinv    = any_dinv || lev_dinv;
        if ( L.fac_end > L.cblk_beg && !lev_dinv ) all_dinv = false;
        for ( int c = 0; c < PA_NGC; c++ ) L.gc_end[c] = gtasks[c].size();
        /* look-ahead order: the tasks updating cblks of the next level
         * first (the next level's panels can start once they are done) */
        for ( int c = 0; c < PA_NGC; c++ ) {
            auto b0  = gtasks[c].begin() + L.gc_beg[c], b1 = gtasks[c].begin() + L.gc_end[c];
            auto mid = std::stable_partition( b0, b1, [&]( const pa_gemm_task &t ) { return level[t.F] == l + 1; } );
            L.gc_mid[c] = mid - gtasks[c].begin();
        }
        /* look-ahead: about PA_LA_SPLIT_FLOP of the remaining tasks (widest
         * class first, the launch order) cover the next level's diagonal
         * factorizations and tile inverses; the rest waits for those, so
         * that it is queued together with the next level's TRSM */
        {
            double acc = 0.;
            for ( int c = PA_NGC - 1; c >= 0; c-- ) {
                int64_t t = L.gc_mid[c];
                for ( ; t < L.gc_end[c] && acc < la_split; t++ ) {
                    const pa_gemm_task &g = gtasks[c][t];
                    acc += 2. * pa_gc_tm[c] * pa_gc_tn[c] * c_ncol[g.k];
                }
                L.gc_split[c] = t;
            }
        }
        L.solve_end = solve_tasks.size();
        L.bsolve_end = bsolve_tasks.size();
        L.dvt_end = dinv_tasks.size();
    }
    dev->lu   = two;
    dev->ldlt = ldlt;
    /* the diagonal class holds the narrow cblks' fused kernels at their
     * model flops; the GEMM tiles count as executed, padding included */
    dev->real_flops = dev->class_flops[0] + dev->class_flops[1] + gemm_real;
    dev->all_dinv = all_dinv && any_dinv;
    if ( ( rc = pa_upload( &dev->d_lev_cblks, lev_cblks ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_trsm_tasks, trsm_tasks ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_tinv_tasks, tinv_tasks ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_c_tinv, c_tinv ) ) ) return rc;
    dev->tinv_tiles = tinv_tiles;
    /* inverses in the widest precision, one region per TRSM mode */
    PA_HIP( hipMalloc( &dev->d_tinv, (size_t)std::max<int64_t>( tinv_tiles, 1 ) * 256 * sizeof( double ) * ( lu ? 2 : 1 ) ) );
    /* inverted diagonal blocks, one region per factor array */
    if ( any_dinv )
        PA_HIP( hipMalloc( &dev->d_dinv, (size_t)std::max<int64_t>( dev->dtnbr, 1 ) * sizeof( double ) * ( lu ? 2 : 1 ) ) );
    dev->S.c_tinv = dev->d_c_tinv;
    /* the task lists of all classes in one array, class after class */
    std::vector<pa_gemm_task> gemm_tasks;
    for ( int c = 0; c < PA_NGC; c++ ) {
        const int64_t base = gemm_tasks.size();
        for ( pa_level_t &L : dev->levels ) {
            L.gc_beg[c] += base;
            L.gc_mid[c] += base;
            L.gc_split[c] += base;
            L.gc_end[c] += base;
        }
        gemm_tasks.insert( gemm_tasks.end(), gtasks[c].begin(), gtasks[c].end() );
    }
    if ( ( rc = pa_upload( &dev->d_gemm_tasks, gemm_tasks ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_solve_tasks, solve_tasks ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_bsolve_tasks, bsolve_tasks ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_dinv_tasks, dinv_tasks ) ) ) return rc;

    /* the scatter-GEMM reads its operands by 32-bit byte offsets inside the
     * source cblk */
    for ( pastix_int_t k = 0; k < cblknbr; k++ )
        if ( (int64_t)c_stride[k] * c_ncol[k] * (int64_t)sizeof( double ) >= ( (int64_t)1 << 32 ) ) {
            fprintf( stderr, "pastix_amd: cblk %lld holds 4 GiB or more of FP64 coefficients; lower IPARM_MAX_BLOCKSIZE\n",
                     (long long)k );
            return PASTIX_ERR_BADPARAMETER;
        }
    /* LDS needs of the diagonal, TRSM and solve kernels */
    size_t need = std::max( pa_trsm_lds<double>( colmax ), (size_t)colmax * 16 * 8 * 2 ) + 4096;
    if ( need > (size_t)dev->lds_limit ) {
        fprintf( stderr, "pastix_amd: cblks of %d columns exceed the LDS budget (%d bytes); lower IPARM_MAX_BLOCKSIZE\n",
                 colmax, dev->lds_limit );
        return PASTIX_ERR_BADPARAMETER;
    }
    return PASTIX_SUCCESS;
}

/* Candidate schedules of the autotuner (pastix_amd_tune), the default first */
static const pa_tune_t pa_tune_cands[] = {
    { 1 << 30, 1, 1, 0 }, { 384, 1, 1, 0 }, { 256, 1, 1, 0 }, { 512, 1, 1, 0 }, { 1 << 30, 0, 1, 0 },
    { 1 << 30, 1, 0, 0 }, { 384, 1, 1, 1 }, { 256, 1, 1, 1 },
};
#define PA_NTUNE ( (int)( sizeof( pa_tune_cands ) / sizeof( pa_tune_cands[0] ) ) )

extern "C" int
pa_dev_tune_ncand( void )
{
    return PA_NTUNE;
}

/* 1 if candidate c gives a schedule of its own for this problem (the
 * 128-row tile threshold matters in FP32 only, the narrow-cblk kernels in
 * LL^t only) */
extern "C" int
pa_dev_tune_applicable( const pastix_data_t *pd, int c )
{
    if ( c < 0 || c >= PA_NTUNE ) return 0;
    const pa_tune_t &t = pa_tune_cands[c], &d = pa_tune_default;
    const int big = ( pd->iparm[IPARM_MIXED] != 0 ) || ( pd->iparm[IPARM_FLOAT] == SpmFloat );
    const int fa  = (int)pd->iparm[IPARM_FACTORIZATION];
    const int lu  = ( fa == PastixFactLU ) || pa_is_ldlt( fa ); /* no narrow-cblk kernels */
    if ( !big && !t.tall64 && t.tm_min != d.tm_min ) return 0; /* no 128-row tiles in FP64 without tall64 */
    if ( big && t.tall64 ) return 0;                            /* FP32 has its 128-row tiles anyway */
    if ( lu && t.small != d.small ) return 0;
    return 1;
}

/* set the schedule choices (candidate c, or params when c < 0) and rebuild
 * the task lists; the factors have to be recomputed afterwards */
extern "C" int
pa_dev_set_tune( pastix_data_t *pd, int c, const int32_t *params )
{
    pa_device_s *dev = DEV( pd );
    if ( !dev ) return PASTIX_ERR_BADPARAMETER;
    pa_tune_t t;
    if ( c >= 0 ) {
        if ( c >= PA_NTUNE ) return PASTIX_ERR_BADPARAMETER;
        t = pa_tune_cands[c];
    }
    else {
        if ( !params || params[0] < 64 || params[1] < 0 || params[1] > 1 || params[2] < 0 || params[2] > 1 )
            return PASTIX_ERR_BADPARAMETER;
        if ( params[3] < 0 || params[3] > 1 ) return PASTIX_ERR_BADPARAMETER;
        t.tm_min = params[0]; t.narrow = params[1]; t.small = params[2]; t.tall64 = params[3];
    }
    dev->tune = t;
    if ( dev->h_c_ncol.empty() ) return PASTIX_SUCCESS; /* applied when the structure is uploaded */
    PA_HIP( hipSetDevice( dev->device ) );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    return pa_build_schedule( pd );
}

extern "C" void
pa_dev_get_tune( const pastix_data_t *pd, int32_t *params )
{
    const pa_device_s *dev = DEV( pd );
    const pa_tune_t   &t   = dev ? dev->tune : pa_tune_default;
    params[0] = t.tm_min; params[1] = t.narrow; params[2] = t.small; params[3] = t.tall64;
}

extern "C" double
pa_dev_real_flops( const pastix_data_t *pd )
{
    return DEV( pd ) ? DEV( pd )->real_flops : 0.;
}

extern "C" pastix_int_t
pa_dev_nlevels( const pastix_data_t *pd )
{
    return DEV( pd ) ? (pastix_int_t)DEV( pd )->levels.size() : 0;
}

extern "C" int
pa_dev_upload_bcsc( pastix_data_t *pd )
{
    pa_device_s     *dev  = DEV( pd );
    const pa_bcsc_t *bcsc = pd->bcsc;
    int rc;
    PA_HIP( hipSetDevice( dev->device ) );
    pa_free_matrix( dev );
    std::vector<int64_t> colptr( bcsc->colptr, bcsc->colptr + bcsc->n + 1 );
    std::vector<int32_t> rowtab( bcsc->nnz );
    for ( pastix_int_t e = 0; e < bcsc->nnz; e++ ) rowtab[e] = (int32_t)bcsc->rowtab[e];
    if ( ( rc = pa_upload( &dev->d_colptr, colptr ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_rowtab, rowtab ) ) ) return rc;
    std::vector<double> Lv( bcsc->Lvalues, bcsc->Lvalues + bcsc->nnz );
    std::vector<double> Uv( bcsc->Uvalues, bcsc->Uvalues + bcsc->nnz );
    if ( ( rc = pa_upload( &dev->d_Lv, Lv ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_Uv, Uv ) ) ) return rc;
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_upload_valmap( pastix_data_t *pd, const pastix_int_t *Lptr, const pastix_int_t *Lsrc,
                      const pastix_int_t *Uptr, const pastix_int_t *Usrc, int64_t nnz )
{
    pa_device_s *dev = DEV( pd );
    int rc;
    PA_HIP( hipSetDevice( dev->device ) );
    hipFree( dev->d_vm_Lptr ); hipFree( dev->d_vm_Lsrc ); hipFree( dev->d_vm_Uptr ); hipFree( dev->d_vm_Usrc );
    dev->d_vm_Lptr = dev->d_vm_Lsrc = dev->d_vm_Uptr = dev->d_vm_Usrc = NULL;
    std::vector<int64_t> lp( Lptr, Lptr + nnz + 1 ), up( Uptr, Uptr + nnz + 1 );
    std::vector<int64_t> ls( Lsrc, Lsrc + Lptr[nnz] ), us( Usrc, Usrc + Uptr[nnz] );
    if ( ( rc = pa_upload( &dev->d_vm_Lptr, lp ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_vm_Uptr, up ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_vm_Lsrc, ls ) ) ) return rc;
    if ( ( rc = pa_upload( &dev->d_vm_Usrc, us ) ) ) return rc;
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_set_values( pastix_data_t *pd, const double *d_values, double *normF )
{
    pa_device_s *dev = DEV( pd );
    const int64_t nnz = pd->bcsc->nnz;
    PA_HIP( hipSetDevice( dev->device ) );
    if ( !dev->d_vm_Lptr || !dev->d_Lv ) return PASTIX_ERR_BADPARAMETER;
    if ( nnz > 0 ) {
        const unsigned grid = (unsigned)std::min<int64_t>( ( nnz + 255 ) / 256, 65536 );
        k_gather_sum<<<grid, 256, 0, dev->stream>>>( nnz, dev->d_vm_Lptr, dev->d_vm_Lsrc, d_values, dev->d_Lv );
        k_gather_sum<<<grid, 256, 0, dev->stream>>>( nnz, dev->d_vm_Uptr, dev->d_vm_Usrc, d_values, dev->d_Uv );
        PA_HIP( hipGetLastError() );
    }
    *normF = pa_dev_nrm2( pd, nnz, dev->d_Lv );
    return PASTIX_SUCCESS;
}

template <typename T>
static int
pa_coeftab_init_t( pa_device_s *dev, int lu, int32_t *overflow, bool fill )
{
    size_t bytes = (size_t)std::max<int64_t>( dev->coefnbr, 1 ) * sizeof( T );
    if ( !dev->d_L ) PA_HIP( hipMalloc( &dev->d_L, bytes ) );
    PA_HIP( hipMemsetAsync( dev->d_L, 0, bytes, dev->stream ) );
    if ( !lu && !dev->d_DT ) {
        /* transposed diagonal blocks of the LL^t solve: only their upper
         * triangles are ever written, the rest stays zero */
        size_t dbytes = (size_t)std::max<int64_t>( dev->dtnbr, 1 ) * sizeof( T );
        PA_HIP( hipMalloc( &dev->d_DT, dbytes ) );
        PA_HIP( hipMemsetAsync( dev->d_DT, 0, dbytes, dev->stream ) );
    }
    if ( lu ) {
        if ( !dev->d_U ) PA_HIP( hipMalloc( &dev->d_U, bytes ) );
        PA_HIP( hipMemsetAsync( dev->d_U, 0, bytes, dev->stream ) );
    }
    PA_HIP( hipMemsetAsync( dev->d_counters, 0, 16 * sizeof( int32_t ), dev->stream ) );
    int32_t n = (int32_t)dev->n;
    if ( n > 0 && fill ) {
        k_fill<T><<<( n + 255 ) / 256, 256, 0, dev->stream>>>(
            dev->S, n, dev->d_colptr, dev->d_rowtab, dev->d_Lv, dev->d_Uv, (T *)dev->d_L,
            lu ? (T *)dev->d_U : nullptr, dev->d_col2cblk, dev->d_counters + 2 );
        PA_HIP( hipGetLastError() );
    }
    int32_t h[16];
    PA_HIP( hipMemcpyAsync( h, dev->d_counters, sizeof( h ), hipMemcpyDeviceToHost, dev->stream ) );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    if ( overflow ) *overflow = h[2];
    return PASTIX_SUCCESS;
}

static int
pa_coeftab_setup( pastix_data_t *pd, int prec, int lu, int32_t *overflow, bool fill )
{
    pa_device_s *dev = DEV( pd );
    PA_HIP( hipSetDevice( dev->device ) );
    /* the task lists, the tile-inverse regions and the narrow-cblk kernels
     * depend on the factorization kind and the arithmetic: a change of
     * IPARM_FACTORIZATION or IPARM_MIXED after the analysis (read again at
     * each numerical factorization, as the reference does in
     * sopalin/pastix_task_sopalin.c:303,509) rebuilds them */
    if ( dev->sched_fact != (int)pd->iparm[IPARM_FACTORIZATION] || dev->sched_big != pa_sched_big( pd ) ) {
        int rc;
        PA_HIP( hipStreamSynchronize( dev->stream ) );
        if ( ( rc = pa_build_schedule( pd ) ) ) return rc;
    }
    if ( ( prec == PA_PREC_FLOAT ) != ( dev->sched_big != 0 ) ||
         lu != ( dev->sched_fact == PastixFactLU || pa_is_ldlt( dev->sched_fact ) ) ) {
        fprintf( stderr, "pastix_amd: coeftab arithmetic or factorization kind does not match the schedule\n" );
        return PASTIX_ERR_BADPARAMETER;
    }
    dev->fact_kind = dev->sched_fact;
    /* the HBM coeftab is kept across numerical factorizations of the same
     * structure and arithmetic; it is only reallocated when they change */
    if ( dev->prec != prec || dev->lu != lu || ( lu && !dev->d_U ) ) pa_free_factors( dev );
    dev->prec = prec;
    dev->lu   = lu;
    if ( prec == PA_PREC_FLOAT ) return pa_coeftab_init_t<float>( dev, lu, overflow, fill );
    return pa_coeftab_init_t<double>( dev, lu, overflow, fill );
}

extern "C" int
pa_dev_coeftab_init( pastix_data_t *pd, int prec, int lu, int32_t *overflow )
{
    return pa_coeftab_setup( pd, prec, lu, overflow, true );
}

/* ------------------------------------------------------------------------- */
/* factor arrays to and from host memory (checkpoints)                       */
/* ------------------------------------------------------------------------- */
/* the factors held in HBM: arithmetic, two arrays, factorization kind and
 * coefficients per array; PASTIX_ERR_BADPARAMETER when there are none */
extern "C" int
pa_dev_factor_info( const pastix_data_t *pd, int *prec, int *two, int *fact, int64_t *coefnbr )
{
    const pa_device_s *dev = DEV( pd );
    if ( !dev || !dev->d_L || dev->fact_kind < 0 ) return PASTIX_ERR_BADPARAMETER;
    *prec    = dev->prec;
    *two     = dev->lu;
    *fact    = dev->fact_kind;
    *coefnbr = dev->coefnbr;
    return PASTIX_SUCCESS;
}

/* copy count coefficients from offset off of factor array `side` (0: L,
 * 1: the second array) to host memory, in the factors' arithmetic */
extern "C" int
pa_dev_factor_get( const pastix_data_t *pd, int side, int64_t off, int64_t count, void *host )
{
    pa_device_s *dev = DEV( pd );
    const void  *src = side ? dev->d_U : dev->d_L;
    const size_t es  = dev->prec == PA_PREC_DOUBLE ? sizeof( double ) : sizeof( float );
    if ( !src || off < 0 || count < 0 || off + count > dev->coefnbr ) return PASTIX_ERR_BADPARAMETER;
    PA_HIP( hipSetDevice( dev->device ) );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    PA_HIP( hipMemcpy( host, (const char *)src + off * es, count * es, hipMemcpyDeviceToHost ) );
    return PASTIX_SUCCESS;
}

/* allocate (or keep) the coeftab for factors of kind IPARM_FACTORIZATION in
 * arithmetic prec that will be copied in with pa_dev_factor_put, without
 * filling it from the matrix */
extern "C" int
pa_dev_factor_begin( pastix_data_t *pd, int prec, int lu )
{
    return pa_coeftab_setup( pd, prec, lu, nullptr, false );
}

extern "C" int
pa_dev_factor_put( pastix_data_t *pd, int side, int64_t off, int64_t count, const void *host )
{
    pa_device_s *dev = DEV( pd );
    void        *dst = side ? dev->d_U : dev->d_L;
    const size_t es  = dev->prec == PA_PREC_DOUBLE ? sizeof( double ) : sizeof( float );
    if ( !dst || off < 0 || count < 0 || off + count > dev->coefnbr ) return PASTIX_ERR_BADPARAMETER;
    PA_HIP( hipSetDevice( dev->device ) );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    PA_HIP( hipMemcpy( (char *)dst + off * es, host, count * es, hipMemcpyHostToDevice ) );
    return PASTIX_SUCCESS;
}

/* after the factor arrays are in place: the transposed diagonal blocks of
 * the LL^t solve, as at the end of a factorization */
template <typename T> static int pa_launch_dinv( pa_device_s *dev, const pa_level_t &lev, hipStream_t st );

extern "C" int
pa_dev_factor_end( pastix_data_t *pd )
{
    pa_device_s *dev = DEV( pd );
    PA_HIP( hipSetDevice( dev->device ) );
    /* the inverted diagonal blocks of the solve, as a factorization leaves them */
    for ( const pa_level_t &lev : dev->levels ) {
        int rc = ( dev->prec == PA_PREC_DOUBLE ) ? pa_launch_dinv<double>( dev, lev, dev->stream )
                                                 : pa_launch_dinv<float>( dev, lev, dev->stream );
        if ( rc ) return rc;
    }
    if ( !dev->lu && dev->S.cblknbr > 0 ) {
        if ( dev->prec == PA_PREC_DOUBLE )
            k_diag_transpose<double><<<dev->S.cblknbr, 256, 0, dev->stream>>>( dev->S, (const double *)dev->d_L,
                                                                               (double *)dev->d_DT );
        else
            k_diag_transpose<float><<<dev->S.cblknbr, 256, 0, dev->stream>>>( dev->S, (const float *)dev->d_L,
                                                                              (float *)dev->d_DT );
        PA_HIP( hipGetLastError() );
    }
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* factorization                                                             */
/* ------------------------------------------------------------------------- */
#define PA_PROF_MARK( slot )                                                                      \
    do {                                                                                          \
        if ( dev->profiling ) PA_HIP( hipEventRecord( dev->prof_ev[6 * li + ( slot )], dev->stream ) ); \
    } while ( 0 )

/* the narrow cblks of a level, one launch per row-slot count */
template <typename T>
static void
pa_small_llt_launch( pa_device_s *dev, const pa_level_t &lev, T *L, const double *crit, int32_t *c0, int32_t *c1 )
{
    int64_t beg = lev.fac_end;
    /* cblks of at most 32 rows: two per wavefront (eight per workgroup) */
    const int32_t np = (int32_t)( lev.sm_pair_end - beg );
    if ( np > 0 ) {
        const int32_t *cb = dev->d_lev_cblks + beg;
        k_small_llt<T, 1, true><<<(unsigned)( ( np + 7 ) / 8 ), 256, 0, dev->stream>>>( dev->S, cb, np, L, crit, c0, c1, c0 + 3 );
        beg = lev.sm_pair_end;
    }
    for ( int ns = 1; ns <= 3; ns++ ) {
        const int32_t n = (int32_t)( lev.sm_end[ns - 1] - beg );
        const int32_t *cb = dev->d_lev_cblks + beg;
        const unsigned g  = (unsigned)( ( n + 3 ) / 4 );
        if ( n > 0 ) {
            if ( ns == 1 )      k_small_llt<T, 1><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, L, crit, c0, c1, c0 + 3 );
            else if ( ns == 2 ) k_small_llt<T, 2><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, L, crit, c0, c1, c0 + 3 );
            else                k_small_llt<T, 3><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, L, crit, c0, c1, c0 + 3 );
        }
        beg = lev.sm_end[ns - 1];
    }
}

template <typename T, bool BWD>
static void
pa_small_solve_launch( pa_device_s *dev, const pa_level_t &lev, const T *C, T *x, int64_t ld, T *xo, int64_t ldo,
                       int32_t nrhs )
{
    int64_t beg = lev.fac_end;
    /* cblks of at most 32 rows: two per wavefront (eight per workgroup) */
    const int32_t np = (int32_t)( lev.sm_pair_end - beg );
    if ( np > 0 ) {
        const int32_t *cb = dev->d_lev_cblks + beg;
        const unsigned g  = (unsigned)( ( np + 7 ) / 8 );
        if ( xo == x ) k_small_solve<T, BWD, 1, true, true><<<g, 256, 0, dev->stream>>>( dev->S, cb, np, C, x, nrhs, ld, xo, ldo );
        else           k_small_solve<T, BWD, 1, false, true><<<g, 256, 0, dev->stream>>>( dev->S, cb, np, C, x, nrhs, ld, xo, ldo );
        beg = lev.sm_pair_end;
    }
    for ( int ns = 1; ns <= 3; ns++ ) {
        const int32_t n = (int32_t)( lev.sm_end[ns - 1] - beg );
        const int32_t *cb = dev->d_lev_cblks + beg;
        const unsigned g  = (unsigned)( ( n + 3 ) / 4 );
        if ( n > 0 && xo == x ) {
            if ( ns == 1 )      k_small_solve<T, BWD, 1, true><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, C, x, nrhs, ld, xo, ldo );
            else if ( ns == 2 ) k_small_solve<T, BWD, 2, true><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, C, x, nrhs, ld, xo, ldo );
            else                k_small_solve<T, BWD, 3, true><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, C, x, nrhs, ld, xo, ldo );
        }
        else if ( n > 0 ) {
            if ( ns == 1 )      k_small_solve<T, BWD, 1, false><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, C, x, nrhs, ld, xo, ldo );
            else if ( ns == 2 ) k_small_solve<T, BWD, 2, false><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, C, x, nrhs, ld, xo, ldo );
            else                k_small_solve<T, BWD, 3, false><<<g, 256, 0, dev->stream>>>( dev->S, cb, n, C, x, nrhs, ld, xo, ldo );
        }
        beg = lev.sm_end[ns - 1];
    }
}

/* Diagonal factorization, tile inverses and panel TRSM of the general
 * cblks of one level, on stream st */
template <typename T>
static int
pa_launch_panel( pa_device_s *dev, const pa_level_t &lev, hipStream_t st )
{
    const double *crit = dev->d_crit;
    T *L = (T *)dev->d_L, *U = (T *)dev->d_U;
    const int lu = dev->lu;
    const int32_t ncb = (int32_t)( lev.fac_end - lev.cblk_beg );
    if ( ncb == 0 ) return PASTIX_SUCCESS;
    /* the panel P, and for LU also the row panel U12 */
    size_t lds_diag = (size_t)pa_diag_nb<T>::value * lev.ncmax_fac * sizeof( T ) * ( lu ? 2 : 1 );
    const int32_t *cb = dev->d_lev_cblks + lev.cblk_beg;
    int32_t *c0 = dev->d_counters, *c1 = dev->d_counters + 1;
    const bool regdiag = !lu && sizeof( T ) == 4 && lev.ncmax_fac > 64 && lev.ncmax_fac <= 16 * PA_DREG_NT;
    if ( regdiag ) {
        k_diag_potrf_reg<<<ncb, 1024, 0, st>>>( dev->S, cb, (float *)L, crit, c0, c1 );
    }
    else if ( lev.ncmax_fac > 64 ) {
        /* 1024 threads fit without spilling (FP64 with 2 trailing tiles
         * per wave and round): the waves of the trailing updates; LDL^t
         * runs the LU kernel on the block symmetrised from its lower
         * triangle */
        if ( dev->ldlt ) k_diag<T, true, 1024, true><<<ncb, 1024, lds_diag, st>>>( dev->S, cb, L, U, crit, c0, c1 );
        else if ( lu )   k_diag<T, true, 1024><<<ncb, 1024, lds_diag, st>>>( dev->S, cb, L, U, crit, c0, c1 );
        else             k_diag<T, false, 1024><<<ncb, 1024, lds_diag, st>>>( dev->S, cb, L, U, crit, c0, c1 );
    }
    else {
        if ( dev->ldlt ) k_diag<T, true, 64, true><<<ncb, 64, lds_diag, st>>>( dev->S, cb, L, U, crit, c0, c1 );
        else if ( lu )   k_diag<T, true, 64><<<ncb, 64, lds_diag, st>>>( dev->S, cb, L, U, crit, c0, c1 );
        else             k_diag<T, false, 64><<<ncb, 64, lds_diag, st>>>( dev->S, cb, L, U, crit, c0, c1 );
    }
    PA_HIP( hipGetLastError() );
    return PASTIX_SUCCESS;
}

/* k_trsm_inv with the accumulators sized for the level's widest cblk */
template <typename T, int MODE>
static void
pa_trsm_inv_launch( int32_t ncmax, dim3 grid, hipStream_t st, const pa_dstruct &S, const int2 *tr, T *L, T *U,
                    const T *Ti, int64_t ti2, T *Di, int64_t di2 )
{
    const int nj = ( ncmax + 63 ) / 64;
    if ( nj <= 1 )      k_trsm_inv<T, MODE, 1><<<grid, 256, 0, st>>>( S, tr, L, U, Ti, ti2, Di, di2 );
    else if ( nj <= 2 ) k_trsm_inv<T, MODE, 2><<<grid, 256, 0, st>>>( S, tr, L, U, Ti, ti2, Di, di2 );
    else if ( nj <= 4 ) k_trsm_inv<T, MODE, 4><<<grid, 256, 0, st>>>( S, tr, L, U, Ti, ti2, Di, di2 );
    else                k_trsm_inv<T, MODE, PA_TI_NJ><<<grid, 256, 0, st>>>( S, tr, L, U, Ti, ti2, Di, di2 );
}

/* parts of a level's TRSM launch sequence: the tile inverses, the TRSM */
enum { PA_TRSM_TINV = 1, PA_TRSM_SLABS = 2, PA_TRSM_ALL = 3 };

template <typename T>
static int
pa_launch_trsm( pa_device_s *dev, const pa_level_t &lev, hipStream_t st, int parts = PA_TRSM_ALL )
{
    T *L = (T *)dev->d_L, *U = (T *)dev->d_U;
    const int lu = dev->lu;
    int64_t ntr = lev.trsm_end - lev.trsm_beg;
    if ( ntr > 0 && lev.trsm_inv ) {
        const int32_t ntl = (int32_t)( lev.tinv_end - lev.tinv_beg );
        const unsigned gti = (unsigned)( ( ntl + 3 ) / 4 );
        const int2   *tt   = dev->d_tinv_tasks + lev.tinv_beg;
        const int2   *tr   = dev->d_trsm_tasks + lev.trsm_beg;
        T            *Ti   = (T *)dev->d_tinv;
        const int64_t ti2  = dev->tinv_tiles * 256; /* LU: the unit-L^T inverses follow the U ones */
        T            *Di   = (T *)dev->d_dinv;      /* inverted diagonal blocks (identity tasks) */
        const int64_t di2  = dev->dtnbr;             /* LU: the U^T array's inverses */
        const bool tinv = ( parts & PA_TRSM_TINV ) != 0, slabs = ( parts & PA_TRSM_SLABS ) != 0;
        if ( dev->ldlt ) {
            /* L panel against D L^t, with its copy L D into the second array */
            if ( tinv ) k_tile_inv<T, 4><<<gti, 256, 0, st>>>( dev->S, tt, ntl, L, U, Ti, 0 );
            if ( slabs ) pa_trsm_inv_launch<T, 4>( lev.ncmax_fac, dim3( (unsigned)ntr ), st, dev->S, tr, L, U, Ti, 0, Di, di2 );
        }
        else if ( lu ) {
            /* both LU modes (L and U^T panels) in one launch each */
            if ( tinv ) k_tile_inv<T, 3><<<dim3( gti, 2 ), 256, 0, st>>>( dev->S, tt, ntl, L, U, Ti, ti2 );
            if ( slabs ) pa_trsm_inv_launch<T, 3>( lev.ncmax_fac, dim3( (unsigned)ntr, 2 ), st, dev->S, tr, L, U, Ti, ti2, Di, di2 );
        }
        else {
            if ( tinv ) k_tile_inv<T, 0><<<gti, 256, 0, st>>>( dev->S, tt, ntl, L, U, Ti, 0 );
            if ( slabs ) pa_trsm_inv_launch<T, 0>( lev.ncmax_fac, dim3( (unsigned)ntr ), st, dev->S, tr, L, U, Ti, 0, Di, di2 );
        }
        PA_HIP( hipGetLastError() );
    }
    else if ( ntr > 0 && ( parts & PA_TRSM_SLABS ) ) {
        size_t lds_trsm = pa_trsm_lds<T>( lev.ncmax );
        if ( dev->ldlt ) {
            k_trsm<T, 1><<<(unsigned)ntr, 256, lds_trsm, st>>>( dev->S, dev->d_trsm_tasks + lev.trsm_beg, L, U );
            k_ldlt_scalo<T><<<(unsigned)ntr, 256, 0, st>>>( dev->S, dev->d_trsm_tasks + lev.trsm_beg, L, U );
        }
        else if ( lu ) {
            k_trsm<T, 1><<<(unsigned)ntr, 256, lds_trsm, st>>>( dev->S, dev->d_trsm_tasks + lev.trsm_beg, L, U );
            k_trsm<T, 2><<<(unsigned)ntr, 256, lds_trsm, st>>>( dev->S, dev->d_trsm_tasks + lev.trsm_beg, L, U );
        }
        else {
            k_trsm<T, 0><<<(unsigned)ntr, 256, lds_trsm, st>>>( dev->S, dev->d_trsm_tasks + lev.trsm_beg, L, U );
        }
        PA_HIP( hipGetLastError() );
    }
    return PASTIX_SUCCESS;
}

/* The tile inverses and the diagonal-block inverses of one level alone (the
 * identity slabs of its TRSM task list), for factors that did not come from
 * a factorization on this instance */
template <typename T>
static int
pa_launch_dinv( pa_device_s *dev, const pa_level_t &lev, hipStream_t st )
{
    T *L = (T *)dev->d_L, *U = (T *)dev->d_U;
    const int64_t nid = lev.trsm_end - lev.trsm_mid;
    if ( !lev.dinv || nid <= 0 ) return PASTIX_SUCCESS;
    const int32_t ntl = (int32_t)( lev.tinv_end - lev.tinv_beg );
    const unsigned gti = (unsigned)( ( ntl + 3 ) / 4 );
    const int2   *tt   = dev->d_tinv_tasks + lev.tinv_beg;
    const int2   *tr   = dev->d_trsm_tasks + lev.trsm_mid;
    T            *Ti   = (T *)dev->d_tinv;
    const int64_t ti2  = dev->tinv_tiles * 256;
    T            *Di   = (T *)dev->d_dinv;
    const int64_t di2  = dev->dtnbr;
    if ( dev->lu ) {
        k_tile_inv<T, 3><<<dim3( gti, 2 ), 256, 0, st>>>( dev->S, tt, ntl, L, U, Ti, ti2 );
        pa_trsm_inv_launch<T, 3>( lev.ncmax_fac, dim3( (unsigned)nid, 2 ), st, dev->S, tr, L, U, Ti, ti2, Di, di2 );
    }
    else {
        k_tile_inv<T, 0><<<gti, 256, 0, st>>>( dev->S, tt, ntl, L, U, Ti, 0 );
        pa_trsm_inv_launch<T, 0>( lev.ncmax_fac, dim3( (unsigned)nid ), st, dev->S, tr, L, U, Ti, 0, Di, di2 );
    }
    PA_HIP( hipGetLastError() );
    return PASTIX_SUCCESS;
}

/* the scatter-GEMM tasks [from[c], to[c]) of every tile class, on the
 * library stream, the widest tiles first (they hold most of the work) */
template <typename T>
static int
pa_launch_gemm( pa_device_s *dev, const int64_t *from, const int64_t *to )
{
    T *L = (T *)dev->d_L, *U = (T *)dev->d_U;
    T *Up = dev->lu ? U : nullptr;
    int32_t *nerr = dev->d_counters + 3;
    for ( int c = PA_NGC - 1; c >= 0; c-- ) {
        const int64_t ng = to[c] - from[c];
        if ( ng <= 0 ) continue;
        const pa_gemm_task *tk = dev->d_gemm_tasks + from[c];
        const unsigned g = (unsigned)ng;
        switch ( c ) {
        case 0: k_gemm_level<T, 64, 64><<<g, 256, 0, dev->stream>>>( dev->S, tk, L, Up, nerr ); break;
        case 1: k_gemm_level<T, 64, 32><<<g, 256, 0, dev->stream>>>( dev->S, tk, L, Up, nerr ); break;
        case 3: k_gemm_level<T, 128, 64><<<g, 256, 0, dev->stream>>>( dev->S, tk, L, Up, nerr ); break;
        case 4: k_gemm_level<T, 128, 32><<<g, 256, 0, dev->stream>>>( dev->S, tk, L, Up, nerr ); break;
        default:
            if constexpr ( sizeof( T ) == 4 ) {
                k_gemm_level<T, 128, 128><<<g, 256, 0, dev->stream>>>( dev->S, tk, L, Up, nerr );
            }
            else {
                return PASTIX_ERR_INTERNAL; /* 128 x 128 tiles are FP32 only */
            }
        }
        PA_HIP( hipGetLastError() );
    }
    return PASTIX_SUCCESS;
}

/*
 * The level-ordered launch sequence.  Per level: the narrow cblks' fused
 * kernel, the diagonal factorizations, the tile inverses and TRSMs, and the
 * scatter-GEMMs.  With look-ahead (the default, off while profiling the
 * kernel classes), the GEMM tasks of level l whose target cblk belongs to
 * level l + 1 run first; then the diagonal factorizations and TRSMs of
 * level l + 1 start on the second stream, concurrently with the rest of
 * level l's GEMM, whose targets lie above level l + 1 and are disjoint from
 * what they read and write; level l + 1's GEMMs wait for both.  Every other
 * input of level l + 1's panels comes from updates ordered before the fork
 * on the library stream.
 */
template <typename T>
static int
pa_factorize_t( pa_device_s *dev )
{
    const double *crit = dev->d_crit;
    T *L = (T *)dev->d_L;
    const size_t nl = dev->levels.size();
    const bool   la = dev->lookahead && !dev->profiling;
    int rc;
    if ( la && dev->la_ev.size() < 3 * nl ) {
        size_t old = dev->la_ev.size();
        dev->la_ev.resize( 3 * nl );
        for ( size_t i = old; i < dev->la_ev.size(); i++ )
            PA_HIP( hipEventCreateWithFlags( &dev->la_ev[i], hipEventDisableTiming ) );
    }
    bool ahead = false; /* the current level's panels were launched ahead */
    for ( size_t li = 0; li < nl; li++ ) {
        const pa_level_t &lev = dev->levels[li];
        int32_t ncb = (int32_t)( lev.fac_end - lev.cblk_beg ), nsm = (int32_t)( lev.cblk_end - lev.fac_end );
        if ( ncb == 0 && nsm == 0 ) continue;
        PA_PROF_MARK( 0 );
        int32_t *c0 = dev->d_counters, *c1 = dev->d_counters + 1;
        if ( nsm > 0 ) {
            pa_small_llt_launch<T>( dev, lev, L, crit, c0, c1 );
            PA_HIP( hipGetLastError() );
        }
        if ( ahead ) {
            /* join: this level's panels were factored on the second stream */
            PA_HIP( hipStreamWaitEvent( dev->stream, dev->la_ev[2 * li + 1], 0 ) );
        }
        else {
            if ( ( rc = pa_launch_panel<T>( dev, lev, dev->stream ) ) ) return rc;
        }
        PA_PROF_MARK( 1 );
        PA_PROF_MARK( 2 );
        if ( !ahead && ( rc = pa_launch_trsm<T>( dev, lev, dev->stream ) ) ) return rc;
        PA_PROF_MARK( 3 );
        PA_PROF_MARK( 4 );
        /* the next level with general cblks, if it directly follows */
        const bool fork = la && li + 1 < nl && dev->levels[li + 1].fac_end > dev->levels[li + 1].cblk_beg;
        if ( fork ) {
            if ( ( rc = pa_launch_gemm<T>( dev, lev.gc_beg, lev.gc_mid ) ) ) return rc;
            const pa_level_t &nx = dev->levels[li + 1];
            PA_HIP( hipEventRecord( dev->la_ev[2 * ( li + 1 )], dev->stream ) );
            PA_HIP( hipStreamWaitEvent( dev->stream2, dev->la_ev[2 * ( li + 1 )], 0 ) );
            if ( ( rc = pa_launch_panel<T>( dev, nx, dev->stream2 ) ) ) return rc;
            if ( ( rc = pa_launch_trsm<T>( dev, nx, dev->stream2, PA_TRSM_TINV ) ) ) return rc;
            /* the rest of this level's GEMM in two parts: the second waits
             * until the next level's TRSM is ready to start, so that the
             * TRSM's workgroups are queued with it instead of after a full
             * GPU of GEMM tiles (its last slabs would otherwise run alone
             * after this level's GEMM, holding up the next one) */
            PA_HIP( hipEventRecord( dev->la_ev[2 * nl + li + 1], dev->stream2 ) );
            if ( ( rc = pa_launch_trsm<T>( dev, nx, dev->stream2, PA_TRSM_SLABS ) ) ) return rc;
            PA_HIP( hipEventRecord( dev->la_ev[2 * ( li + 1 ) + 1], dev->stream2 ) );
            if ( ( rc = pa_launch_gemm<T>( dev, lev.gc_mid, lev.gc_split ) ) ) return rc;
            PA_HIP( hipStreamWaitEvent( dev->stream, dev->la_ev[2 * nl + li + 1], 0 ) );
            if ( ( rc = pa_launch_gemm<T>( dev, lev.gc_split, lev.gc_end ) ) ) return rc;
        }
        else {
            if ( ( rc = pa_launch_gemm<T>( dev, lev.gc_beg, lev.gc_end ) ) ) return rc;
        }
        ahead = fork;
        PA_PROF_MARK( 5 );
    }
    if ( ahead ) return PASTIX_ERR_INTERNAL; /* a fork past the last level: not possible */
    if ( dev->lu == 0 && dev->S.cblknbr > 0 && !dev->all_dinv ) {
        /* transposed diagonal blocks for the substitution backward sweeps
         * (with inverted diagonal blocks on every level, a solve that needs
         * them makes them itself) */
        k_diag_transpose<T><<<dev->S.cblknbr, 256, 0, dev->stream>>>( dev->S, L, (T *)dev->d_DT );
        PA_HIP( hipGetLastError() );
    }
    return PASTIX_SUCCESS;
}

static inline int64_t lu_launches( const pa_device_s *dev ) { return dev->lu ? 2 : 1; }

/* accumulate the per-class kernel times of the last factorization */
static int
pa_collect_profile( pa_device_s *dev )
{
    for ( size_t li = 0; li < dev->levels.size(); li++ ) {
        const pa_level_t &lev = dev->levels[li];
        if ( lev.cblk_end == lev.cblk_beg ) continue;
        for ( int c = 0; c < 3; c++ ) {
            float ms = 0.f;
            PA_HIP( hipEventElapsedTime( &ms, dev->prof_ev[6 * li + 2 * c], dev->prof_ev[6 * li + 2 * c + 1] ) );
            dev->prof_ms[c] += ms;
        }
        if ( getenv( "PASTIX_AMD_LEVEL_REPORT" ) ) {
            /* per-level timing report on stderr: the critical path by level */
            float t[3];
            for ( int c = 0; c < 3; c++ )
                PA_HIP( hipEventElapsedTime( &t[c], dev->prof_ev[6 * li + 2 * c], dev->prof_ev[6 * li + 2 * c + 1] ) );
            fprintf( stderr, "level %3zu cblks %6d ncmax %4d trsm %7d gemm %7d/%7d/%7d/%7d/%7d | diag %8.3f trsm %8.3f gemm %8.3f ms"
                             " | gemm %6.1f TF trsm %6.1f TF\n",
                     li, (int)( lev.cblk_end - lev.cblk_beg ), lev.ncmax, (int)( lev.trsm_end - lev.trsm_beg ),
                     (int)( lev.gc_end[0] - lev.gc_beg[0] ), (int)( lev.gc_end[1] - lev.gc_beg[1] ),
                     (int)( lev.gc_end[2] - lev.gc_beg[2] ), (int)( lev.gc_end[3] - lev.gc_beg[3] ),
                     (int)( lev.gc_end[4] - lev.gc_beg[4] ), t[0], t[1], t[2],
                     lev.flops[2] / ( t[2] * 1e9 ), lev.flops[1] / ( t[1] * 1e9 ) );
        }
        dev->prof_launch[0] += 1;
        /* inverted-tile TRSM: k_tile_inv + k_trsm_inv (both sides in one
         * launch each); substitution TRSM: one launch per side */
        dev->prof_launch[1] += ( lev.trsm_end > lev.trsm_beg ) ? ( lev.trsm_inv ? 2 : lu_launches( dev ) ) : 0;
        for ( int c = 0; c < PA_NGC; c++ ) dev->prof_launch[2] += ( lev.gc_end[c] > lev.gc_beg[c] ) ? 1 : 0;
    }
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_factorize( pastix_data_t *pd, double crit, int32_t *npiv, int32_t *nneg, double *seconds )
{
    pa_device_s *dev = DEV( pd );
    int rc;
    PA_HIP( hipSetDevice( dev->device ) );
    /* pivots, negative terms, (fill overflow), dropped updates */
    PA_HIP( hipMemsetAsync( dev->d_counters, 0, 4 * sizeof( int32_t ), dev->stream ) );
    k_set_f64<<<1, 64, 0, dev->stream>>>( dev->d_crit, crit );
    PA_HIP( hipGetLastError() );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    auto t0 = std::chrono::steady_clock::now();
    PA_HIP( hipEventRecord( dev->ev0, dev->stream ) );
    auto launch = [&]() {
        return ( dev->prec == PA_PREC_FLOAT ) ? pa_factorize_t<float>( dev ) : pa_factorize_t<double>( dev );
    };
    if ( dev->profiling && dev->prof_ev.size() < 6 * dev->levels.size() ) {
        /* the per-level timing events exist before any capture */
        size_t old = dev->prof_ev.size();
        dev->prof_ev.resize( 6 * dev->levels.size() );
        for ( size_t i = old; i < dev->prof_ev.size(); i++ ) PA_HIP( hipEventCreate( &dev->prof_ev[i] ) );
    }
    if ( dev->use_graphs && !dev->profiling ) {
        /* the launch sequence depends on the factors' addresses only (the
         * criterion is read from d_crit); the per-level timing events are
         * recorded by the eager launches only (captured events cannot be
         * timed) */
        pa_graph_entry &g = dev->fact_graph;
        bool hit = g.exec && g.key[0] == dev->d_L && g.key[1] == dev->d_U && g.key[2] == dev->d_DT &&
                   g.ikey[0] == dev->prec && g.ikey[1] == dev->lu && g.ikey[2] == dev->profiling;
        rc = pa_graph_run( dev, g, hit, launch );
        if ( rc ) return rc;
        g.key[0] = dev->d_L; g.key[1] = dev->d_U; g.key[2] = dev->d_DT;
        g.ikey[0] = dev->prec; g.ikey[1] = dev->lu; g.ikey[2] = dev->profiling;
    }
    else {
        rc = launch();
    }
    if ( rc ) return rc;
    PA_HIP( hipEventRecord( dev->ev1, dev->stream ) );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    auto t1 = std::chrono::steady_clock::now();
    if ( dev->profiling ) {
        if ( ( rc = pa_collect_profile( dev ) ) ) return rc;
    }
    float ms = 0.f;
    hipEventElapsedTime( &ms, dev->ev0, dev->ev1 );
    dev->last_fact_event_s = ms * 1e-3;
    if ( seconds ) *seconds = std::chrono::duration<double>( t1 - t0 ).count();
    int32_t h[4];
    PA_HIP( hipMemcpy( h, dev->d_counters, sizeof( h ), hipMemcpyDeviceToHost ) );
    if ( npiv ) *npiv = h[0];
    if ( nneg ) *nneg = h[1];
    if ( h[3] != 0 ) {
        fprintf( stderr, "pastix_amd: %d update rows/columns fall outside the block structure of their target "
                         "(inconsistent symbolic factorization)\n", h[3] );
        return PASTIX_ERR_INTERNAL;
    }
    return PASTIX_SUCCESS;
}

extern "C" double
pa_dev_last_fact_event_seconds( const pastix_data_t *pd )
{
    return DEV( pd ) ? DEV( pd )->last_fact_event_s : 0.;
}

/* ------------------------------------------------------------------------- */
/* solve                                                                     */
/* ------------------------------------------------------------------------- */
/* largest cblk width handled by the register-prefetching diagonal solves */
#define PA_SD_MAXCOL 512

template <typename T, bool BWD, bool UNIT, int TH>
static void
pa_solve_diag_launch( pa_device_s *dev, dim3 grid, size_t lds, const int32_t *cb, const T *C, const T *CT,
                      const int64_t *ct_off, const T *x, int64_t ld, T *xo, int64_t ldo, int32_t nrhs, int32_t rchunk )
{
    if ( BWD ) k_solve_diag_bwd<T, UNIT, TH><<<grid, TH, lds, dev->stream>>>( dev->S, cb, CT, ct_off, x, nrhs, rchunk, ld, xo, ldo );
    else       k_solve_diag_fwd<T, UNIT, TH><<<grid, TH, lds, dev->stream>>>( dev->S, cb, C, x, nrhs, rchunk, ld, xo, ldo );
}

/* diagonal solves of one level, from x (the right-hand sides with every
 * update of the other cblks applied) to xo: the products with the inverted
 * diagonal blocks on dinv levels (64-entry blocks, 256 or 512 threads; in
 * place, xo == x, one workgroup per cblk), else the register-prefetching
 * kernels (workgroup of the level's widest cblk rounded up to 64/128/256/512
 * threads) for cblks of up to PA_SD_MAXCOL columns, the LDS-blocked kernel
 * above that */
template <typename T, bool BWD, bool UNIT>
static int
pa_solve_diag_level( pa_device_s *dev, const pa_level_t &lev, const T *C, const T *CT, const int64_t *ct_off,
                     const T *x, int64_t ld, T *xo, int64_t ldo, int32_t nrhs, const T *Dinv )
{
    /* the general cblks of the level; the narrow ones go to k_small_solve */
    const int32_t ncb = (int32_t)( lev.fac_end - lev.cblk_beg );
    const int32_t *cb = dev->d_lev_cblks + lev.cblk_beg;
    const int32_t ncmax = lev.ncmax_fac;
    if ( ncb == 0 ) return PASTIX_SUCCESS;
    if ( Dinv && lev.dinv && xo == x ) {
        const size_t  avail  = (size_t)dev->lds_limit - 4096;
        const int32_t rchunk = (int32_t)std::max<size_t>( 1, std::min<size_t>( nrhs, avail / ( 2 * (size_t)ncmax * sizeof( T ) ) ) );
        dim3          grid( ncb, ( nrhs + rchunk - 1 ) / rchunk );
        size_t        lds = 2 * (size_t)ncmax * rchunk * sizeof( T );
        T            *xi  = xo;
        if ( ncmax <= 64 )       k_solve_dinv_cblk<T, BWD, 64><<<grid, 64, lds, dev->stream>>>( dev->S, cb, Dinv, xi, nrhs, rchunk, ld );
        else if ( ncmax <= 128 ) k_solve_dinv_cblk<T, BWD, 128><<<grid, 128, lds, dev->stream>>>( dev->S, cb, Dinv, xi, nrhs, rchunk, ld );
        else if ( ncmax <= 256 ) k_solve_dinv_cblk<T, BWD, 256><<<grid, 256, lds, dev->stream>>>( dev->S, cb, Dinv, xi, nrhs, rchunk, ld );
        else                     k_solve_dinv_cblk<T, BWD, 512><<<grid, 512, lds, dev->stream>>>( dev->S, cb, Dinv, xi, nrhs, rchunk, ld );
    }
    else if ( Dinv && lev.dinv ) {
        /* products with the inverted diagonal blocks */
        const int     th     = ncmax <= 128 ? 256 : 512;
        const size_t  per    = ( (size_t)ncmax + ( th / 64 ) * 64 ) * sizeof( T );
        const size_t  avail  = (size_t)dev->lds_limit - 4096;
        const int32_t rchunk = (int32_t)std::max<size_t>( 1, std::min<size_t>( nrhs, avail / per ) );
        const int64_t nt     = lev.dvt_end - lev.dvt_beg;
        dim3          grid( (unsigned)nt, ( nrhs + rchunk - 1 ) / rchunk );
        const size_t  lds = per * rchunk;
        const int2   *tk  = dev->d_dinv_tasks + lev.dvt_beg;
        if ( nt <= 0 ) return PASTIX_SUCCESS;
        if ( th == 256 ) k_solve_dinv<T, BWD, 256><<<grid, 256, lds, dev->stream>>>( dev->S, tk, Dinv, x, ld, xo, ldo, nrhs, rchunk );
        else             k_solve_dinv<T, BWD, 512><<<grid, 512, lds, dev->stream>>>( dev->S, tk, Dinv, x, ld, xo, ldo, nrhs, rchunk );
    }
    else if ( ncmax <= PA_SD_MAXCOL ) {
        const size_t  avail  = (size_t)dev->lds_limit - 4096;
        const int32_t rchunk = (int32_t)std::max<size_t>( 1, std::min<size_t>( nrhs, avail / ( (size_t)ncmax * sizeof( T ) ) ) );
        dim3          grid( ncb, ( nrhs + rchunk - 1 ) / rchunk );
        size_t        lds = (size_t)ncmax * rchunk * sizeof( T );
        if ( ncmax <= 64 )       pa_solve_diag_launch<T, BWD, UNIT, 64>( dev, grid, lds, cb, C, CT, ct_off, x, ld, xo, ldo, nrhs, rchunk );
        else if ( ncmax <= 128 ) pa_solve_diag_launch<T, BWD, UNIT, 128>( dev, grid, lds, cb, C, CT, ct_off, x, ld, xo, ldo, nrhs, rchunk );
        else if ( ncmax <= 256 ) pa_solve_diag_launch<T, BWD, UNIT, 256>( dev, grid, lds, cb, C, CT, ct_off, x, ld, xo, ldo, nrhs, rchunk );
        else                     pa_solve_diag_launch<T, BWD, UNIT, 512>( dev, grid, lds, cb, C, CT, ct_off, x, ld, xo, ldo, nrhs, rchunk );
    }
    else {
        size_t lds = (size_t)ncmax * nrhs * sizeof( T );
        if ( lds > (size_t)dev->lds_limit - 4096 ) return PASTIX_ERR_BADPARAMETER;
        k_solve_diag<T, BWD, UNIT><<<ncb, 256, lds, dev->stream>>>( dev->S, cb, C, x, nrhs, ld, xo, ldo );
    }
    PA_HIP( hipGetLastError() );
    return PASTIX_SUCCESS;
}

/* One forward sweep (op(C) y = b, C in lower 2D-tile storage, C's own
 * diagonal blocks) or backward sweep (op(C)^T x = y; CT/ct_off locate the
 * transposes of C's diagonal blocks) over every level, from the right-hand
 * sides in x to the solution in xo.  x is the working array: the updates
 * are added to it and the diagonal steps read it; the diagonal steps write
 * their results to xo, which the updates of the same sweep then read
 * (forward: the solved x_k of the level's own cblks; backward: the solved
 * rows of the cblks above).  With xo another array, every level's diagonal
 * step can split a cblk over several workgroups; with xo == x the sweep
 * runs in place, one workgroup per cblk. */
template <typename T>
static int
pa_sweep_fwd_t( pa_device_s *dev, const T *Cf, bool unit_f, T *x, int64_t ld, T *xo, int64_t ldo, int32_t nrhs,
                const T *Dinv )
{
    const size_t nl = dev->levels.size();
    int rc;
    for ( size_t li = 0; li < nl; li++ ) {
        const pa_level_t &lev = dev->levels[li];
        if ( lev.cblk_end == lev.cblk_beg ) continue;
        rc = unit_f ? pa_solve_diag_level<T, false, true>( dev, lev, Cf, nullptr, nullptr, x, ld, xo, ldo, nrhs, Dinv )
                    : pa_solve_diag_level<T, false, false>( dev, lev, Cf, nullptr, nullptr, x, ld, xo, ldo, nrhs, Dinv );
        if ( rc ) return rc;
        int64_t ns = lev.solve_end - lev.solve_beg;
        if ( ns > 0 && nrhs == 1 )
            k_solve_fwd_update<T, 1><<<(unsigned)ns, 256, 0, dev->stream>>>( dev->S, dev->d_solve_tasks + lev.solve_beg, Cf, x,
                                                                            nrhs, ld, xo, ldo );
        else if ( ns > 0 )
            k_solve_fwd_update<T, 4><<<(unsigned)ns, 256, 0, dev->stream>>>( dev->S, dev->d_solve_tasks + lev.solve_beg, Cf, x,
                                                                            nrhs, ld, xo, ldo );
        const int32_t nsm = (int32_t)( lev.cblk_end - lev.fac_end );
        if ( nsm > 0 ) pa_small_solve_launch<T, false>( dev, lev, Cf, x, ld, xo, ldo, nrhs );
        PA_HIP( hipGetLastError() );
    }
    return PASTIX_SUCCESS;
}

template <typename T>
static int
pa_sweep_bwd_t( pa_device_s *dev, const T *Cb, bool unit_b, const T *CbT, const int64_t *cbt_off, T *x, int64_t ld,
                T *xo, int64_t ldo, int32_t nrhs, const T *Dinv )
{
    const size_t nl = dev->levels.size();
    int rc;
    for ( size_t li = nl; li-- > 0; ) {
        const pa_level_t &lev = dev->levels[li];
        if ( lev.cblk_end == lev.cblk_beg ) continue;
        int64_t ns = lev.bsolve_end - lev.bsolve_beg;
        if ( ns > 0 && nrhs == 1 )
            k_solve_bwd_update<T, false><<<(unsigned)ns, 256, 0, dev->stream>>>( dev->S, dev->d_bsolve_tasks + lev.bsolve_beg, Cb,
                                                                                x, nrhs, ld, xo, ldo );
        else if ( ns > 0 )
            k_solve_bwd_update<T, true><<<(unsigned)ns, 256, 0, dev->stream>>>( dev->S, dev->d_bsolve_tasks + lev.bsolve_beg, Cb,
                                                                               x, nrhs, ld, xo, ldo );
        const int32_t nsm = (int32_t)( lev.cblk_end - lev.fac_end );
        if ( nsm > 0 ) pa_small_solve_launch<T, true>( dev, lev, Cb, x, ld, xo, ldo, nrhs );
        PA_HIP( hipGetLastError() );
        rc = unit_b ? pa_solve_diag_level<T, true, true>( dev, lev, Cb, CbT, cbt_off, x, ld, xo, ldo, nrhs, Dinv )
                    : pa_solve_diag_level<T, true, false>( dev, lev, Cb, CbT, cbt_off, x, ld, xo, ldo, nrhs, Dinv );
        if ( rc ) return rc;
    }
    return PASTIX_SUCCESS;
}

static unsigned
pa_grid( int64_t n )
{
    int64_t g = ( n + 255 ) / 256;
    return (unsigned)std::max<int64_t>( 1, std::min<int64_t>( g, 65536 ) );
}

/* A step of a solve (reference sopalin/pastix_task_solve.c:596-649 and
 * pastix_subtask_trsm / pastix_subtask_diag): a forward or backward sweep over
 * one factor array (side 0: L, side 1: the second array, U^T for LU), or the
 * diagonal scaling x <- D^{-1} x with D the diagonal of the L diagonal
 * blocks (reference kernels/solve_ztrsmsp.c:664-716). */
enum { PA_OP_FWD = 1, PA_OP_BWD = 2, PA_OP_DIAG = 3 };
struct pa_sweep_op {
    int kind, side, unit;
};

/* Run the steps on x (n x nrhs, leading dimension ld).  One right-hand
 * side: alt (n, leading dimension n) is the second array of the sweeps,
 * each sweep reads one of the two and leaves its result in the other, so
 * that the diagonal steps spread the wide cblks of the top levels over
 * several workgroups.  Several right-hand sides: the sweeps run in place on
 * x (alt is not used): the work of a level is wide enough, and the narrow
 * cblks' kernels then update each right-hand-side entry where they read it
 * instead of streaming a second array.  On return *res / *ldres give the
 * array that holds the result (x after an even number of sweeps). */
template <typename T>
static int
pa_run_ops_t( pa_device_s *dev, const pa_sweep_op *ops, int nops, T *x, int32_t nrhs, int64_t ld, T *alt, T **res,
              int64_t *ldres )
{
    int rc;
    const bool inplace = nrhs > 1 || alt == nullptr;
    T      *cur = x, *oth = inplace ? x : alt;
    int64_t ldc = ld, ldt = inplace ? ld : dev->n;
    for ( int o = 0; o < nops; o++ ) {
        const pa_sweep_op &op = ops[o];
        const T *C = (const T *)( op.side ? dev->d_U : dev->d_L );
        /* the inverted diagonal blocks hold the inverse of the side's lower
         * triangle with the diagonal the factorization gives it (LL^t: L
         * non-unit; LU: L unit, U^T non-unit); other requests substitute */
        const int natural_unit = ( dev->fact_kind == PastixFactLU && op.side == 0 ) ? 1 : 0;
        const T  *Dinv = ( dev->d_dinv && !dev->ldlt && op.unit == natural_unit )
                             ? (const T *)dev->d_dinv + ( op.side ? dev->dtnbr : 0 ) : nullptr;
        if ( op.kind == PA_OP_FWD ) {
            if ( ( rc = pa_sweep_fwd_t<T>( dev, C, op.unit != 0, cur, ldc, oth, ldt, nrhs, Dinv ) ) ) return rc;
            std::swap( cur, oth );
            std::swap( ldc, ldt );
        }
        else if ( op.kind == PA_OP_BWD ) {
            /* transposed diagonal blocks of the swept factor: for two-array
             * factorizations the other array's diagonal square, for LL^t the
             * separate transposed array */
            const bool     two = dev->d_U != nullptr;
            const int64_t *ct_off = two ? dev->d_c_coef : dev->d_c_dtoff;
            const T       *ct     = (const T *)( two ? ( op.side ? dev->d_L : dev->d_U ) : dev->d_DT );
            if ( !two && dev->all_dinv && !Dinv && dev->S.cblknbr > 0 ) {
                /* a substitution sweep where the factorization left only the
                 * inverted blocks: transpose the diagonal blocks first */
                k_diag_transpose<T><<<dev->S.cblknbr, 256, 0, dev->stream>>>( dev->S, (const T *)dev->d_L,
                                                                             (T *)dev->d_DT );
                PA_HIP( hipGetLastError() );
            }
            if ( ( rc = pa_sweep_bwd_t<T>( dev, C, op.unit != 0, ct, ct_off, cur, ldc, oth, ldt, nrhs, Dinv ) ) ) return rc;
            std::swap( cur, oth );
            std::swap( ldc, ldt );
        }
        else {
            if ( dev->n > 0 ) {
                dim3 grid( pa_grid( dev->n ), (unsigned)std::min<int32_t>( nrhs, 64 ) );
                k_solve_dscale<T><<<grid, 256, 0, dev->stream>>>( dev->S, (int32_t)dev->n, dev->d_col2cblk,
                                                                 (const T *)dev->d_L, cur, nrhs, ldc );
                PA_HIP( hipGetLastError() );
            }
        }
    }
    *res   = cur;
    *ldres = ldc;
    return PASTIX_SUCCESS;
}

/* Launch a sequence of solve steps on a device FP64 buffer in the
 * factorization numbering, the result back in d_x.  FP64 factors use alt
 * (n*nrhs doubles) as the second array of the sweeps; FP32 factors work on
 * the FP32 copy xs with alt (n*nrhs floats) as the second array (down- and
 * up-conversion, reference sopalin/pastix_task_solve.c:583-592,
 * sopalin/pastix_rhs.c:162-242). */
static int
pa_launch_ops( pa_device_s *dev, const pa_sweep_op *ops, int nops, int32_t nrhs, double *d_x, int64_t ld, float *xs,
               void *alt )
{
    const int64_t n = dev->n;
    int           rc;
    if ( dev->prec == PA_PREC_DOUBLE ) {
        double *res;
        int64_t ldr;
        if ( ( rc = pa_run_ops_t<double>( dev, ops, nops, d_x, nrhs, ld, (double *)alt, &res, &ldr ) ) ) return rc;
        if ( res != d_x )
            for ( int32_t r = 0; r < nrhs; r++ )
                k_convert<double, double><<<pa_grid( n ), 256, 0, dev->stream>>>( n, res + r * ldr, d_x + r * ld );
        PA_HIP( hipGetLastError() );
        return PASTIX_SUCCESS;
    }
    for ( int32_t r = 0; r < nrhs; r++ )
        k_convert<float, double><<<pa_grid( n ), 256, 0, dev->stream>>>( n, d_x + r * ld, xs + r * n );
    float  *res;
    int64_t ldr;
    if ( ( rc = pa_run_ops_t<float>( dev, ops, nops, xs, nrhs, n, (float *)alt, &res, &ldr ) ) ) return rc;
    for ( int32_t r = 0; r < nrhs; r++ )
        k_convert<double, float><<<pa_grid( n ), 256, 0, dev->stream>>>( n, res + r * ldr, d_x + r * ld );
    PA_HIP( hipGetLastError() );
    return PASTIX_SUCCESS;
}

static int
pa_check_ops( const pa_device_s *dev, const pa_sweep_op *ops, int nops )
{
    if ( !dev->d_L ) return PASTIX_ERR_BADPARAMETER;
    for ( int o = 0; o < nops; o++ )
        if ( ops[o].side && !dev->d_U ) return PASTIX_ERR_BADPARAMETER;
    return PASTIX_SUCCESS;
}

/* Run a sequence of solve steps (pa_launch_ops) with the handle's FP32
 * buffer and second sweep array (one right-hand side), replayed from a
 * cached hipGraph per (buffer, nrhs, ld, steps). */
static int
pa_dev_run_ops( pastix_data_t *pd, const pa_sweep_op *ops, int nops, int32_t nrhs, double *d_x, int64_t ld )
{
    pa_device_s *dev = DEV( pd );
    PA_HIP( hipSetDevice( dev->device ) );
    if ( pa_check_ops( dev, ops, nops ) ) return PASTIX_ERR_BADPARAMETER;
    int64_t len = dev->n * (int64_t)nrhs;
    if ( dev->prec != PA_PREC_DOUBLE && dev->xs_len < len ) {
        hipFree( dev->d_xs );
        dev->d_xs = nullptr;
        PA_HIP( hipMalloc( (void **)&dev->d_xs, std::max<int64_t>( len, 1 ) * sizeof( float ) ) );
        dev->xs_len = len;
    }
    /* the second array of the one-right-hand-side sweeps */
    const size_t alt = (size_t)std::max<int64_t>( dev->n, 1 ) * ( dev->prec == PA_PREC_DOUBLE ? sizeof( double ) : sizeof( float ) );
    if ( nrhs == 1 && dev->alt_bytes < alt ) {
        hipFree( dev->d_alt );
        dev->d_alt = nullptr;
        PA_HIP( hipMalloc( &dev->d_alt, alt ) );
        dev->alt_bytes = alt;
    }
    int64_t code = 0;
    for ( int o = 0; o < nops; o++ ) code = code * 16 + ops[o].kind * 4 + ops[o].side * 2 + ops[o].unit;
    auto launch = [&]() -> int { return pa_launch_ops( dev, ops, nops, nrhs, d_x, ld, dev->d_xs, dev->d_alt ); };
    if ( !dev->use_graphs ) return launch();
    /* cached graph per (buffer, nrhs, ld, steps) among the most recent ones */
    pa_graph_entry *slot = nullptr, *oldest = &dev->solve_graphs[0];
    for ( auto &g : dev->solve_graphs ) {
        if ( g.exec && g.key[0] == d_x && g.key[1] == dev->d_xs && g.key[2] == dev->d_L && g.key[3] == dev->d_U &&
             g.key[4] == dev->d_alt &&
             g.ikey[0] == nrhs && g.ikey[1] == ld && g.ikey[2] == code && g.ikey[3] == dev->prec ) {
            slot = &g;
            break;
        }
        if ( g.stamp < oldest->stamp ) oldest = &g;
    }
    bool hit = slot != nullptr;
    if ( !hit ) slot = oldest;
    int rc = pa_graph_run( dev, *slot, hit, launch );
    if ( rc ) return rc;
    slot->key[0] = d_x; slot->key[1] = dev->d_xs; slot->key[2] = dev->d_L; slot->key[3] = dev->d_U;
    slot->key[4] = dev->d_alt;
    slot->ikey[0] = nrhs; slot->ikey[1] = ld; slot->ikey[2] = code; slot->ikey[3] = dev->prec;
    slot->stamp = ++dev->graph_clock;
    return PASTIX_SUCCESS;
}

/*
 * Steps of the full solve (reference sopalin/pastix_task_solve.c:543-649):
 *   LL^t:        L y = b (non-unit),  L^t x = y (non-unit)
 *   LDL^t:       L y = b (unit),      D z = y,  L^t x = z (unit)
 *   LU:          L y = b (unit),      U x = y (U^T tiles, non-unit)
 *   LU, trans:   U^t y = b (U^T tiles, non-unit), L^t x = y (unit)
 */
static int
pa_solve_ops( int fact, pastix_trans_t trans, pa_sweep_op *ops )
{
    const bool tr   = ( trans != PastixNoTrans );
    int        nops = 0;
    if ( fact == PastixFactLU ) {
        ops[nops++] = { PA_OP_FWD, tr ? 1 : 0, tr ? 0 : 1 };
        ops[nops++] = { PA_OP_BWD, tr ? 0 : 1, tr ? 1 : 0 };
    }
    else if ( pa_is_ldlt( fact ) ) {
        ops[nops++] = { PA_OP_FWD, 0, 1 };
        ops[nops++] = { PA_OP_DIAG, 0, 0 };
        ops[nops++] = { PA_OP_BWD, 0, 1 };
    }
    else {
        ops[nops++] = { PA_OP_FWD, 0, 0 };
        ops[nops++] = { PA_OP_BWD, 0, 0 };
    }
    return nops;
}

extern "C" int
pa_dev_solve_perm( pastix_data_t *pd, pastix_trans_t trans, int32_t nrhs, double *d_x, int64_t ld )
{
    pa_sweep_op ops[3];
    const int   nops = pa_solve_ops( DEV( pd )->fact_kind, trans, ops );
    return pa_dev_run_ops( pd, ops, nops, nrhs, d_x, ld );
}

/* Scratch bytes the allocation-free solve needs for n * nrhs entries:
 * FP64 copy and second array (FP64 factors), or FP64 copy, FP32 copy and
 * FP32 second array (FP32 factors); 16 bytes per entry either way */
static size_t
pa_noalloc_bytes( const pa_device_s *dev, int64_t len, size_t *xb, size_t *sb )
{
    *xb = ( (size_t)len * sizeof( double ) + 255 ) & ~(size_t)255;
    *sb = ( dev->prec != PA_PREC_DOUBLE ) ? ( ( (size_t)len * sizeof( float ) + 255 ) & ~(size_t)255 ) : 0;
    return *xb + *sb + (size_t)len * ( dev->prec != PA_PREC_DOUBLE ? sizeof( float ) : sizeof( double ) );
}

/* Largest n * nrhs the allocation-free solve takes */
extern "C" int64_t
pa_dev_noalloc_max( pastix_data_t *pd )
{
    return (int64_t)( ( PA_SCRATCH_BYTES - 512 ) / 16 );
}

/* The full solve of pa_dev_solve_perm plus the permutations P b and P^T x,
 * on d_b in the user numbering, with no device memory allocated: the
 * permuted right-hand sides, their FP32 copy and the second array of the
 * sweeps live in this device's `scratch` array (every entry is written
 * before it is read: nothing needs clearing).  Eager launches on the handle's stream, then a stream
 * synchronization.  Calls that overlap on one device share the array and
 * race. */
extern "C" int
pa_dev_solve_noalloc( pastix_data_t *pd, pastix_trans_t trans, int32_t nrhs, double *d_b, int64_t ldb )
{
    pa_device_s *dev = DEV( pd );
    PA_HIP( hipSetDevice( dev->device ) );
    pa_sweep_op  ops[3];
    const int    nops = pa_solve_ops( dev->fact_kind, trans, ops );
    if ( pa_check_ops( dev, ops, nops ) ) return PASTIX_ERR_BADPARAMETER;
    const int64_t n = dev->n;
    size_t        xb, sb;
    const size_t  need = pa_noalloc_bytes( dev, n * (int64_t)nrhs, &xb, &sb );
    if ( nrhs < 1 || need > PA_SCRATCH_BYTES || !dev->d_scratch ) {
        fprintf( stderr,
                 "pastix_amd_solve_dev_noalloc: n * nrhs = %lld needs %zu bytes of scratch, the limit is %zu\n",
                 (long long)( n * nrhs ), need, (size_t)PA_SCRATCH_BYTES );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( n == 0 ) return PASTIX_SUCCESS;
    double *xp  = (double *)dev->d_scratch;
    float  *xs  = sb ? (float *)( (char *)dev->d_scratch + xb ) : nullptr;
    void   *alt = (char *)dev->d_scratch + xb + sb;
    int     rc  = pa_dev_apply_perm( pd, 1, nrhs, d_b, ldb, xp, n );
    if ( !rc ) rc = pa_launch_ops( dev, ops, nops, nrhs, xp, n, xs, alt );
    if ( !rc ) rc = pa_dev_apply_perm( pd, 0, nrhs, xp, n, d_b, ldb );
    if ( rc ) return rc;
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    return PASTIX_SUCCESS;
}

/* One triangular sweep in the factorization numbering (reference
 * pastix_subtask_trsm, sopalin/pastix_task_solve.c:296 and the cases of
 * kernels/solve_ztrsmsp.c:326-350, 491-510 with side = Left):
 *   (Lower, NoTrans) forward over L, (Lower, Trans) backward over L,
 *   (Upper, Trans) forward over the stored U^T, (Upper, NoTrans) backward
 *   over the stored U^T. */
extern "C" int
pa_dev_trsm_perm( pastix_data_t *pd, int upper, int transposed, int unit, int32_t nrhs, double *d_x, int64_t ld )
{
    pa_device_s *dev = DEV( pd );
    if ( upper && !dev->d_U ) return PASTIX_ERR_BADPARAMETER;
    pa_sweep_op op = { ( upper ? !transposed : transposed ) ? PA_OP_BWD : PA_OP_FWD, upper ? 1 : 0, unit ? 1 : 0 };
    return pa_dev_run_ops( pd, &op, 1, nrhs, d_x, ld );
}

/* x <- D^{-1} x (reference pastix_subtask_diag, sopalin/pastix_task_solve.c:422) */
extern "C" int
pa_dev_diag_perm( pastix_data_t *pd, int32_t nrhs, double *d_x, int64_t ld )
{
    pa_sweep_op op = { PA_OP_DIAG, 0, 0 };
    return pa_dev_run_ops( pd, &op, 1, nrhs, d_x, ld );
}

extern "C" int
pa_dev_spmv_perm( pastix_data_t *pd, pastix_trans_t trans, double alpha, const double *d_x, double beta, double *d_y )
{
    pa_device_s *dev = DEV( pd );
    int64_t n = dev->n;
    /* NoTrans uses A^T stored by columns (Uvalues), Trans uses Lvalues */
    const double *vals = ( trans == PastixNoTrans ) ? dev->d_Uv : dev->d_Lv;
    unsigned grid = (unsigned)( ( n * 8 + 255 ) / 256 );
    if ( n > 0 )
        k_spmv<<<grid, 256, 0, dev->stream>>>( (int32_t)n, dev->d_colptr, dev->d_rowtab, vals, alpha, d_x, beta, d_y );
    PA_HIP( hipGetLastError() );
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_apply_perm( pastix_data_t *pd, int forward, int32_t nrhs, const double *d_src, int64_t lds, double *d_dst, int64_t ldd )
{
    pa_device_s *dev = DEV( pd );
    int64_t n = dev->n;
    if ( n == 0 ) return PASTIX_SUCCESS;
    /* forward: x_perm[i] = x[peritab[i]]  (P b); backward: x[j] = x_perm[permtab[j]]
     * (both as gathers: coalesced writes) */
    dim3 grid( std::min<unsigned>( pa_grid( n ), 4096 ), (unsigned)std::min<int32_t>( nrhs, 64 ) );
    k_gather<double, double><<<grid, 256, 0, dev->stream>>>( n, nrhs, forward ? dev->d_invp : dev->d_perm, d_src, lds,
                                                             d_dst, ldd );
    PA_HIP( hipGetLastError() );
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* vector helpers                                                            */
/* ------------------------------------------------------------------------- */
extern "C" void *
pa_dev_malloc( size_t bytes )
{
    void *p = NULL;
    if ( hipMalloc( &p, std::max<size_t>( bytes, 8 ) ) != hipSuccess ) return NULL;
    return p;
}
extern "C" void
pa_dev_free( void *p )
{
    if ( p ) hipFree( p );
}
extern "C" void *
pa_dev_wget( pastix_data_t *pd, size_t bytes )
{
    pa_device_s *dev = DEV( pd );
    if ( !dev ) return NULL;
    bytes = ( std::max<size_t>( bytes, 256 ) + 255 ) & ~(size_t)255;
    int best = -1;
    for ( size_t i = 0; i < dev->wpool.size(); i++ ) {
        const pa_wblock &w = dev->wpool[i];
        if ( !w.used && w.bytes >= bytes && ( best < 0 || w.bytes < dev->wpool[best].bytes ) ) best = (int)i;
    }
    if ( best < 0 ) {
        void *p = NULL;
        if ( hipMalloc( &p, bytes ) != hipSuccess ) {
            /* release the free blocks (after the stream has drained) and retry once */
            hipGetLastError();
            hipStreamSynchronize( dev->stream );
            for ( size_t i = 0; i < dev->wpool.size(); ) {
                if ( !dev->wpool[i].used ) {
                    hipFree( dev->wpool[i].p );
                    dev->wpool.erase( dev->wpool.begin() + i );
                }
                else {
                    i++;
                }
            }
            pa_graphs_drop( dev );
            if ( hipMalloc( &p, bytes ) != hipSuccess ) {
                hipGetLastError();
                return NULL;
            }
        }
        dev->wpool.push_back( { p, bytes, 0 } );
        best = (int)dev->wpool.size() - 1;
    }
    dev->wpool[best].used = 1;
    return dev->wpool[best].p;
}
extern "C" void
pa_dev_wput( pastix_data_t *pd, void *p )
{
    pa_device_s *dev = DEV( pd );
    if ( !dev || !p ) return;
    for ( pa_wblock &w : dev->wpool ) {
        if ( w.p == p ) {
            w.used = 0;
            return;
        }
    }
}
extern "C" int
pa_dev_memcpy_h2d( pastix_data_t *pd, void *d, const void *h, size_t bytes )
{
    PA_HIP( hipMemcpyAsync( d, h, bytes, hipMemcpyHostToDevice, DEV( pd )->stream ) );
    PA_HIP( hipStreamSynchronize( DEV( pd )->stream ) );
    return PASTIX_SUCCESS;
}
extern "C" int
pa_dev_memcpy_d2h( pastix_data_t *pd, void *h, const void *d, size_t bytes )
{
    PA_HIP( hipMemcpyAsync( h, d, bytes, hipMemcpyDeviceToHost, DEV( pd )->stream ) );
    PA_HIP( hipStreamSynchronize( DEV( pd )->stream ) );
    return PASTIX_SUCCESS;
}
extern "C" int
pa_dev_memcpy_d2d( pastix_data_t *pd, void *d, const void *s, size_t bytes )
{
    PA_HIP( hipMemcpyAsync( d, s, bytes, hipMemcpyDeviceToDevice, DEV( pd )->stream ) );
    return PASTIX_SUCCESS;
}
extern "C" int
pa_dev_sync( pastix_data_t *pd )
{
    PA_HIP( hipStreamSynchronize( DEV( pd )->stream ) );
    return PASTIX_SUCCESS;
}

static int
pa_reduce_partials( pa_device_s *dev, int count, double *out, int nout )
{
    PA_HIP( hipMemcpyAsync( dev->h_partial, dev->d_partial, (size_t)count * nout * sizeof( double ),
                            hipMemcpyDeviceToHost, dev->stream ) );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    for ( int j = 0; j < nout; j++ ) {
        double s = 0.;
        for ( int i = 0; i < count; i++ ) s += dev->h_partial[j * count + i];
        out[j] = s;
    }
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_mdot( pastix_data_t *pd, int64_t n, int32_t m, const double *X, const double *v, double *out )
{
    pa_device_s *dev = DEV( pd );
    int blocks = (int)std::min<int64_t>( ( n + 255 ) / 256, 4096 / std::max( m, 1 ) );
    blocks = std::max( blocks, 1 );
    blocks = std::min( blocks, 256 );
    dim3 grid( blocks, m );
    k_mdot<<<grid, 256, 0, dev->stream>>>( n, X, n, v, dev->d_partial );
    PA_HIP( hipGetLastError() );
    return pa_reduce_partials( dev, blocks, out, m );
}

extern "C" double
pa_dev_dot( pastix_data_t *pd, int64_t n, const double *x, const double *y )
{
    double out = 0.;
    pa_device_s *dev = DEV( pd );
    int blocks = (int)std::max<int64_t>( 1, std::min<int64_t>( ( n + 255 ) / 256, 256 ) );
    k_mdot<<<dim3( blocks, 1 ), 256, 0, dev->stream>>>( n, x, n, y, dev->d_partial );
    pa_reduce_partials( dev, blocks, &out, 1 );
    return out;
}

extern "C" double
pa_dev_nrm2( pastix_data_t *pd, int64_t n, const double *x )
{
    return sqrt( pa_dev_dot( pd, n, x, x ) );
}

extern "C" int
pa_dev_axpby( pastix_data_t *pd, int64_t n, double a, const double *x, double b, double *y )
{
    k_axpby<<<pa_grid( n ), 256, 0, DEV( pd )->stream>>>( n, a, x, b, y );
    PA_HIP( hipGetLastError() );
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_gemv( pastix_data_t *pd, int64_t n, int32_t m, const double *X, const double *c, double alpha, double *y )
{
    pa_device_s *dev = DEV( pd );
    double *dc = dev->d_partial; /* m <= 4096 coefficients */
    PA_HIP( hipMemcpyAsync( dc, c, m * sizeof( double ), hipMemcpyHostToDevice, dev->stream ) );
    k_gemv_n<<<pa_grid( n ), 256, 0, dev->stream>>>( n, m, X, n, dc, alpha, y );
    PA_HIP( hipGetLastError() );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    return PASTIX_SUCCESS;
}

extern "C" int
pa_dev_get_coeftab( const pastix_data_t *pd, int side, double *host )
{
    pa_device_s *dev = DEV( pd );
    void *src = side ? dev->d_U : dev->d_L;
    if ( !src ) return PASTIX_ERR_BADPARAMETER;
    PA_HIP( hipSetDevice( dev->device ) );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    if ( dev->prec == PA_PREC_DOUBLE ) {
        PA_HIP( hipMemcpy( host, src, dev->coefnbr * sizeof( double ), hipMemcpyDeviceToHost ) );
    }
    else {
        std::vector<float> tmp( dev->coefnbr );
        PA_HIP( hipMemcpy( tmp.data(), src, dev->coefnbr * sizeof( float ), hipMemcpyDeviceToHost ) );
        for ( int64_t i = 0; i < dev->coefnbr; i++ ) host[i] = tmp[i];
    }
    return PASTIX_SUCCESS;
}

/* profiling of the factorization kernels by class (diag, trsm, gemm) */
extern "C" int
pa_dev_profile( pastix_data_t *pd, int enable, double *out )
{
    pa_device_s *dev = DEV( pd );
    if ( !dev ) return PASTIX_ERR_BADPARAMETER;
    if ( enable >= 0 ) {
        dev->profiling = enable;
        for ( int c = 0; c < 3; c++ ) { dev->prof_ms[c] = 0.; dev->prof_launch[c] = 0; }
    }
    if ( out ) {
        for ( int c = 0; c < 3; c++ ) {
            out[c]     = dev->prof_ms[c];
            out[3 + c] = (double)dev->prof_launch[c];
            out[6 + c] = dev->class_flops[c];
        }
    }
    return PASTIX_SUCCESS;
}

/* out[e] = alpha sum_r u[row(e), r] v[col(e), r] on a CSC pattern in the user
 * numbering (device arrays), on the library stream */
extern "C" int
pastix_amd_sampled_outer( pastix_data_t *pd, const void *d_colptr, const void *d_rows, pastix_int_t nrhs,
                          const void *d_u, pastix_int_t ldu, const void *d_v, pastix_int_t ldv, double alpha,
                          void *d_out )
{
    if ( pd == NULL || DEV( pd ) == NULL ) return PASTIX_ERR_BADPARAMETER;
    pa_device_s *dev = DEV( pd );
    const int64_t n = dev->n;
    if ( n <= 0 || nrhs < 1 || ldu < n || ldv < n ) return PASTIX_ERR_BADPARAMETER;
    PA_HIP( hipSetDevice( dev->device ) );
    k_sampled_outer<<<(unsigned)( ( n + 255 ) / 256 ), 256, 0, dev->stream>>>(
        n, (const int64_t *)d_colptr, (const int64_t *)d_rows, (int32_t)nrhs, (const double *)d_u, ldu,
        (const double *)d_v, ldv, alpha, (double *)d_out );
    PA_HIP( hipGetLastError() );
    PA_HIP( hipStreamSynchronize( dev->stream ) );
    return PASTIX_SUCCESS;
}
