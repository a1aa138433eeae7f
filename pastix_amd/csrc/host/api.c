/*
 * api.c -- the PaStiX public API on top of the CPU analysis and the GPU
 * numerical engine.
 *
 * Step order, parameter defaults, error codes and output parameters follow
 * the reference (common/api.c:420-603, 905-960; blend/pastix_task_analyze.c:53;
 * sopalin/pastix_task_sopalin.c:93-710; sopalin/pastix_task_solve.c:169-809).
 */
#include <math.h>
#include <stdarg.h>
#include <time.h>
#include <float.h>
#include <unistd.h>
#include "internal.h"
#include "../hip/pa_device.h"

void
pa_error( const char *fmt, ... )
{
    va_list ap;
    va_start( ap, fmt );
    fprintf( stderr, "pastix_amd error: " );
    vfprintf( stderr, fmt, ap );
    fprintf( stderr, "\n" );
    va_end( ap );
}

void
pa_warning( const char *fmt, ... )
{
    va_list ap;
    va_start( ap, fmt );
    fprintf( stderr, "pastix_amd warning: " );
    vfprintf( stderr, fmt, ap );
    fprintf( stderr, "\n" );
    va_end( ap );
}

double
pa_wtime( void )
{
    struct timespec ts;
    clock_gettime( CLOCK_MONOTONIC, &ts );
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Defaults of reference common/api.c:420-603.  This build has no Scotch or
 * Metis: the default ordering is the personal one (a grid ordering given to
 * pastix_subtask_order, or the elimination-tree fallback). */
void
pastixInitParam( pastix_int_t *iparm, double *dparm )
{
    memset( iparm, 0, IPARM_SIZE * sizeof( pastix_int_t ) );
    memset( dparm, 0, DPARM_SIZE * sizeof( double ) );
    iparm[IPARM_VERBOSE]                        = PastixVerboseNo;
    iparm[IPARM_IO_STRATEGY]                    = PastixIONo;
    iparm[IPARM_TRACE]                          = PastixTraceNot;
    iparm[IPARM_ORDERING]                       = PastixOrderPersonal;
    iparm[IPARM_ORDERING_DEFAULT]               = 1;
    iparm[IPARM_SCOTCH_MT]                      = 1;
    iparm[IPARM_SCOTCH_SWITCH_LEVEL]            = 120;
    iparm[IPARM_SCOTCH_CMIN]                    = 0;
    iparm[IPARM_SCOTCH_CMAX]                    = 100000;
    iparm[IPARM_SCOTCH_FRAT]                    = 8;
    iparm[IPARM_METIS_CTYPE]                    = -1;
    iparm[IPARM_METIS_RTYPE]                    = -1;
    iparm[IPARM_METIS_NSEPS]                    = 1;
    iparm[IPARM_METIS_NITER]                    = 10;
    iparm[IPARM_METIS_UFACTOR]                  = 200;
    iparm[IPARM_METIS_COMPRESS]                 = 1;
    iparm[IPARM_METIS_SEED]                     = 3452;
    iparm[IPARM_AMALGAMATION_LVLCBLK]           = 5;
    iparm[IPARM_AMALGAMATION_LVLBLAS]           = 5;
    iparm[IPARM_REORDERING_SPLIT]               = 0;
    iparm[IPARM_REORDERING_STOP]                = PASTIX_INT_MAX;
    iparm[IPARM_SPLITTING_STRATEGY]             = PastixSplitKway;
    iparm[IPARM_SPLITTING_LEVELS_PROJECTIONS]   = 0;
    iparm[IPARM_SPLITTING_LEVELS_KWAY]          = PASTIX_INT_MAX;
    iparm[IPARM_SPLITTING_PROJECTIONS_DEPTH]    = 3;
    iparm[IPARM_SPLITTING_PROJECTIONS_DISTANCE] = 3;
    iparm[IPARM_SPLITTING_PROJECTIONS_WIDTH]    = 1;
    iparm[IPARM_MIN_BLOCKSIZE]                  = 160;
    iparm[IPARM_MAX_BLOCKSIZE]                  = 320;
    iparm[IPARM_TASKS2D_LEVEL]                  = -1;
    iparm[IPARM_TASKS2D_WIDTH]                  = iparm[IPARM_MIN_BLOCKSIZE];
    iparm[IPARM_FACTORIZATION]                  = PastixFactLU;
    iparm[IPARM_FACTO_LOOK_SIDE]                = PastixFactRightLooking;
    iparm[IPARM_SCHUR_FACT_MODE]                = PastixFactModeLocal;
    iparm[IPARM_TRANSPOSE_SOLVE]                = PastixNoTrans;
    iparm[IPARM_SCHUR_SOLV_MODE]                = PastixSolvModeLocal;
    iparm[IPARM_APPLYPERM_WS]                   = 1;
    iparm[IPARM_REFINEMENT]                     = PastixRefineGMRES;
    iparm[IPARM_ITERMAX]                        = 250;
    iparm[IPARM_GMRES_IM]                       = 25;
    iparm[IPARM_SCHEDULER]                      = PastixSchedDynamic;
    iparm[IPARM_THREAD_NBR]                     = -1;
    iparm[IPARM_SOCKET_NBR]                     = -1;
    iparm[IPARM_GPU_NBR]                        = 1;
    iparm[IPARM_GPU_MEMORY_PERCENTAGE]          = 95;
    iparm[IPARM_GPU_MEMORY_BLOCK_SIZE]          = 32 * 1024;
    iparm[IPARM_GLOBAL_ALLOCATION]              = 1;
    iparm[IPARM_COMPRESS_MIN_WIDTH]             = 128;
    iparm[IPARM_COMPRESS_MIN_HEIGHT]            = 20;
    iparm[IPARM_COMPRESS_WHEN]                  = PastixCompressNever;
    iparm[IPARM_COMPRESS_METHOD]                = PastixCompressMethodPQRCP;
    iparm[IPARM_COMPRESS_ORTHO]                 = PastixCompressOrthoCGS;
    iparm[IPARM_COMPRESS_PRESELECT]             = 1;
    iparm[IPARM_COMPRESS_ILUK]                  = -2;
    iparm[IPARM_MIXED]                          = 0;
    iparm[IPARM_MPI_THREAD_LEVEL]               = PastixMpiNone;
    iparm[IPARM_MODIFY_PARAMETER]               = 1;
    iparm[IPARM_START_TASK]                     = PastixTaskOrdering;
    iparm[IPARM_END_TASK]                       = PastixTaskClean;
    iparm[IPARM_FLOAT]                          = -1;
    iparm[IPARM_MTX_TYPE]                       = -1;
    iparm[IPARM_DOF_NBR]                        = 1;
    dparm[DPARM_EPSILON_REFINEMENT]             = -1.;
    dparm[DPARM_RELATIVE_ERROR]                 = -1.;
    dparm[DPARM_EPSILON_MAGN_CTRL]              = 0.;
    dparm[DPARM_A_NORM]                         = -1.;
    dparm[DPARM_COMPRESS_TOLERANCE]             = 1e-8;
    dparm[DPARM_COMPRESS_MIN_RATIO]             = 1.;
}

int
pastixCheckParam( const pastix_int_t *iparm, const double *dparm )
{
    (void)dparm;
    if ( iparm[IPARM_FACTORIZATION] < PastixFactPOTRF || iparm[IPARM_FACTORIZATION] > PastixFactHETRF ) {
        pa_error( "factorization %ld is not a factorization kind (POTRF, SYTRF, GETRF, PXTRF, HETRF)",
                  (long)iparm[IPARM_FACTORIZATION] );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( iparm[IPARM_INCOMPLETE] != 0 ) {
        pa_error( "incomplete (ILU(k)) factorization is not available in this build" );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    if ( iparm[IPARM_MIN_BLOCKSIZE] < 1 || iparm[IPARM_MAX_BLOCKSIZE] < iparm[IPARM_MIN_BLOCKSIZE] ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( iparm[IPARM_REFINEMENT] < PastixRefineGMRES || iparm[IPARM_REFINEMENT] > PastixRefineBiCGSTAB ) {
        pa_error( "refinement %ld is not a refinement method (GMRES, CG, SR, BiCGSTAB)",
                  (long)iparm[IPARM_REFINEMENT] );
        return PASTIX_ERR_BADPARAMETER;
    }
    return PASTIX_SUCCESS;
}

static pastix_int_t pa_instance_counter = 0;

void
pastixInitWithAffinity( pastix_data_t **pastix_data, PASTIX_Comm comm, pastix_int_t *iparm,
                        double *dparm, const int *bindtab )
{
    pastix_data_t *pd = calloc( 1, sizeof( pastix_data_t ) );
    (void)bindtab;
    if ( iparm[IPARM_MODIFY_PARAMETER] == 0 ) {
        pastixInitParam( iparm, dparm );
    }
    pd->id          = pa_instance_counter++;
    pd->iparm       = iparm;
    pd->dparm       = dparm;
    pd->pastix_comm = comm;
    pd->steps       = STEP_INIT;
    if ( pa_dev_open( pd ) != PASTIX_SUCCESS ) {
        pa_error( "pastixInit: could not open the GPU" );
    }
    *pastix_data = pd;
}

void
pastixInit( pastix_data_t **pastix_data, PASTIX_Comm comm, pastix_int_t *iparm, double *dparm )
{
    pastixInitWithAffinity( pastix_data, comm, iparm, dparm, NULL );
}

static void
pa_clear_analysis( pastix_data_t *pd )
{
    if ( pd->ordemesh ) { pastixOrderExit( pd->ordemesh ); free( pd->ordemesh ); pd->ordemesh = NULL; }
    if ( pd->symbmtx )  { pa_symbol_exit( pd->symbmtx ); free( pd->symbmtx ); pd->symbmtx = NULL; }
    if ( pd->solvmatr ) { pa_solver_exit( pd->solvmatr ); free( pd->solvmatr ); pd->solvmatr = NULL; }
    if ( pd->graph )    { pa_graph_exit( pd->graph ); free( pd->graph ); pd->graph = NULL; }
}

void
pastixFinalize( pastix_data_t **pastix_data )
{
    pastix_data_t *pd;
    if ( pastix_data == NULL || *pastix_data == NULL ) return;
    pd = *pastix_data;
    pa_dev_close( pd );
    pa_clear_analysis( pd );
    if ( pd->bcsc ) { pa_bcsc_exit( pd->bcsc ); free( pd->bcsc ); }
    free( pd->legacy_spm );
    free( pd->b );
    free( pd->x0 );
    free( pd );
    *pastix_data = NULL;
}

static int
pa_check_spm( const spmatrix_t *spm )
{
    if ( spm == NULL || spm->n < 0 ) return PASTIX_ERR_BADPARAMETER;
    if ( spm->dof != 1 || spm->dofs != NULL ) {
        pa_error( "matrices with dof != 1 are not supported" );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    if ( spm->loc2glob != NULL ) {
        pa_error( "distributed matrices are not supported (single process build)" );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    if ( spm->fmttype != SpmCSC && spm->fmttype != SpmCSR ) {
        pa_error( "only CSC/CSR matrices are supported" );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    return PASTIX_SUCCESS;
}

/* Ordering step (reference order/pastix_subtask_order.c:97-470). */
int
pastix_subtask_order( pastix_data_t *pd, const spmatrix_t *spm, pastix_order_t *myorder )
{
    pastix_int_t  n;
    pastix_order_t *o;
    double t0 = pa_wtime();
    int rc;

    if ( pd == NULL || spm == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( !( pd->steps & STEP_INIT ) ) return PASTIX_ERR_BADPARAMETER;
    if ( ( rc = pa_check_spm( spm ) ) ) return rc;
    pd->iparm[IPARM_FLOAT] = spm->flttype;
    pd->iparm[IPARM_MTX_TYPE] = spm->mtxtype;
    pa_clear_analysis( pd );
    pd->csc = spm;
    n = spm->n;

    pd->graph = calloc( 1, sizeof( pa_graph_t ) );
    if ( ( rc = pa_graph_from_spm( pd->graph, spm ) ) ) return rc;

    if ( pd->iparm[IPARM_ORDERING] != PastixOrderPersonal ) {
        pa_error( "pastix_subtask_order: ordering %ld needs Scotch/Metis, which this build does not "
                  "provide; use PastixOrderPersonal (e.g. with pastixOrderGrid)",
                  (long)pd->iparm[IPARM_ORDERING] );
        return PASTIX_ERR_BADPARAMETER;
    }

    o = calloc( 1, sizeof( pastix_order_t ) );
    pd->ordemesh = o;
    if ( pd->iparm[IPARM_IO_STRATEGY] & PastixIOLoad ) {
        /* the ordering of a file (reference order/pastix_subtask_order.c:62) */
        if ( ( rc = pastixOrderLoad( pd, o ) ) ) return rc;
        pastixOrderBase( o, 0 );
        if ( o->vertnbr != n ) {
            pa_error( "pastix_subtask_order: the ordering file has %ld unknowns, the matrix %ld", (long)o->vertnbr,
                      (long)n );
            return PASTIX_ERR_BADPARAMETER;
        }
        goto ordered;
    }
    pastixOrderAlloc( o, n, 0 );
    if ( myorder != NULL ) {
        pastixOrderBase( myorder, 0 );
    }
    /* permutation (reference order/order_compute_personal.c) */
    if ( myorder == NULL || ( myorder->permtab == NULL && myorder->peritab == NULL ) ) {
        pastix_int_t i;
        for ( i = 0; i < n; i++ ) o->permtab[i] = o->peritab[i] = i;
    }
    else if ( myorder->permtab == NULL ) {
        pastix_int_t i;
        for ( i = 0; i < n; i++ ) o->permtab[myorder->peritab[i]] = i;
        memcpy( o->peritab, myorder->peritab, n * sizeof( pastix_int_t ) );
    }
    else if ( myorder->peritab == NULL ) {
        pastix_int_t i;
        for ( i = 0; i < n; i++ ) o->peritab[myorder->permtab[i]] = i;
        memcpy( o->permtab, myorder->permtab, n * sizeof( pastix_int_t ) );
    }
    else {
        memcpy( o->permtab, myorder->permtab, n * sizeof( pastix_int_t ) );
        memcpy( o->peritab, myorder->peritab, n * sizeof( pastix_int_t ) );
    }
    if ( myorder != NULL && myorder->rangtab != NULL && myorder->treetab != NULL ) {
        o->cblknbr = myorder->cblknbr;
        o->rangtab = malloc( ( o->cblknbr + 1 ) * sizeof( pastix_int_t ) );
        o->treetab = malloc( ( o->cblknbr + 1 ) * sizeof( pastix_int_t ) );
        memcpy( o->rangtab, myorder->rangtab, ( o->cblknbr + 1 ) * sizeof( pastix_int_t ) );
        memcpy( o->treetab, myorder->treetab, o->cblknbr * sizeof( pastix_int_t ) );
    }
ordered:
    if ( o->rangtab == NULL || o->treetab == NULL ) {
        /* a permutation without its supernode partition and tree: fundamental
         * supernodes of the postordered elimination tree, then amalgamation
         * (reference order/pastix_subtask_order.c:362-381) */
        if ( ( rc = pa_order_find_supernodes( o, n, pd->graph->colptr, pd->graph->rows ) ) ) return rc;
        if ( ( rc = pa_order_amalgamate( o, n, pd->graph->colptr, pd->graph->rows, 0,
                                         pd->iparm[IPARM_AMALGAMATION_LVLCBLK],
                                         pd->iparm[IPARM_AMALGAMATION_LVLBLAS] ) ) )
            return rc;
    }
    if ( pastixOrderCheck( o ) != PASTIX_SUCCESS ) {
        pa_error( "pastix_subtask_order: inconsistent ordering" );
        return PASTIX_ERR_BADPARAMETER;
    }
    pa_order_apply_level_order( o, pd->iparm[IPARM_TASKS2D_LEVEL], pd->iparm[IPARM_TASKS2D_WIDTH] );
    if ( pd->iparm[IPARM_IO_STRATEGY] & PastixIOSave ) {
        /* saved after the level ordering (reference pastix_subtask_order.c:485) */
        if ( ( rc = pastixOrderSave( pd, o ) ) ) return rc;
    }
    o->sndenbr = o->cblknbr;
    o->sndetab = malloc( ( o->cblknbr + 1 ) * sizeof( pastix_int_t ) );
    memcpy( o->sndetab, o->rangtab, ( o->cblknbr + 1 ) * sizeof( pastix_int_t ) );

    pd->dparm[DPARM_ORDER_TIME] = pa_wtime() - t0;
    pd->steps |= STEP_ORDERING;
    return PASTIX_SUCCESS;
}

/* Symbolic factorization (reference symbol/pastix_subtask_symbfact.c:86-330). */
int
pastix_subtask_symbfact( pastix_data_t *pd )
{
    double t0 = pa_wtime();
    int rc;
    if ( pd == NULL || !( pd->steps & STEP_ORDERING ) ) return PASTIX_ERR_BADPARAMETER;
    if ( pd->symbmtx ) { pa_symbol_exit( pd->symbmtx ); free( pd->symbmtx ); }
    pd->symbmtx = calloc( 1, sizeof( symbol_matrix_t ) );
    if ( ( rc = pa_symbol_fax( pd->symbmtx, pd->graph, pd->ordemesh ) ) ) return rc;
    pa_symbol_build_browtab( pd->symbmtx );
    if ( pa_symbol_check( pd->symbmtx ) ) {
        pa_error( "symbolic factorization produced an inconsistent structure" );
        return PASTIX_ERR_INTERNAL;
    }
    pd->iparm[IPARM_NNZEROS] = pa_symbol_nnz( pd->symbmtx );
    pd->dparm[DPARM_FACT_THFLOPS] = pa_symbol_flops( pd->symbmtx, (int)pd->iparm[IPARM_FACTORIZATION] );
    pd->dparm[DPARM_SYMBFACT_TIME] = pa_wtime() - t0;
    pd->steps |= STEP_SYMBFACT;
    return PASTIX_SUCCESS;
}

/* The reference reorders the rows inside each supernode to gather coupling
 * blocks (symbol/symbol_reordering.c); this build keeps the ordering as is. */
int
pastix_subtask_reordering( pastix_data_t *pd )
{
    if ( pd == NULL || !( pd->steps & STEP_SYMBFACT ) ) return PASTIX_ERR_BADPARAMETER;
    pd->dparm[DPARM_REORDER_TIME] = 0.;
    return PASTIX_SUCCESS;
}

/* Splitting and solver structure (reference blend/pastix_subtask_blend.c:121-600). */
int
pastix_subtask_blend( pastix_data_t *pd )
{
    double t0 = pa_wtime();
    int rc, flt;
    if ( pd == NULL || !( pd->steps & STEP_SYMBFACT ) ) return PASTIX_ERR_BADPARAMETER;
    if ( ( rc = pastixCheckParam( pd->iparm, pd->dparm ) ) ) return rc;
    if ( ( rc = pa_split_symbol( pd->symbmtx, pd->iparm[IPARM_MIN_BLOCKSIZE], pd->iparm[IPARM_MAX_BLOCKSIZE] ) ) )
        return rc;
    pa_symbol_build_browtab( pd->symbmtx );
    if ( pa_symbol_check( pd->symbmtx ) ) {
        pa_error( "splitting produced an inconsistent structure" );
        return PASTIX_ERR_INTERNAL;
    }
    pd->dparm[DPARM_FACT_THFLOPS] = pa_symbol_flops( pd->symbmtx, (int)pd->iparm[IPARM_FACTORIZATION] );
    pd->iparm[IPARM_NNZEROS]      = pa_symbol_nnz( pd->symbmtx );
    if ( pd->solvmatr ) { pa_solver_exit( pd->solvmatr ); free( pd->solvmatr ); }
    pd->solvmatr = calloc( 1, sizeof( SolverMatrix ) );
    flt = pd->iparm[IPARM_MIXED] ? PastixFloat : PastixDouble;
    if ( ( rc = pa_solver_gen( pd->solvmatr, pd->symbmtx, (int)pd->iparm[IPARM_FACTORIZATION], flt ) ) ) return rc;
    /* static schedule of the CPU factorization over IPARM_THREAD_NBR threads
     * (the host's online processors when it is not set): proportional
     * mapping + simulation, as the reference's blend */
    {
        pastix_int_t thr = pd->iparm[IPARM_THREAD_NBR];
        double       pred = 0.;
        if ( thr < 1 ) thr = (pastix_int_t)sysconf( _SC_NPROCESSORS_ONLN );
        if ( ( rc = pa_blend_static( pd->solvmatr, (int)pd->iparm[IPARM_FACTORIZATION], thr,
                                     (int)pd->iparm[IPARM_ALLCAND], &pred ) ) )
            return rc;
        pd->dparm[DPARM_PRED_FACT_TIME] = pred;
    }
    /* the analysis runs without a GPU; the structure is uploaded as soon as
     * one is available (here, or at the first numerical step) */
    if ( pd->dev != NULL ) {
        if ( ( rc = pa_dev_setup_structure( pd ) ) ) return rc;
    }
    pd->dparm[DPARM_BLEND_TIME] = pa_wtime() - t0;
    pd->steps |= STEP_ANALYSE;
    return PASTIX_SUCCESS;
}

int
pastix_task_analyze( pastix_data_t *pd, const spmatrix_t *spm )
{
    int rc;
    double t0 = pa_wtime();
    if ( pd == NULL || spm == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( !( pd->steps & STEP_ORDERING ) || pd->csc != spm ) {
        if ( ( rc = pastix_subtask_order( pd, spm, NULL ) ) ) return rc;
    }
    if ( ( rc = pastix_subtask_symbfact( pd ) ) ) return rc;
    if ( ( rc = pastix_subtask_reordering( pd ) ) ) return rc;
    if ( ( rc = pastix_subtask_blend( pd ) ) ) return rc;
    pd->dparm[DPARM_ANALYZE_TIME] = pa_wtime() - t0;
    return PASTIX_SUCCESS;
}

/* Permuted CSC, ||A||_F (reference sopalin/pastix_task_sopalin.c:93-210). */
int
pastix_subtask_spm2bcsc( pastix_data_t *pd, spmatrix_t *spm )
{
    int rc;
    if ( pd == NULL || spm == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( !( pd->steps & STEP_ANALYSE ) ) {
        pa_error( "pastix_subtask_spm2bcsc: the analysis has to be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( spm->n != pd->ordemesh->vertnbr ) return PASTIX_ERR_BADPARAMETER;
    if ( pd->dev == NULL ) {
        pa_error( "pastix_task_numfact: no GPU is available; this engine has no CPU path" );
        return PASTIX_ERR_INTERNAL;
    }
    if ( pa_dev_nlevels( pd ) == 0 && pd->solvmatr->cblknbr > 0 ) {
        if ( ( rc = pa_dev_setup_structure( pd ) ) ) return rc;
    }
    if ( pd->bcsc ) { pa_bcsc_exit( pd->bcsc ); free( pd->bcsc ); }
    pd->bcsc = calloc( 1, sizeof( pa_bcsc_t ) );
    if ( ( rc = pa_bcsc_init( pd->bcsc, spm, pd->ordemesh ) ) ) return rc;
    pd->csc = spm;
    pd->dparm[DPARM_A_NORM] = pd->bcsc->normF;
    if ( ( rc = pa_dev_upload_bcsc( pd ) ) ) return rc;
    pd->valmap_ready = 0;
    pd->steps &= ~( STEP_BCSC2CTAB | STEP_NUMFACT );
    pd->steps |= STEP_CSC2BCSC;
    return PASTIX_SUCCESS;
}

/* Device-resident counterpart of re-entering pastix_subtask_spm2bcsc with new
 * values on the same pattern (reference sopalin/pastix_task_sopalin.c:93-210):
 * the permuted CSC is gathered on the device and ||A||_F recomputed there. */
int
pastix_amd_values_dev( pastix_data_t *pd, const void *d_values )
{
    int rc;
    double normF = 0.;
    if ( pd == NULL || d_values == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( !( pd->steps & STEP_CSC2BCSC ) || pd->csc == NULL || pd->bcsc == NULL ) {
        pa_error( "pastix_amd_values_dev: pastix_task_numfact() or pastix_subtask_spm2bcsc() has to be called first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !pd->valmap_ready ) {
        /* the map is built once from the spm of the last spm2bcsc */
        pa_bcsc_t   tmp;
        pa_valmap_t map;
        memset( &tmp, 0, sizeof( tmp ) );
        if ( ( rc = pa_bcsc_build( &tmp, pd->csc, pd->ordemesh, &map ) ) ) return rc;
        if ( tmp.nnz != pd->bcsc->nnz ) {
            pa_bcsc_exit( &tmp );
            pa_valmap_exit( &map );
            pa_error( "pastix_amd_values_dev: the matrix given to spm2bcsc has changed" );
            return PASTIX_ERR_BADPARAMETER;
        }
        rc = pa_dev_upload_valmap( pd, map.Lptr, map.Lsrc, map.Uptr, map.Usrc, map.nnz );
        pa_bcsc_exit( &tmp );
        pa_valmap_exit( &map );
        if ( rc ) return rc;
        pd->valmap_ready = 1;
    }
    if ( ( rc = pa_dev_set_values( pd, (const double *)d_values, &normF ) ) ) return rc;
    pd->bcsc->normF         = normF;
    pd->dparm[DPARM_A_NORM] = normF;
    pd->steps &= ~( STEP_BCSC2CTAB | STEP_NUMFACT );
    return PASTIX_SUCCESS;
}

/* Coeftab allocation and fill in HBM (reference :213-410, sopalin/coeftab.c). */
int
pastix_subtask_bcsc2ctab( pastix_data_t *pd )
{
    int32_t overflow = 0;
    int prec, rc;
    if ( pd == NULL || !( pd->steps & STEP_CSC2BCSC ) ) return PASTIX_ERR_BADPARAMETER;
    if ( ( rc = pastixCheckParam( pd->iparm, pd->dparm ) ) ) return rc;
    prec = pd->iparm[IPARM_MIXED] ? PA_PREC_FLOAT : PA_PREC_DOUBLE;
    if ( pd->iparm[IPARM_FLOAT] == SpmFloat ) prec = PA_PREC_FLOAT;
    pd->solvmatr->flttype = ( prec == PA_PREC_FLOAT ) ? PastixFloat : PastixDouble;
    if ( pd->solvmatr->factotype != pd->iparm[IPARM_FACTORIZATION] ) {
        /* the factorization kind changed after the analysis: the flop count
         * of the new kind (the device schedule follows at the fill) */
        pd->solvmatr->factotype       = (int)pd->iparm[IPARM_FACTORIZATION];
        pd->dparm[DPARM_FACT_THFLOPS] = pa_symbol_flops( pd->symbmtx, (int)pd->iparm[IPARM_FACTORIZATION] );
    }
    /* a second factor array: U^T of LU, the D L^t copy of LDL^t / LDL^h
     * (real arithmetic: LL^h is LL^t and LDL^h is LDL^t) */
    rc = pa_dev_coeftab_init( pd, prec,
                              pd->iparm[IPARM_FACTORIZATION] == PastixFactLU ||
                                  pd->iparm[IPARM_FACTORIZATION] == PastixFactLDLT ||
                                  pd->iparm[IPARM_FACTORIZATION] == PastixFactLDLH,
                              &overflow );
    if ( rc ) return rc;
    if ( overflow ) {
        pa_warning( "%d coefficients overflow the single precision range", overflow );
    }
    pd->steps |= STEP_BCSC2CTAB;
    return PASTIX_SUCCESS;
}

/* Numerical factorization (reference :412-650). */
int
pastix_subtask_sopalin( pastix_data_t *pd )
{
    double threshold, seconds = 0.;
    int32_t npiv = 0, nneg = 0;
    int rc;
    if ( pd == NULL || !( pd->steps & STEP_ANALYSE ) || !( pd->steps & STEP_CSC2BCSC ) ||
         !( pd->steps & STEP_BCSC2CTAB ) ) {
        pa_error( "pastix_subtask_sopalin: all steps up to pastix_subtask_bcsc2ctab() must be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    /* static pivoting threshold: eps * ||A||_F with the FP64 epsilon, the
     * matrix being given in double (reference :465-494) */
    if ( pd->dparm[DPARM_EPSILON_MAGN_CTRL] < 0. ) {
        threshold = -pd->dparm[DPARM_EPSILON_MAGN_CTRL];
    }
    else if ( pd->dparm[DPARM_EPSILON_MAGN_CTRL] == 0. ) {
        double eps = ( pd->iparm[IPARM_FLOAT] == SpmFloat ) ? (double)FLT_EPSILON / 2. : DBL_EPSILON / 2.;
        threshold = eps * pd->dparm[DPARM_A_NORM];
    }
    else {
        threshold = pd->dparm[DPARM_EPSILON_MAGN_CTRL] * pd->dparm[DPARM_A_NORM];
    }
    pd->solvmatr->diagthreshold = threshold;
    if ( ( rc = pa_dev_factorize( pd, threshold, &npiv, &nneg, &seconds ) ) ) return rc;
    pd->solvmatr->nbpivots = npiv;
    if ( nneg > 0 ) {
        pa_error( "Negative diagonal term (%d occurrence(s)) in the Cholesky factorization", nneg );
    }
    pd->iparm[IPARM_STATIC_PIVOTING] = npiv;
    pd->dparm[DPARM_FACT_TIME]  = seconds;
    pd->dparm[DPARM_FACT_FLOPS] = seconds > 0. ? pd->dparm[DPARM_FACT_THFLOPS] / seconds / 1073741824. : 0.;
    /* performed flops (reference kernels/kernels_trace.c:349 accumulates the
     * kernels' own counts): the device kernels' work, GEMM padding included */
    pd->dparm[DPARM_FACT_RLFLOPS] = pa_dev_real_flops( pd );
    if ( pd->iparm[IPARM_VERBOSE] > PastixVerboseNot ) {
        fprintf( stdout, "    Time to factorize                      %e s\n", seconds );
        fprintf( stdout, "    Pivot thresholds: %d static pivot(s)\n", npiv );
    }
    pd->steps |= STEP_NUMFACT;
    return PASTIX_SUCCESS;
}

int
pastix_task_numfact( pastix_data_t *pd, spmatrix_t *spm )
{
    int rc;
    if ( pd == NULL || spm == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( !( pd->steps & STEP_ANALYSE ) ) {
        pa_error( "pastix_task_numfact: pastix_task_analyze() has to be called first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( ( rc = pastix_subtask_spm2bcsc( pd, spm ) ) ) return rc;
    if ( ( rc = pastix_subtask_bcsc2ctab( pd ) ) ) return rc;
    if ( ( rc = pastix_subtask_sopalin( pd ) ) ) return rc;
    return PASTIX_SUCCESS;
}

/* Autotuner of the launch schedule (no reference counterpart: it replaces
 * the StarPU/PaRSEC runtime's choices for one GPU).  Every applicable
 * candidate factors the current matrix twice (the second, graph-replayed
 * run is timed with device events); the fastest is kept and its factors are
 * left in place. */
int
pastix_amd_tune( pastix_data_t *pd, double *times, pastix_int_t *params )
{
    int c, best = -1, ncand = pa_dev_tune_ncand(), rc, rep;
    double tbest = 0.;
    if ( pd == NULL || pd->dev == NULL || !( pd->steps & STEP_CSC2BCSC ) ) {
        pa_error( "pastix_amd_tune: pastix_task_numfact() has to be called first" );
        return -PASTIX_ERR_BADPARAMETER;
    }
    for ( c = 0; c < ncand; c++ ) {
        double t = 0.;
        if ( times ) times[c] = 0.;
        if ( !pa_dev_tune_applicable( pd, c ) ) continue;
        if ( ( rc = pa_dev_set_tune( pd, c, NULL ) ) ) return -rc;
        for ( rep = 0; rep < 2; rep++ ) {
            if ( ( rc = pastix_subtask_bcsc2ctab( pd ) ) ) return -rc;
            if ( ( rc = pastix_subtask_sopalin( pd ) ) ) return -rc;
            t = pa_dev_last_fact_event_seconds( pd );
        }
        if ( times ) times[c] = t;
        if ( best < 0 || t < tbest ) {
            best  = c;
            tbest = t;
        }
    }
    if ( ( rc = pa_dev_set_tune( pd, best, NULL ) ) ) return -rc;
    if ( ( rc = pastix_subtask_bcsc2ctab( pd ) ) ) return -rc;
    if ( ( rc = pastix_subtask_sopalin( pd ) ) ) return -rc;
    if ( params ) {
        int32_t p[4];
        pa_dev_get_tune( pd, p );
        for ( c = 0; c < 4; c++ ) params[c] = p[c];
    }
    return best;
}

int
pastix_amd_set_tuning( pastix_data_t *pd, const pastix_int_t *params )
{
    int32_t p[4];
    int     c, rc;
    if ( pd == NULL || pd->dev == NULL || params == NULL ) return PASTIX_ERR_BADPARAMETER;
    for ( c = 0; c < 4; c++ ) p[c] = (int32_t)params[c];
    if ( ( rc = pa_dev_set_tune( pd, -1, p ) ) ) return rc;
    /* the factors of the previous schedule stay valid; a new factorization
     * uses the new task lists */
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* Right-hand sides (reference sopalin/pastix_rhs.c)                         */
/* ------------------------------------------------------------------------- */
int
pastixRhsInit( pastix_rhs_t *rhs )
{
    pastix_rhs_t B;
    if ( rhs == NULL ) return PASTIX_ERR_BADPARAMETER;
    B = calloc( 1, sizeof( struct pastix_rhs_s ) );
    if ( B == NULL ) return PASTIX_ERR_ALLOC;
    B->allocated = -1;
    B->flttype   = PastixPattern;
    B->m = B->n = B->ld = -1;
    *rhs = B;
    return PASTIX_SUCCESS;
}

int
pastixRhsFinalize( pastix_rhs_t rhs )
{
    if ( rhs == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( rhs->b != NULL ) {
        if ( rhs->allocated > 0 ) {
            free( rhs->b );
        }
        else {
            pa_warning( "pastixRhsFinalize called before the ordering of b was restored "
                        "(pastix_subtask_applyorder with PastixDirBackward)" );
        }
    }
    free( rhs );
    return PASTIX_SUCCESS;
}

static size_t
pa_flt_size( int flttype )
{
    return ( flttype == PastixFloat ) ? sizeof( float ) : sizeof( double );
}

/* FP64 -> FP32 copy of an RHS (reference pastixRhsDoubletoSingle,
 * sopalin/pastix_rhs.c:162-236): an entry outside the single precision range
 * is an error, as LAPACK dlag2s reports it. */
int
pastixRhsDoubletoSingle( const pastix_rhs_t dB, pastix_rhs_t sB )
{
    pastix_int_t i, j;
    int          ovf = 0, tofree = 0;
    if ( dB == NULL || sB == NULL || dB->flttype != PastixDouble || dB->b == NULL ) {
        pa_error( "pastixRhsDoubletoSingle: invalid float type for mixed precision" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( sB->allocated == -1 ) {
        *sB           = *dB;
        sB->allocated = 1;
        sB->flttype   = PastixFloat;
        sB->b         = malloc( (size_t)dB->ld * dB->n * sizeof( float ) );
        if ( sB->b == NULL ) return PASTIX_ERR_ALLOC;
        tofree = 1;
    }
    if ( sB->flttype != PastixFloat || sB->b == NULL || sB->m != dB->m || sB->n != dB->n || sB->ld < sB->m )
        return PASTIX_ERR_BADPARAMETER;
    for ( j = 0; j < dB->n; j++ ) {
        const double *src = (const double *)dB->b + j * dB->ld;
        float        *dst = (float *)sB->b + j * sB->ld;
        for ( i = 0; i < dB->m; i++ ) {
            if ( fabs( src[i] ) > (double)FLT_MAX ) ovf = 1;
            dst[i] = (float)src[i];
        }
    }
    if ( ovf ) {
        if ( tofree ) {
            free( sB->b );
            sB->b         = NULL;
            sB->allocated = -1;
        }
        return PASTIX_ERR_INTERNAL;
    }
    return PASTIX_SUCCESS;
}

/* FP32 -> FP64 copy of an RHS (reference pastixRhsSingleToDouble,
 * sopalin/pastix_rhs.c:242-316) */
int
pastixRhsSingleToDouble( const pastix_rhs_t sB, pastix_rhs_t dB )
{
    pastix_int_t i, j;
    if ( sB == NULL || dB == NULL || sB->flttype != PastixFloat || sB->b == NULL ) {
        pa_error( "pastixRhsSingleToDouble: invalid float type for mixed precision" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( dB->allocated == -1 ) {
        *dB           = *sB;
        dB->allocated = 1;
        dB->flttype   = PastixDouble;
        dB->b         = malloc( (size_t)sB->ld * sB->n * sizeof( double ) );
        if ( dB->b == NULL ) return PASTIX_ERR_ALLOC;
    }
    if ( dB->flttype != PastixDouble || dB->b == NULL || dB->m != sB->m || dB->n != sB->n || dB->ld < dB->m )
        return PASTIX_ERR_BADPARAMETER;
    for ( j = 0; j < sB->n; j++ ) {
        const float *src = (const float *)sB->b + j * sB->ld;
        double      *dst = (double *)dB->b + j * dB->ld;
        for ( i = 0; i < sB->m; i++ ) dst[i] = (double)src[i];
    }
    return PASTIX_SUCCESS;
}

/* Host RHS of the caller's arithmetic (FP64, or FP32 for a single precision
 * matrix) <-> device FP64 work array with leading dimension m. */
static int
pa_rhs_upload( pastix_data_t *pd, const void *B, int flttype, pastix_int_t m, pastix_int_t nrhs, pastix_int_t ldb,
               double *dB )
{
    pastix_int_t i, j;
    int rc = PASTIX_SUCCESS;
    if ( flttype != PastixFloat ) {
        for ( j = 0; j < nrhs && !rc; j++ )
            rc = pa_dev_memcpy_h2d( pd, dB + j * m, (const double *)B + j * ldb, m * sizeof( double ) );
        return rc;
    }
    {
        double *tmp = malloc( (size_t)pa_imax( m, 1 ) * sizeof( double ) );
        if ( tmp == NULL ) return PASTIX_ERR_ALLOC;
        for ( j = 0; j < nrhs && !rc; j++ ) {
            const float *src = (const float *)B + j * ldb;
            for ( i = 0; i < m; i++ ) tmp[i] = (double)src[i];
            rc = pa_dev_memcpy_h2d( pd, dB + j * m, tmp, m * sizeof( double ) );
        }
        free( tmp );
    }
    return rc;
}

static int
pa_rhs_download( pastix_data_t *pd, void *B, int flttype, pastix_int_t m, pastix_int_t nrhs, pastix_int_t ldb,
                 const double *dB )
{
    pastix_int_t i, j;
    int rc = PASTIX_SUCCESS;
    if ( flttype != PastixFloat ) {
        for ( j = 0; j < nrhs && !rc; j++ )
            rc = pa_dev_memcpy_d2h( pd, (double *)B + j * ldb, dB + j * m, m * sizeof( double ) );
        return rc;
    }
    {
        double *tmp = malloc( (size_t)pa_imax( m, 1 ) * sizeof( double ) );
        if ( tmp == NULL ) return PASTIX_ERR_ALLOC;
        for ( j = 0; j < nrhs && !rc; j++ ) {
            float *dst = (float *)B + j * ldb;
            rc = pa_dev_memcpy_d2h( pd, tmp, dB + j * m, m * sizeof( double ) );
            for ( i = 0; !rc && i < m; i++ ) dst[i] = (float)tmp[i];
        }
        free( tmp );
    }
    return rc;
}

/* arithmetic of the caller's vectors: the one of the matrix given to the
 * analysis (reference sopalin/pastix_task_solve.c:200, csc->flttype) */
static int
pa_user_flttype( const pastix_data_t *pd )
{
    return ( pd->iparm[IPARM_FLOAT] == SpmFloat ) ? PastixFloat : PastixDouble;
}

/* In-place device solve on an HBM buffer in the user numbering. */
int
pastix_amd_solve_dev( pastix_data_t *pd, pastix_trans_t transA, pastix_int_t nrhs, void *dB, pastix_int_t ldb )
{
    double *tmp;
    pastix_int_t n;
    int rc;
    double t0;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_task_solve: all steps up to pastix_task_numfact() must be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    n = pd->solvmatr->nodenbr;
    if ( nrhs < 1 || ldb < n || dB == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( transA != PastixNoTrans && transA != PastixTrans ) return PASTIX_ERR_BADPARAMETER;
    t0  = pa_wtime();
    tmp = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !tmp ) return PASTIX_ERR_OUTOFMEMORY;
    rc = pa_dev_apply_perm( pd, 1, (int32_t)nrhs, dB, ldb, tmp, n );
    if ( !rc ) rc = pa_dev_solve_perm( pd, transA, (int32_t)nrhs, tmp, n );
    if ( !rc ) rc = pa_dev_apply_perm( pd, 0, (int32_t)nrhs, tmp, n, dB, ldb );
    if ( !rc ) rc = pa_dev_sync( pd );
    pa_dev_wput( pd, tmp );
    pd->dparm[DPARM_SOLV_TIME] = pa_wtime() - t0;
    if ( !rc ) pd->steps |= STEP_SOLVE;
    return rc;
}

/* pastix_amd_solve_dev with no device allocation: the device layer's static
 * scratch array holds the permuted right-hand sides (no counterpart in the
 * reference; for callers that cannot allocate). */
int
pastix_amd_solve_dev_noalloc( pastix_data_t *pd, pastix_trans_t transA, pastix_int_t nrhs, void *dB,
                              pastix_int_t ldb )
{
    pastix_int_t n;
    int rc;
    double t0;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_amd_solve_dev_noalloc: all steps up to pastix_task_numfact() must be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    n = pd->solvmatr->nodenbr;
    if ( nrhs < 1 || ldb < n || dB == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( transA != PastixNoTrans && transA != PastixTrans ) return PASTIX_ERR_BADPARAMETER;
    if ( (int64_t)n * nrhs > pa_dev_noalloc_max( pd ) ) {
        pa_error( "pastix_amd_solve_dev_noalloc: n * nrhs = %ld exceeds the scratch limit %ld", (long)( n * nrhs ),
                  (long)pa_dev_noalloc_max( pd ) );
        return PASTIX_ERR_BADPARAMETER;
    }
    t0 = pa_wtime();
    rc = pa_dev_solve_noalloc( pd, transA, (int32_t)nrhs, dB, ldb );
    pd->dparm[DPARM_SOLV_TIME] = pa_wtime() - t0;
    if ( !rc ) pd->steps |= STEP_SOLVE;
    return rc;
}

/* Host-buffer solve (reference sopalin/pastix_task_solve.c:758): B is
 * overwritten by the solution; data crosses PCIe both ways.  The
 * permutations P b and P^T x run on the device. */
int
pastix_task_solve( pastix_data_t *pd, pastix_int_t m, pastix_int_t nrhs, void *B, pastix_int_t ldb )
{
    double *dB;
    pastix_int_t n;
    int rc, flt;
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "pastix_task_solve: all steps up to pastix_task_numfact() must be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    n = pd->solvmatr->nodenbr;
    if ( m != n || ldb < n || B == NULL || nrhs < 1 ) return PASTIX_ERR_BADPARAMETER;
    flt = pa_user_flttype( pd );
    dB  = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !dB ) return PASTIX_ERR_OUTOFMEMORY;
    rc = pa_rhs_upload( pd, B, flt, n, nrhs, ldb, dB );
    if ( !rc ) rc = pastix_amd_solve_dev( pd, (pastix_trans_t)pd->iparm[IPARM_TRANSPOSE_SOLVE], nrhs, dB, n );
    if ( !rc ) rc = pa_rhs_download( pd, B, flt, n, nrhs, ldb, dB );
    pa_dev_wput( pd, dB );
    return rc;
}

/* Permutation of the right-hand sides (reference
 * pastix_subtask_applyorder, sopalin/pastix_task_solve.c:169-290, and
 * bvec_zlapmr's shared-memory case, bcsc/bvec_zlapmr.c:816-910): Forward
 * permutes b in place (b[i] <- b[peritab[i]]) and records it in Bp; Backward
 * restores the user numbering in place and resets Bp. */
int
pastix_subtask_applyorder( pastix_data_t *pd, pastix_dir_t dir, pastix_int_t m, pastix_int_t n, void *b,
                           pastix_int_t ldb, pastix_rhs_t Bp )
{
    const pastix_order_t *o;
    pastix_int_t i, j, N;
    int          flt;
    size_t       sz;
    char        *tmp;
    if ( pd == NULL ) {
        pa_error( "pastix_subtask_applyorder: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( Bp == NULL ) {
        pa_error( "pastix_subtask_applyorder: wrong Bp parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    o = pd->ordemesh;
    if ( o == NULL || !( pd->steps & STEP_ORDERING ) ) {
        pa_error( "pastix_subtask_applyorder: the ordering step has to be done first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( o->baseval != 0 ) {
        pa_error( "pastix_subtask_applyorder: ordermesh must be 0-based" );
        return PASTIX_ERR_BADPARAMETER;
    }
    N   = o->vertnbr;
    flt = pa_user_flttype( pd );
    if ( m != N || ldb < m || n < 0 || ( b == NULL && n > 0 ) ) return PASTIX_ERR_BADPARAMETER;
    if ( dir == PastixDirForward ) {
        Bp->flttype   = flt;
        Bp->m         = N;
        Bp->n         = n;
        Bp->allocated = 0;
        Bp->ld        = ldb;
        Bp->b         = b;
    }
    else if ( dir == PastixDirBackward ) {
        if ( Bp->allocated != 0 || Bp->b != b || Bp->m != N || Bp->n != n || Bp->ld != ldb || Bp->flttype != flt ) {
            pa_error( "pastix_subtask_applyorder: Bp does not describe b" );
            return PASTIX_ERR_BADPARAMETER;
        }
    }
    else {
        return PASTIX_ERR_BADPARAMETER;
    }
    sz  = pa_flt_size( flt );
    tmp = malloc( (size_t)pa_imax( N, 1 ) * sz );
    if ( tmp == NULL ) return PASTIX_ERR_ALLOC;
    for ( j = 0; j < n; j++ ) {
        char *col = (char *)b + (size_t)j * ldb * sz;
        if ( dir == PastixDirForward )
            for ( i = 0; i < N; i++ ) memcpy( tmp + i * sz, col + o->peritab[i] * sz, sz );
        else
            for ( i = 0; i < N; i++ ) memcpy( tmp + i * sz, col + o->permtab[i] * sz, sz );
        memcpy( col, tmp, (size_t)N * sz );
    }
    free( tmp );
    if ( dir == PastixDirBackward ) {
        Bp->allocated = -1;
        Bp->flttype   = PastixPattern;
        Bp->m = Bp->n = Bp->ld = -1;
        Bp->b         = NULL;
    }
    return PASTIX_SUCCESS;
}

/* checks shared by the subtasks that work on a permuted RHS */
static int
pa_check_rhs( pastix_data_t *pd, const pastix_rhs_t Bp, const char *who )
{
    if ( pd == NULL ) {
        pa_error( "%s: wrong pastix_data parameter", who );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( Bp == NULL || Bp->b == NULL ) {
        pa_error( "%s: wrong Bp parameter", who );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        pa_error( "%s: all steps up to pastix_task_numfact() have to be done first", who );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( Bp->m != pd->solvmatr->nodenbr || Bp->ld < Bp->m || Bp->n < 0 ||
         ( Bp->flttype != PastixDouble && Bp->flttype != PastixFloat ) ) {
        pa_error( "%s: Bp does not match the factorized matrix", who );
        return PASTIX_ERR_BADPARAMETER;
    }
    return PASTIX_SUCCESS;
}

/* the device steps on a permuted host RHS: upload, run, download */
typedef int ( *pa_perm_op_t )( pastix_data_t *, void *, int32_t, double *, int64_t );

static int
pa_rhs_device_op( pastix_data_t *pd, pastix_rhs_t Bp, pa_perm_op_t op, void *arg )
{
    const pastix_int_t n = Bp->m, nrhs = Bp->n;
    double *dB;
    int rc;
    if ( nrhs == 0 ) return PASTIX_SUCCESS;
    dB = pa_dev_wget( pd, (size_t)n * nrhs * sizeof( double ) );
    if ( !dB ) return PASTIX_ERR_OUTOFMEMORY;
    rc = pa_rhs_upload( pd, Bp->b, Bp->flttype, n, nrhs, Bp->ld, dB );
    if ( !rc ) rc = op( pd, arg, (int32_t)nrhs, dB, n );
    if ( !rc ) rc = pa_dev_sync( pd );
    if ( !rc ) rc = pa_rhs_download( pd, Bp->b, Bp->flttype, n, nrhs, Bp->ld, dB );
    pa_dev_wput( pd, dB );
    return rc;
}

typedef struct { int upper, transposed, unit; } pa_trsm_arg_t;

static int
pa_op_trsm( pastix_data_t *pd, void *arg, int32_t nrhs, double *dB, int64_t ld )
{
    const pa_trsm_arg_t *a = (const pa_trsm_arg_t *)arg;
    return pa_dev_trsm_perm( pd, a->upper, a->transposed, a->unit, nrhs, dB, ld );
}

static int
pa_op_diag( pastix_data_t *pd, void *arg, int32_t nrhs, double *dB, int64_t ld )
{
    (void)arg;
    return pa_dev_diag_perm( pd, nrhs, dB, ld );
}

static int
pa_op_solve( pastix_data_t *pd, void *arg, int32_t nrhs, double *dB, int64_t ld )
{
    return pa_dev_solve_perm( pd, *(const pastix_trans_t *)arg, nrhs, dB, ld );
}

/* Triangular solve on a permuted RHS (reference pastix_subtask_trsm,
 * sopalin/pastix_task_solve.c:296-400).  Left side only: the reference does
 * not handle the right side either (kernels/solve_ztrsmsp.c:326-340).  The
 * Upper cases sweep the stored U^T and need an LU factorization. */
int
pastix_subtask_trsm( pastix_data_t *pd, pastix_side_t side, pastix_uplo_t uplo, pastix_trans_t trans,
                     pastix_diag_t diag, pastix_rhs_t Bp )
{
    pa_trsm_arg_t a;
    int rc;
    /* the instance and the handle must exist (reference :310-318); the side
     * is checked before the handle's contents */
    if ( pd == NULL || Bp == NULL ) {
        pa_error( "pastix_subtask_trsm: wrong pastix_data or Bp parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( side != PastixLeft ) {
        pa_error( "pastix_subtask_trsm: only the left side is handled" );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    if ( ( rc = pa_check_rhs( pd, Bp, "pastix_subtask_trsm" ) ) ) return rc;
    if ( ( uplo != PastixLower && uplo != PastixUpper ) || ( diag != PastixUnit && diag != PastixNonUnit ) ||
         ( trans != PastixNoTrans && trans != PastixTrans && trans != PastixConjTrans ) )
        return PASTIX_ERR_BADPARAMETER;
    if ( uplo == PastixUpper && pd->iparm[IPARM_FACTORIZATION] != PastixFactLU ) {
        pa_error( "pastix_subtask_trsm: the upper factor is only stored by an LU factorization" );
        return PASTIX_ERR_BADPARAMETER;
    }
    a.upper      = ( uplo == PastixUpper );
    a.transposed = ( trans != PastixNoTrans ); /* real arithmetic: ConjTrans == Trans */
    a.unit       = ( diag == PastixUnit );
    return pa_rhs_device_op( pd, Bp, pa_op_trsm, &a );
}

/* y <- D^{-1} y (reference pastix_subtask_diag, sopalin/pastix_task_solve.c:422-470) */
int
pastix_subtask_diag( pastix_data_t *pd, pastix_rhs_t Bp )
{
    int rc;
    if ( ( rc = pa_check_rhs( pd, Bp, "pastix_subtask_diag" ) ) ) return rc;
    return pa_rhs_device_op( pd, Bp, pa_op_diag, NULL );
}

/* Solve on an RHS already in the factorization numbering (reference
 * sopalin/pastix_task_solve.c:512-707).  The FP32 sweeps of a mixed
 * precision factorization run on an FP32 copy made on the device. */
int
pastix_subtask_solve_adv( pastix_data_t *pd, pastix_trans_t transA, pastix_rhs_t Bp )
{
    double t0;
    int    rc;
    if ( ( rc = pa_check_rhs( pd, Bp, "pastix_subtask_solve_adv" ) ) ) return rc;
    if ( transA == PastixConjTrans ) transA = PastixTrans; /* real arithmetic */
    if ( transA != PastixNoTrans && transA != PastixTrans ) return PASTIX_ERR_BADPARAMETER;
    t0 = pa_wtime();
    rc = pa_rhs_device_op( pd, Bp, pa_op_solve, &transA );
    pd->dparm[DPARM_SOLV_TIME] = pa_wtime() - t0;
    if ( !rc ) pd->steps |= STEP_SOLVE;
    return rc;
}

int
pastix_subtask_solve( pastix_data_t *pd, pastix_rhs_t Bp )
{
    if ( pd == NULL ) return PASTIX_ERR_BADPARAMETER;
    return pastix_subtask_solve_adv( pd, (pastix_trans_t)pd->iparm[IPARM_TRANSPOSE_SOLVE], Bp );
}

int
pastix_amd_profile( pastix_data_t *pd, int enable, double *out )
{
    if ( pd == NULL || pd->dev == NULL ) return PASTIX_ERR_BADPARAMETER;
    return pa_dev_profile( pd, enable, out );
}

double
pastix_amd_last_fact_seconds( pastix_data_t *pd )
{
    return pd ? pa_dev_last_fact_event_seconds( pd ) : 0.;
}

int
pastix_amd_get_dims( const pastix_data_t *pd, pastix_int_t *dims )
{
    if ( pd == NULL || pd->solvmatr == NULL || dims == NULL ) return PASTIX_ERR_BADPARAMETER;
    dims[0] = pd->solvmatr->nodenbr;
    dims[1] = pd->solvmatr->cblknbr;
    dims[2] = pd->solvmatr->bloknbr;
    dims[3] = pd->solvmatr->coefnbr;
    dims[4] = pd->solvmatr->brownbr;
    dims[5] = pa_dev_nlevels( pd );
    dims[6] = pd->iparm[IPARM_FACTORIZATION];
    dims[7] = pd->solvmatr->flttype;
    return PASTIX_SUCCESS;
}

const SolverMatrix *
pastix_amd_get_solver( const pastix_data_t *pd )
{
    return pd ? pd->solvmatr : NULL;
}

int
pastix_amd_get_schedule( const pastix_data_t *pd, pastix_int_t *thrdnbr, pastix_int_t *tptr, pastix_int_t *ttsk )
{
    const SolverMatrix *s;
    pastix_int_t        t, i, n = 0;
    if ( pd == NULL || pd->solvmatr == NULL || thrdnbr == NULL ) return PASTIX_ERR_BADPARAMETER;
    s        = pd->solvmatr;
    *thrdnbr = s->thrdnbr;
    if ( tptr == NULL ) return PASTIX_SUCCESS;
    tptr[0] = 0;
    for ( t = 0; t < s->thrdnbr; t++ ) {
        for ( i = 0; i < s->ttsknbr[t]; i++ ) {
            if ( ttsk ) ttsk[n] = s->tasktab[s->ttsktab[t][i]].cblknum;
            n++;
        }
        tptr[t + 1] = n;
    }
    return PASTIX_SUCCESS;
}

int
pastix_amd_get_structure( const pastix_data_t *pd, pastix_int_t *cblk, pastix_int_t *blok,
                          pastix_int_t *perm, pastix_int_t *invp )
{
    const SolverMatrix *s;
    pastix_int_t k, b, off = 0;
    if ( pd == NULL || pd->solvmatr == NULL ) return PASTIX_ERR_BADPARAMETER;
    s = pd->solvmatr;
    for ( k = 0; cblk && k < s->cblknbr; k++ ) {
        const SolverCblk *c = s->cblktab + k;
        cblk[5 * k + 0] = c->fcolnum;
        cblk[5 * k + 1] = c->lcolnum;
        cblk[5 * k + 2] = c->fblokptr - s->bloktab;
        cblk[5 * k + 3] = c->stride;
        cblk[5 * k + 4] = off;
        off += c->stride * cblk_colnbr( c );
    }
    for ( b = 0; blok && b < s->bloknbr; b++ ) {
        const SolverBlok *bl = s->bloktab + b;
        blok[5 * b + 0] = bl->frownum;
        blok[5 * b + 1] = bl->lrownum;
        blok[5 * b + 2] = bl->lcblknm;
        blok[5 * b + 3] = bl->fcblknm;
        blok[5 * b + 4] = bl->coefind;
    }
    if ( perm ) memcpy( perm, pd->ordemesh->permtab, s->nodenbr * sizeof( pastix_int_t ) );
    if ( invp ) memcpy( invp, pd->ordemesh->peritab, s->nodenbr * sizeof( pastix_int_t ) );
    return PASTIX_SUCCESS;
}

int
pastix_amd_get_coeftab( const pastix_data_t *pd, int side, double *host )
{
    if ( pd == NULL || !( pd->steps & STEP_BCSC2CTAB ) ) return PASTIX_ERR_BADPARAMETER;
    return pa_dev_get_coeftab( pd, side, host );
}

pastix_order_t *
pastixOrderGet( const pastix_data_t *pd )
{
    return pd ? pd->ordemesh : NULL;
}

void
pastixDumpParam( const pastix_data_t *pd )
{
    int i;
    if ( !pd ) return;
    for ( i = 0; i < IPARM_SIZE; i++ ) fprintf( stdout, "iparm[%d],%ld\n", i, (long)pd->iparm[i] );
    for ( i = 0; i < DPARM_SIZE; i++ ) fprintf( stdout, "dparm[%d],%e\n", i, pd->dparm[i] );
}

/* Legacy one-call interface (reference sopalin/pastix.c:103-487): runs the
 * steps IPARM_START_TASK..IPARM_END_TASK on a CSC given by arrays, advancing
 * IPARM_START_TASK as each step completes.  The spm describing the caller's
 * arrays belongs to the instance (no state shared between instances, as the
 * reference's example/reentrant.c requires); b is kept between the solve
 * and the refinement steps so that the refinement sees the original right
 * hand side. */
int
pastix( pastix_data_t **pastix_data, PASTIX_Comm comm, pastix_int_t n, pastix_int_t *colptr,
        pastix_int_t *rowptr, void *values, pastix_int_t *perm, pastix_int_t *invp, void *B,
        pastix_int_t nrhs, pastix_int_t *iparm, double *dparm )
{
    pastix_data_t *pd;
    spmatrix_t    *spm;
    size_t         size;
    int            rc;

    if ( iparm[IPARM_MODIFY_PARAMETER] == 0 ) {
        pastixInitParam( iparm, dparm );
        iparm[IPARM_MODIFY_PARAMETER] = 1;
        return PASTIX_SUCCESS;
    }
    if ( iparm[IPARM_END_TASK] < PastixTaskInit ) return PASTIX_SUCCESS;
    if ( iparm[IPARM_START_TASK] == PastixTaskInit ) {
        if ( *pastix_data != NULL ) pastixFinalize( pastix_data ); /* restart with new parameters */
        pastixInit( pastix_data, comm, iparm, dparm );
        iparm[IPARM_START_TASK]++;
    }
    if ( iparm[IPARM_END_TASK] < PastixTaskOrdering ) return PASTIX_SUCCESS;
    pd = *pastix_data;
    if ( pd == NULL ) {
        pa_error( "pastix: the initialization step was not run before the following steps" );
        return PASTIX_ERR_BADPARAMETER;
    }

    /* the spm of the caller's arrays (rebuilt when they change) */
    if ( iparm[IPARM_START_TASK] <= PastixTaskNumfact ) {
        spm = pd->legacy_spm;
        if ( spm != NULL && ( spm->n != n || spm->nnz != colptr[n] - colptr[0] || spm->colptr != colptr ||
                              spm->rowptr != rowptr ) ) {
            if ( pd->csc == spm ) pd->csc = NULL;
            free( spm );
            spm = pd->legacy_spm = NULL;
        }
        if ( spm == NULL ) {
            if ( iparm[IPARM_FLOAT] == -1 ) {
                pa_error( "pastix: iparm[IPARM_FLOAT] has to be set" );
                return PASTIX_ERR_BADPARAMETER;
            }
            if ( iparm[IPARM_MTX_TYPE] == -1 ) {
                pa_error( "pastix: iparm[IPARM_MTX_TYPE] has to be set" );
                return PASTIX_ERR_BADPARAMETER;
            }
            if ( iparm[IPARM_DOF_NBR] < 1 ) {
                pa_error( "pastix: variable dofs are not supported by the legacy interface" );
                return PASTIX_ERR_BADPARAMETER;
            }
            spm = calloc( 1, sizeof( spmatrix_t ) );
            if ( spm == NULL ) return PASTIX_ERR_ALLOC;
            spm->mtxtype  = (spm_mtxtype_t)iparm[IPARM_MTX_TYPE];
            spm->flttype  = (spm_coeftype_t)iparm[IPARM_FLOAT];
            spm->fmttype  = SpmCSC;
            spm->baseval  = colptr[0];
            spm->n = spm->gN = spm->nexp = spm->gNexp = n;
            spm->nnz = spm->gnnz = spm->nnzexp = spm->gnnzexp = colptr[n] - colptr[0];
            spm->dof      = iparm[IPARM_DOF_NBR];
            spm->layout   = SpmColMajor;
            spm->colptr   = colptr;
            spm->rowptr   = rowptr;
            spm->values   = values;
            spm->clustnbr = 1;
            pd->legacy_spm = spm;
        }
        /* values given only at the numerical steps */
        if ( spm->values == NULL ) spm->values = values;
    }
    else {
        spm = pd->legacy_spm;
    }
    if ( spm == NULL ) return PASTIX_ERR_BADPARAMETER;
    size = ( spm->flttype == SpmFloat ? sizeof( float ) : sizeof( double ) ) * (size_t)spm->n * ( nrhs > 0 ? nrhs : 1 );

    if ( iparm[IPARM_START_TASK] == PastixTaskOrdering ) {
        pastix_order_t o, *po = NULL;
        if ( perm != NULL || invp != NULL ) {
            memset( &o, 0, sizeof( o ) );
            o.vertnbr = n;
            o.permtab = perm;
            o.peritab = invp;
            po        = &o;
        }
        if ( ( rc = pastix_subtask_order( pd, spm, po ) ) ) return rc;
        if ( perm ) memcpy( perm, pd->ordemesh->permtab, n * sizeof( pastix_int_t ) );
        if ( invp ) memcpy( invp, pd->ordemesh->peritab, n * sizeof( pastix_int_t ) );
        iparm[IPARM_START_TASK]++;
    }
    if ( iparm[IPARM_END_TASK] < PastixTaskSymbfact ) return PASTIX_SUCCESS;
    if ( iparm[IPARM_START_TASK] == PastixTaskSymbfact ) {
        if ( ( rc = pastix_subtask_symbfact( pd ) ) ) return rc;
        iparm[IPARM_START_TASK]++;
    }
    if ( iparm[IPARM_END_TASK] < PastixTaskAnalyze ) return PASTIX_SUCCESS;
    if ( iparm[IPARM_START_TASK] == PastixTaskAnalyze ) {
        if ( ( rc = pastix_subtask_reordering( pd ) ) ) return rc;
        if ( ( rc = pastix_subtask_blend( pd ) ) ) return rc;
        iparm[IPARM_START_TASK]++;
    }
    if ( iparm[IPARM_END_TASK] < PastixTaskNumfact ) return PASTIX_SUCCESS;
    if ( iparm[IPARM_START_TASK] == PastixTaskNumfact ) {
        if ( ( rc = pastix_task_numfact( pd, spm ) ) ) return rc;
        iparm[IPARM_START_TASK]++;
    }
    if ( iparm[IPARM_END_TASK] < PastixTaskSolve ) return PASTIX_SUCCESS;
    if ( iparm[IPARM_START_TASK] == PastixTaskSolve ) {
        free( pd->x0 ); pd->x0 = NULL;
        free( pd->b );  pd->b  = NULL;
        /* keep b for a refinement in this call */
        if ( iparm[IPARM_END_TASK] > PastixTaskSolve ) {
            if ( ( pd->b = malloc( size ) ) == NULL ) return PASTIX_ERR_ALLOC;
            memcpy( pd->b, B, size );
        }
        if ( ( rc = pastix_task_solve( pd, spm->n, nrhs, B, spm->n ) ) ) return rc;
        iparm[IPARM_START_TASK]++;
        /* keep x0 for a refinement in a later call */
        if ( iparm[IPARM_END_TASK] == PastixTaskSolve ) {
            if ( ( pd->x0 = malloc( size ) ) == NULL ) return PASTIX_ERR_ALLOC;
            memcpy( pd->x0, B, size );
        }
    }
    if ( iparm[IPARM_END_TASK] < PastixTaskRefine ) return PASTIX_SUCCESS;
    if ( iparm[IPARM_START_TASK] == PastixTaskRefine ) {
        void *rb = pd->b, *rx = pd->x0;
        if ( rb == NULL && rx == NULL ) {
            pa_error( "pastix: neither b nor x0 was kept for the refinement" );
            return PASTIX_ERR_UNKNOWN;
        }
        if ( rb != NULL && rx != NULL ) {
            pa_error( "pastix: both b and x0 were kept, which cannot happen" );
            return PASTIX_ERR_UNKNOWN;
        }
        /* x0 kept from an earlier call: B holds the original b; b kept in
         * this call: B holds x0 */
        if ( rb == NULL ) rb = B;
        else              rx = B;
        if ( ( rc = pastix_task_refine( pd, spm->n, nrhs, rb, spm->n, rx, spm->n ) ) ) return rc;
        iparm[IPARM_START_TASK]++;
        if ( B != rx ) memcpy( B, rx, size );
    }
    free( pd->x0 ); pd->x0 = NULL;
    free( pd->b );  pd->b  = NULL;
    if ( iparm[IPARM_END_TASK] < PastixTaskClean ) return PASTIX_SUCCESS;
    if ( iparm[IPARM_START_TASK] == PastixTaskClean ) {
        pastixFinalize( pastix_data );
        iparm[IPARM_START_TASK]++;
    }
    return PASTIX_SUCCESS;
}
