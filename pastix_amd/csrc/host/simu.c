/*
 * simu.c -- static scheduling of the CPU factorization: the analysis step
 * that gives each of thrdnbr threads the ordered list of the cblk tasks it
 * runs (SolverMatrix tasktab / ttsktab), as the reference's blend does for
 * its static scheduler (sopalin/sequential_zpotrf.c:104-145 walks these
 * lists).  The GPU engine does not use these lists: they describe the CPU
 * baseline's schedule and are part of the saved structure.
 *
 * Steps, on the final (split) symbol matrix as the reference runs them with
 * its default build options (blend/pastix_subtask_blend.c:230-470,
 * PASTIX_BLEND_PROPMAP_2STEPS off):
 *   1. cost model: per block time of the diagonal factorization + TRSM and
 *      of each block update, from the fitted BLAS kernel models
 *      (symbol/symbol_cost.c:195-250 sum2dext, symbol/symbol_cost_perfs.c,
 *      blend/perf.h);
 *   2. elimination tree of the cblks with a virtual root over the forest
 *      (blend/elimintree.c:478-517), depth levels and subtree costs
 *      (blend/cand.c:250-310 candSubTreeBuild);
 *   3. proportional mapping of the threads over the subtrees
 *      (blend/propmap.c:120-458, no cross-processor restriction);
 *   4. simulation of the execution: every ready task goes to the candidate
 *      thread that can start it earliest; contributions make the facing
 *      cblk ready (blend/simu_run.c:340-1200, one process).
 */
#include <math.h>
#include "internal.h"

/* ---- fitted kernel time models (blend/perf.h: "AMD 6180 MKL") ---------- */
static double pa_perf_gemm( double i, double j, double k )
{
    return 2.429169e-10 * i * j * k + 2.724804e-10 * i * j + 1.328900e-09 * j * k + 1.148989e-07 * i -
           2.704179e-10 * j + 1.216278e-06;
}
static double pa_perf_geam( double i, double j )
{
    return 1.358111e-09 * i * j - 4.416379e-09 * i + 2.270780e-08 * j - 3.335563e-07;
}
static double pa_perf_trsm( double i, double j ) { return 2.626177e-10 * i * i * j + 3.976198e-08 * i + 3.255168e-06; }
static double pa_perf_potrf( double i )
{
    return 2.439599e-11 * i * i * i + 1.707006e-08 * i * i - 1.469893e-07 * i + 4.071507e-07;
}
static double pa_perf_scal( double i ) { return 4.371793e-10 * i + 2.052399e-07; }
static double pa_perf_copy( double i ) { return 9.177969e-10 * i + 2.266129e-07; }
static double pa_pos( double t ) { return t > 0. ? t : 0.; }

/* per factorization kind (symbol/symbol_cost_perfs.c:20-250; LL^h as
 * LL^t and LDL^h as LDL^t in real arithmetic) */
static double pa_cost_diag( int fact, double n )
{
    if ( fact == PastixFactLU ) return 2. * pa_pos( pa_perf_potrf( n ) );
    if ( fact == PastixFactLDLT || fact == PastixFactLDLH ) return pa_pos( pa_perf_potrf( n ) + n * pa_perf_copy( n ) );
    return pa_pos( pa_perf_potrf( n ) );
}
static double pa_cost_trsm( int fact, double m, double n )
{
    if ( fact == PastixFactLU ) return pa_pos( 2. * pa_perf_trsm( n, m ) );
    if ( fact == PastixFactLDLT || fact == PastixFactLDLH )
        return pa_pos( pa_perf_trsm( n, m ) + n * ( pa_perf_scal( m ) + pa_perf_copy( m ) ) );
    return pa_pos( pa_perf_trsm( n, m ) );
}
/* called as blkupdate(K, M, N) by sum2dext: the model's (M, N, K) are the
 * cblk width, the rows below and the block height */
static double pa_cost_update( int fact, double m, double n, double k )
{
    double t = pa_perf_gemm( m, n, k ) + pa_perf_geam( m, n );
    if ( fact == PastixFactLU ) t += pa_perf_gemm( m - n, n, k ) + pa_perf_geam( m - n, n );
    return pa_pos( t );
}

/* ---- binary min-heap on (key1, key2) (the reference's pqueue order,
 * kernels/queue.c:45-57) ------------------------------------------------- */
typedef struct { double k1, k2; pastix_int_t e; } pa_hitem;
typedef struct { pa_hitem *t; pastix_int_t n, cap; } pa_heap;

static int pa_hless( const pa_hitem *a, const pa_hitem *b )
{
    return ( a->k1 == b->k1 ) ? ( a->k2 < b->k2 ) : ( a->k1 < b->k1 );
}
static int pa_hpush( pa_heap *h, pastix_int_t e, double k1, double k2 )
{
    pastix_int_t i;
    if ( h->n == h->cap ) {
        pastix_int_t nc = h->cap ? 2 * h->cap : 16;
        pa_hitem    *nt = realloc( h->t, nc * sizeof( pa_hitem ) );
        if ( !nt ) return -1;
        h->t = nt; h->cap = nc;
    }
    i = h->n++;
    h->t[i].k1 = k1; h->t[i].k2 = k2; h->t[i].e = e;
    while ( i > 0 && pa_hless( &h->t[i], &h->t[( i - 1 ) / 2] ) ) {
        pa_hitem x = h->t[i]; h->t[i] = h->t[( i - 1 ) / 2]; h->t[( i - 1 ) / 2] = x;
        i = ( i - 1 ) / 2;
    }
    return 0;
}
static void pa_hpop( pa_heap *h )
{
    pastix_int_t i = 0;
    h->t[0] = h->t[--h->n];
    for ( ;; ) {
        pastix_int_t l = 2 * i + 1, r = l + 1, m = i;
        if ( l < h->n && pa_hless( &h->t[l], &h->t[m] ) ) m = l;
        if ( r < h->n && pa_hless( &h->t[r], &h->t[m] ) ) m = r;
        if ( m == i ) break;
        { pa_hitem x = h->t[i]; h->t[i] = h->t[m]; h->t[m] = x; }
        i = m;
    }
}

/* ---- proportional mapping (blend/propmap.c) ------------------------------ */
typedef struct {
    pastix_int_t *fson, *sons, *nsons;   /* children lists, node cblknbr = root */
    double       *ndecost, *subcost;
    pastix_int_t *fcand, *lcand;
    pastix_int_t *stack;                 /* cblknbr + 1 entries */
} pa_pmap;

#define PA_CROSS_TOLERANCE 0.1

/* the whole subtree of node to candidates [fc, lc] (propMappSubtreeOn1P):
 * the subtree of a node is the contiguous index range ending at the node
 * (sons precede their father), walked here without recursion */
static void pa_pmap_on1( pa_pmap *pm, pastix_int_t node, pastix_int_t fc, pastix_int_t lc )
{
    pastix_int_t top = 1, *st = pm->stack;
    st[0] = node;
    while ( top > 0 ) {
        pastix_int_t v = st[--top], i;
        pm->fcand[v] = fc;
        pm->lcand[v] = lc;
        for ( i = 0; i < pm->nsons[v]; i++ ) st[top++] = pm->sons[pm->fson[v] + i];
    }
}

static int pa_pmap_subtree( pa_pmap *pm, pastix_int_t node, pastix_int_t fcandnum, pastix_int_t lcandnum,
                            double *cost_remain )
{
    pastix_int_t candnbr, fcand, lcand, i, p, scandnbr, nsons;
    double       isocost, aspt_cost, cumul_cost, epsilon;
    pa_heap      qtree = { 0 }, qproc = { 0 };
    /* a chain of single sons is walked iteratively (the reference recurses
     * with the same candidates and budget) */
    for ( ;; ) {
        candnbr = lcandnum - fcandnum + 1;
        nsons   = pm->nsons[node];
        if ( candnbr == 1 ) {
            pa_pmap_on1( pm, node, fcandnum, lcandnum );
            return 0;
        }
        pm->fcand[node] = fcandnum;
        pm->lcand[node] = lcandnum;
        if ( nsons == 0 ) return 0;
        /* the node's own work is shared by all candidates */
        isocost = pm->ndecost[node] / candnbr;
        for ( p = 0; p < candnbr; p++ ) cost_remain[p] -= isocost;
        aspt_cost = pm->subcost[node] - pm->ndecost[node];
        /* candidates that already have more than their share leave the set */
        fcand = ( cost_remain[0] <= 0 ) ? 1 : 0;
        if ( cost_remain[candnbr - 1] <= 0 ) candnbr--;
        /* (the reference asserts fcand < candnbr; rounding can break it when
         * both ends are exhausted: keep the whole set then) */
        if ( fcand >= candnbr ) { fcand = 0; candnbr = lcandnum - fcandnum + 1; }
        cumul_cost = 0;
        for ( i = fcand; i < candnbr; i++ ) cumul_cost += cost_remain[i];
        if ( cumul_cost < aspt_cost && cumul_cost > 0 ) {
            double ratio = aspt_cost / cumul_cost;
            for ( i = fcand; i < candnbr; i++ ) cost_remain[i] *= ratio;
        }
        epsilon = ( PA_CROSS_TOLERANCE * cumul_cost ) / (double)candnbr;
        if ( nsons != 1 ) break;
        node = pm->sons[pm->fson[node]];
    }

    /* sons by decreasing subtree cost, then by decreasing own cost truncated
     * to an integer (the reference's key, propmap.c:217-226) */
    for ( i = 0; i < nsons; i++ ) {
        pastix_int_t s = pm->sons[pm->fson[node] + i];
        if ( pa_hpush( &qtree, i, -pm->subcost[s], (double)(pastix_int_t)( -pm->ndecost[s] ) ) )
            return PASTIX_ERR_OUTOFMEMORY;
    }
    /* first stage: subtrees that need several candidates */
    while ( qtree.n > 0 ) {
        pastix_int_t son;
        double      *sub;
        i          = qtree.t[0].e;
        cumul_cost = -qtree.t[0].k1;
        pa_hpop( &qtree );
        son = pm->sons[pm->fson[node] + i];
        if ( cumul_cost <= cost_remain[fcand] ) {
            cost_remain[fcand] -= cumul_cost;
            pa_pmap_on1( pm, son, fcandnum + fcand, fcandnum + fcand );
            break;
        }
        if ( fcand < candnbr - 1 && cost_remain[fcand] <= epsilon ) fcand++;
        lcand      = fcand;
        scandnbr   = 1;
        cumul_cost -= cost_remain[fcand];
        while ( cumul_cost > epsilon && lcand < candnbr - 1 ) {
            lcand++;
            scandnbr++;
            cumul_cost -= cost_remain[lcand];
        }
        sub = malloc( ( lcand - fcand + 1 ) * sizeof( double ) );
        if ( !sub ) return PASTIX_ERR_OUTOFMEMORY;
        if ( cumul_cost > 0 ) {
            isocost = cumul_cost / scandnbr;
            for ( p = 0; p < scandnbr; p++ ) {
                sub[p]                 = cost_remain[fcand + p] + isocost;
                cost_remain[fcand + p] = -isocost;
            }
        }
        else {
            for ( p = 0; p < scandnbr - 1; p++ ) {
                sub[p]                 = cost_remain[fcand + p];
                cost_remain[fcand + p] = 0.0;
            }
            sub[scandnbr - 1]                     = cost_remain[lcand] + cumul_cost;
            cost_remain[fcand + scandnbr - 1] = -cumul_cost;
        }
        if ( pa_pmap_subtree( pm, son, fcandnum + fcand, fcandnum + lcand, sub ) ) {
            free( sub );
            return PASTIX_ERR_OUTOFMEMORY;
        }
        free( sub );
        fcand = ( lcand < candnbr - 1 && cost_remain[lcand] < epsilon ) ? lcand + 1 : lcand;
    }
    /* second stage: the remaining single-candidate subtrees, largest first,
     * each to the candidate with the most remaining budget */
    if ( qtree.n > 0 ) {
        for ( i = 0; i < candnbr; i++ )
            if ( pa_hpush( &qproc, i, -cost_remain[i], 0. ) ) return PASTIX_ERR_OUTOFMEMORY;
        while ( qtree.n > 0 ) {
            pastix_int_t son;
            i          = qtree.t[0].e;
            cumul_cost = -qtree.t[0].k1;
            pa_hpop( &qtree );
            son   = pm->sons[pm->fson[node] + i];
            fcand = qproc.t[0].e;
            pa_hpop( &qproc );
            pa_pmap_on1( pm, son, fcandnum + fcand, fcandnum + fcand );
            cost_remain[fcand] -= cumul_cost;
            if ( pa_hpush( &qproc, fcand, -cost_remain[fcand], 0. ) ) return PASTIX_ERR_OUTOFMEMORY;
        }
    }
    free( qtree.t );
    free( qproc.t );
    return 0;
}

/* ---- the static schedule -------------------------------------------------- */
int
pa_blend_static( SolverMatrix *solv, int fact, pastix_int_t thrdnbr, int allcand, double *pred_time )
{
    const pastix_int_t cblknbr = solv->cblknbr, root = cblknbr;
    pastix_int_t       k, b, i, j, p, nmapped = 0, prio = 0;
    double            *blokcost, *treelevel, *ttime, *ptimer;
    pastix_int_t      *ctrb, *owner, *fcb, *cnt;
    pa_pmap            pm;
    pa_heap           *ready, *future;
    int                rc = PASTIX_SUCCESS;

    if ( thrdnbr < 1 ) thrdnbr = 1;
    /* previous schedule */
    if ( solv->ttsktab )
        for ( p = 0; p < solv->thrdnbr; p++ ) free( solv->ttsktab[p] );
    free( solv->ttsktab );
    free( solv->ttsknbr );
    free( (void *)solv->tasktab );
    solv->ttsktab = NULL; solv->ttsknbr = NULL; solv->tasktab = NULL;

    memset( &pm, 0, sizeof( pm ) );
    blokcost     = calloc( solv->bloknbr + 1, sizeof( double ) );
    treelevel    = calloc( cblknbr + 1, sizeof( double ) );
    ttime        = calloc( cblknbr + 1, sizeof( double ) );
    ptimer       = calloc( thrdnbr, sizeof( double ) );
    ctrb         = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
    owner        = malloc( ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    fcb          = malloc( ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    cnt          = calloc( thrdnbr + 1, sizeof( pastix_int_t ) );
    pm.fson      = calloc( cblknbr + 2, sizeof( pastix_int_t ) );
    pm.sons      = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
    pm.nsons     = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
    pm.ndecost   = calloc( cblknbr + 1, sizeof( double ) );
    pm.subcost   = calloc( cblknbr + 1, sizeof( double ) );
    pm.fcand     = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
    pm.lcand     = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
    pm.stack     = calloc( cblknbr + 2, sizeof( pastix_int_t ) );
    ready        = calloc( thrdnbr, sizeof( pa_heap ) );
    future       = calloc( thrdnbr, sizeof( pa_heap ) );
    solv->tasktab = calloc( cblknbr + 1, sizeof( Task ) );
    solv->ttsknbr = calloc( thrdnbr, sizeof( pastix_int_t ) );
    solv->ttsktab = calloc( thrdnbr, sizeof( pastix_int_t * ) );
    if ( !blokcost || !treelevel || !ttime || !ptimer || !ctrb || !owner || !fcb || !cnt || !pm.fson || !pm.sons ||
         !pm.nsons || !pm.ndecost || !pm.subcost || !pm.fcand || !pm.lcand || !pm.stack || !ready || !future || !solv->tasktab ||
         !solv->ttsknbr || !solv->ttsktab ) {
        rc = PASTIX_ERR_OUTOFMEMORY;
        goto done;
    }
    solv->thrdnbr = thrdnbr;
    solv->procnbr = thrdnbr;
    solv->bublnbr = thrdnbr;
    solv->clustnbr = 1;
    solv->clustnum = 0;
    solv->tasknbr = solv->tasknbr_1dp = cblknbr;

    /* 1. costs (symbol/symbol_cost.c:195-250) */
    for ( k = 0; k < cblknbr; k++ ) {
        const SolverCblk *c  = solv->cblktab + k;
        const pastix_int_t b0 = c->fblokptr - solv->bloktab, b1 = c[1].fblokptr - solv->bloktab;
        double             N = (double)cblk_colnbr( c ), M = 0., K = N, tot;
        for ( b = b0 + 1; b < b1; b++ ) M += (double)blok_rownbr( solv->bloktab + b );
        tot = pa_cost_diag( fact, N );
        if ( M > 0 ) tot += pa_cost_trsm( fact, M, N );
        blokcost[b0] = tot;
        for ( b = b0 + 1; b < b1; b++ ) {
            double nb = (double)blok_rownbr( solv->bloktab + b );
            blokcost[b] = pa_cost_update( fact, K, M, nb );
            tot += blokcost[b];
            M -= nb;
        }
        pm.ndecost[k] = tot;
    }
    /* 2. elimination tree (father = cblk faced by the first off-diagonal
     * block; roots under a virtual node), sons by increasing index */
    for ( k = 0; k < cblknbr; k++ ) {
        const SolverCblk *c = solv->cblktab + k;
        fcb[k] = ( c->fblokptr + 1 < c[1].fblokptr ) ? c->fblokptr[1].fcblknm : root;
        pm.nsons[fcb[k]]++;
    }
    pm.fson[0] = 0;
    for ( k = 0; k <= cblknbr; k++ ) pm.fson[k + 1] = pm.fson[k] + pm.nsons[k];
    {
        pastix_int_t *pos = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
        if ( !pos ) { rc = PASTIX_ERR_OUTOFMEMORY; goto done; }
        for ( k = 0; k <= cblknbr; k++ ) pos[k] = pm.fson[k];
        for ( k = 0; k < cblknbr; k++ ) pm.sons[pos[fcb[k]]++] = k;
        free( pos );
    }
    /* depth levels and subtree costs (blend/cand.c:250-310): a father's
     * index is larger than its sons', so an increasing pass completes each
     * subtree cost before it is added to the father, and a decreasing pass
     * gives the depths from the root down */
    pm.ndecost[root] = 0.;
    for ( k = 0; k <= cblknbr; k++ ) pm.subcost[k] = pm.ndecost[k];
    for ( k = 0; k < cblknbr; k++ ) pm.subcost[fcb[k]] += pm.subcost[k];
    treelevel[root] = 0.;
    for ( k = cblknbr - 1; k >= 0; k-- ) treelevel[k] = treelevel[fcb[k]] - 1.;

    /* 3. proportional mapping (blend/propmap.c:419-458) */
    if ( allcand ) {
        pa_pmap_on1( &pm, root, 0, thrdnbr - 1 );
    }
    else {
        double *remain = malloc( thrdnbr * sizeof( double ) );
        if ( !remain ) { rc = PASTIX_ERR_OUTOFMEMORY; goto done; }
        for ( p = 0; p < thrdnbr; p++ ) remain[p] = pm.subcost[root] / thrdnbr;
        rc = pa_pmap_subtree( &pm, root, 0, thrdnbr - 1, remain );
        free( remain );
        if ( rc ) goto done;
    }

    /* 4. simulation (blend/simu_run.c): contributions per facing cblk, one
     * per pair (i, j >= i) of off-diagonal blocks of a source cblk */
    for ( k = 0; k < cblknbr; k++ ) {
        const SolverCblk  *c  = solv->cblktab + k;
        const pastix_int_t b0 = c->fblokptr - solv->bloktab, b1 = c[1].fblokptr - solv->bloktab;
        for ( i = b0 + 1; i < b1; i++ ) ctrb[solv->bloktab[i].fcblknm] += b1 - i;
        owner[k] = -1;
    }
    for ( k = 0; k < cblknbr; k++ ) {
        if ( ctrb[k] != 0 ) continue;
        for ( p = pm.fcand[k]; p <= pm.lcand[k]; p++ )
            if ( pa_hpush( &ready[p], k, treelevel[k], (double)( solv->cblktab[k].fblokptr - solv->bloktab ) ) ) {
                rc = PASTIX_ERR_OUTOFMEMORY;
                goto done;
            }
    }
    while ( nmapped < cblknbr ) {
        pastix_int_t pr = -1, task = -1;
        double       early = 1e300, earlyproc = 1e300;
        /* the task a thread can start first (simu_getNextTaskNextProc) */
        for ( p = 0; p < thrdnbr; p++ ) {
            pastix_int_t t = -1;
            double       tr;
            while ( ready[p].n > 0 && owner[ready[p].t[0].e] >= 0 ) pa_hpop( &ready[p] );
            if ( ready[p].n > 0 ) t = ready[p].t[0].e;
            else {
                while ( future[p].n > 0 && owner[future[p].t[0].e] >= 0 ) pa_hpop( &future[p] );
                if ( future[p].n > 0 ) t = future[p].t[0].e;
            }
            if ( t < 0 ) continue;
            tr = ptimer[p] > ttime[t] ? ptimer[p] : ttime[t];
            if ( tr == early && ptimer[p] < earlyproc ) {
                pr = p; earlyproc = ptimer[p]; task = t; early = tr;
            }
            if ( tr < early ) {
                pr = p; task = t; early = tr;
            }
        }
        if ( task < 0 ) {
            pa_error( "static schedule: the simulation found no ready task (%ld of %ld mapped)", (long)nmapped,
                      (long)cblknbr );
            rc = PASTIX_ERR_INTERNAL;
            goto done;
        }
        owner[task] = pr;
        nmapped++;
        solv->tasktab[task].prionum = prio++;
        cnt[pr]++;
        if ( ttime[task] > ptimer[pr] ) ptimer[pr] = ttime[task];
        /* the task's diagonal block, TRSM and updates (simu_computeTask) */
        {
            const SolverCblk  *c  = solv->cblktab + task;
            const pastix_int_t b0 = c->fblokptr - solv->bloktab, b1 = c[1].fblokptr - solv->bloktab;
            ptimer[pr] += blokcost[b0];
            for ( i = b0 + 1; i < b1; i++ ) {
                const pastix_int_t f = solv->bloktab[i].fcblknm;
                ptimer[pr] += blokcost[i];
                for ( j = i; j < b1; j++ ) {
                    ctrb[f]--;
                    if ( ptimer[pr] > ttime[f] ) ttime[f] = ptimer[pr];
                    if ( ctrb[f] == 0 ) {
                        /* ready: into every candidate's ready or future heap */
                        const double lvl = treelevel[f];
                        const double fb  = (double)( solv->cblktab[f].fblokptr - solv->bloktab );
                        for ( p = pm.fcand[f]; p <= pm.lcand[f]; p++ ) {
                            int e = ( ttime[f] > ptimer[p] ) ? pa_hpush( &future[p], f, ttime[f], lvl )
                                                              : pa_hpush( &ready[p], f, lvl, fb );
                            if ( e ) { rc = PASTIX_ERR_OUTOFMEMORY; goto done; }
                        }
                    }
                }
            }
        }
        /* future tasks of this thread that are now ready (simu_pushToReadyHeap) */
        while ( future[pr].n > 0 ) {
            pastix_int_t t = future[pr].t[0].e;
            if ( owner[t] >= 0 ) { pa_hpop( &future[pr] ); continue; }
            if ( !( ttime[t] < ptimer[pr] ) ) break;
            pa_hpop( &future[pr] );
            if ( pa_hpush( &ready[pr], t, treelevel[t], (double)( solv->cblktab[t].fblokptr - solv->bloktab ) ) ) {
                rc = PASTIX_ERR_OUTOFMEMORY;
                goto done;
            }
        }
    }

    /* tasks (one COMP_1D task per cblk) and the per-thread lists in the
     * simulated order */
    for ( k = 0; k < cblknbr; k++ ) {
        Task *t    = solv->tasktab + k;
        t->taskid  = 0;
        t->cblknum = k;
        t->bloknum = solv->cblktab[k].fblokptr - solv->bloktab;
        t->ctrbcnt = solv->cblktab[k + 1].brownum - solv->cblktab[k].brownum;
        solv->cblktab[k].threadid = (int)owner[k];
        solv->cblktab[k].ownerid  = 0;
        solv->cblktab[k].priority = t->prionum;
    }
    for ( p = 0; p < thrdnbr; p++ ) {
        solv->ttsktab[p] = malloc( ( cnt[p] + 1 ) * sizeof( pastix_int_t ) );
        if ( !solv->ttsktab[p] ) { rc = PASTIX_ERR_OUTOFMEMORY; goto done; }
        solv->ttsknbr[p] = 0;
    }
    {
        /* tasks by increasing simulated rank */
        pastix_int_t *byprio = malloc( ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
        if ( !byprio ) { rc = PASTIX_ERR_OUTOFMEMORY; goto done; }
        for ( k = 0; k < cblknbr; k++ ) byprio[solv->tasktab[k].prionum] = k;
        for ( i = 0; i < cblknbr; i++ ) {
            k = byprio[i];
            p = owner[k];
            solv->ttsktab[p][solv->ttsknbr[p]++] = k;
        }
        free( byprio );
    }
    if ( pred_time ) {
        double tmax = 0.;
        for ( p = 0; p < thrdnbr; p++ ) tmax = ptimer[p] > tmax ? ptimer[p] : tmax;
        *pred_time = tmax;
    }

done:
    free( blokcost ); free( treelevel ); free( ttime ); free( ptimer ); free( ctrb ); free( owner ); free( fcb );
    free( cnt ); free( pm.fson ); free( pm.sons ); free( pm.nsons ); free( pm.ndecost ); free( pm.subcost );
    free( pm.fcand ); free( pm.lcand ); free( pm.stack );
    if ( ready ) for ( p = 0; p < thrdnbr; p++ ) free( ready[p].t );
    if ( future ) for ( p = 0; p < thrdnbr; p++ ) free( future[p].t );
    free( ready ); free( future );
    return rc;
}
