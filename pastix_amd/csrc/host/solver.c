/*
 * solver.c -- SolverMatrix generation in the 2D tile layout and the
 * elimination-tree levels that drive the batched device schedule.
 */
#include <assert.h>
#include "internal.h"

/*
 * Build the SolverMatrix of a single-process run from the final symbol matrix
 * (reference blend/solver_matrix_gen.c:536-640 with
 * solver_matrix_gen_utils.c:1003-1057).  Every cblk uses the 2D layout
 * (CBLK_LAYOUT_2D | CBLK_TASKS_2D, reference blend/cand.c:65): block b of a
 * cblk with nc columns is a column-major M_b x nc tile with leading dimension
 * M_b that starts at coefind = (rows of the blocks above b) * nc
 * (reference blend/solver_matrix_gen_utils.h:93).
 */
int
pa_solver_gen( SolverMatrix *solv, const symbol_matrix_t *symb, int factotype, int flttype )
{
    pastix_int_t k, b, cblknbr = symb->cblknbr, bloknbr = symb->bloknbr;
    pastix_int_t coefnbr = 0;

    memset( solv, 0, sizeof( *solv ) );
    solv->nodenbr   = symb->nodenbr;
    solv->cblknbr   = cblknbr;
    solv->gcblknbr  = cblknbr;
    solv->bloknbr   = bloknbr;
    solv->gbloknbr  = bloknbr;
    solv->factotype = factotype;
    solv->flttype   = flttype;
    solv->cblktab   = calloc( cblknbr + 1, sizeof( SolverCblk ) );
    solv->bloktab   = calloc( bloknbr + 1, sizeof( SolverBlok ) );

    for ( k = 0; k < cblknbr; k++ ) {
        const symbol_cblk_t *sc = symb->cblktab + k;
        SolverCblk *c = solv->cblktab + k;
        pastix_int_t nc = sc->lcolnum - sc->fcolnum + 1, stride = 0, nblk;

        c->cblktype = CBLK_LAYOUT_2D | CBLK_TASKS_2D;
        c->fcolnum  = sc->fcolnum;
        c->lcolnum  = sc->lcolnum;
        c->fblokptr = solv->bloktab + sc->bloknum;
        c->lcolidx  = sc->fcolnum;
        c->gcblknum = k;
        c->bcscnum  = k;
        c->sndeidx  = k;
        c->gfaninnum = -1;
        for ( b = sc->bloknum; b < symb->cblktab[k + 1].bloknum; b++ ) {
            const symbol_blok_t *sb = symb->bloktab + b;
            SolverBlok *bl = solv->bloktab + b;
            bl->lcblknm  = k;
            bl->fcblknm  = sb->fcblknm;
            bl->gbloknm  = b;
            bl->gfaninnm = -1;
            bl->frownum  = sb->frownum;
            bl->lrownum  = sb->lrownum;
            bl->coefind  = stride * nc;
            stride += sb->lrownum - sb->frownum + 1;
        }
        c->stride = stride;
        nblk = symb->cblktab[k + 1].bloknum - sc->bloknum;
        solv->cblkmaxblk = pa_imax( solv->cblkmaxblk, nblk );
        solv->colmax     = pa_imax( solv->colmax, nc );
        solv->gemmmax    = pa_imax( solv->gemmmax, stride * nc );
        for ( b = sc->bloknum + 1; b < symb->cblktab[k + 1].bloknum; b++ ) {
            solv->blokmax = pa_imax( solv->blokmax, ( symb->bloktab[b].lrownum - symb->bloktab[b].frownum + 1 ) * nc );
        }
        coefnbr += stride * nc;
    }
    solv->cblktab[cblknbr].fcolnum  = symb->nodenbr;
    solv->cblktab[cblknbr].lcolnum  = symb->nodenbr;
    solv->cblktab[cblknbr].fblokptr = solv->bloktab + bloknbr;
    solv->cblktab[cblknbr].lcolidx  = symb->nodenbr;
    solv->coefnbr = coefnbr;

    return pa_solver_build_browtab( solv );
}

/* browtab: the incoming blocks of each cblk by increasing source cblk, the
 * blocks' browind and the cblks' contribution counts (reference
 * blend/solver_matrix_gen.c:571-620) */
int
pa_solver_build_browtab( SolverMatrix *solv )
{
    pastix_int_t k, b, cblknbr = solv->cblknbr, bloknbr = solv->bloknbr, *pos;
    {
        pastix_int_t *cnt = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
        for ( b = 0; b < bloknbr; b++ ) {
            if ( solv->bloktab[b].fcblknm != solv->bloktab[b].lcblknm ) cnt[solv->bloktab[b].fcblknm]++;
        }
        solv->cblktab[0].brownum = 0;
        for ( k = 0; k < cblknbr; k++ ) {
            solv->cblktab[k + 1].brownum = solv->cblktab[k].brownum + cnt[k];
            solv->cblktab[k].ctrbcnt     = (int32_t)cnt[k];
        }
        solv->brownbr = solv->cblktab[cblknbr].brownum;
        solv->browtab = malloc( ( solv->brownbr + 1 ) * sizeof( pastix_int_t ) );
        pos = cnt;
        for ( k = 0; k < cblknbr; k++ ) pos[k] = solv->cblktab[k].brownum;
        for ( b = 0; b < bloknbr; b++ ) {
            SolverBlok *bl = solv->bloktab + b;
            if ( bl->fcblknm != bl->lcblknm ) {
                bl->browind = pos[bl->fcblknm];
                solv->browtab[pos[bl->fcblknm]++] = b;
            }
            else {
                bl->browind = -1;
            }
        }
        for ( k = 0; k < cblknbr; k++ ) solv->cblktab[k].brown2d = solv->cblktab[k].brownum;
        free( cnt );
    }
    return PASTIX_SUCCESS;
}

void
pa_solver_exit( SolverMatrix *solv )
{
    pastix_int_t t;
    if ( solv ) {
        free( solv->cblktab );
        free( solv->bloktab );
        free( solv->browtab );
        if ( solv->ttsktab )
            for ( t = 0; t < solv->thrdnbr; t++ ) free( solv->ttsktab[t] );
        free( solv->ttsktab );
        free( solv->ttsknbr );
        free( (void *)solv->tasktab );
        memset( solv, 0, sizeof( *solv ) );
    }
}

void
solverExit( SolverMatrix *solv )
{
    pa_solver_exit( solv );
}

/*
 * Height of each cblk in the supernodal elimination tree, whose father is the
 * cblk faced by the first off-diagonal block: leaves are at level 0 and a
 * father is one level above its highest son.  Two cblks of the same level are
 * never ancestor and descendant, so a level can be factored in one batched
 * pass once all lower levels are done.  Returns the number of levels.
 */
pastix_int_t
pa_solver_levels( const SolverMatrix *solv, pastix_int_t *level )
{
    pastix_int_t k, nlev = 0;
    for ( k = 0; k < solv->cblknbr; k++ ) level[k] = 0;
    for ( k = 0; k < solv->cblknbr; k++ ) {
        const SolverCblk *c = solv->cblktab + k;
        if ( c->fblokptr + 1 < c[1].fblokptr ) {
            pastix_int_t fa = c->fblokptr[1].fcblknm;
            assert( fa > k );
            if ( level[fa] < level[k] + 1 ) level[fa] = level[k] + 1;
        }
        nlev = pa_imax( nlev, level[k] + 1 );
    }
    return nlev;
}
