/*
 * solver_io.c -- SolverMatrix structure I/O in the reference's text format
 * (reference blend/solver_io.c:42-355).
 *
 * Layout (whitespace separated integers, as the reference's solverSave
 * writes them):
 *   2                                   version
 *   cblknbr bloknbr nodenbr baseval
 *   per cblk:  fcolnum lcolnum first-block-index
 *   per block: frownum lrownum fcblknm
 *   coefnbr gemmmax 0 0 clustnum clustnbr tasknbr procnbr thrdnbr
 *   per cblk:  stride
 *   per block: coefind
 *   per task:  taskid prionum cblknum bloknum ctrbcnt
 *   per thread: ttsknbr, then its ttsknbr task indices
 *
 * The reference's loader refuses version 2, the only version its saver
 * writes (blend/solver_io.c:79-82); this loader reads version 2 so that a
 * saved structure can be loaded back, and rebuilds the fields the file does
 * not hold (the cblk of each block, browtab, the 2D layout flags, the
 * per-cblk maxima).
 */
#include <inttypes.h>
#include "internal.h"

static int
pa_int_load( FILE *f, pastix_int_t *v )
{
    long long x;
    if ( fscanf( f, "%lld", &x ) != 1 ) return 0;
    *v = (pastix_int_t)x;
    return 1;
}

int
solverSave( const SolverMatrix *s, FILE *f )
{
    pastix_int_t k, b, t, i;
    int          o = 0;
    if ( s == NULL || f == NULL ) return PASTIX_ERR_BADPARAMETER;
    o |= fprintf( f, "2\n%ld\t%ld\t%ld\t%ld\n", (long)s->cblknbr, (long)s->bloknbr, (long)s->nodenbr,
                  (long)s->baseval ) < 0;
    for ( k = 0; k < s->cblknbr && !o; k++ ) {
        const SolverCblk *c = s->cblktab + k;
        o |= fprintf( f, "%ld\t%ld\t%ld\n", (long)c->fcolnum, (long)c->lcolnum, (long)( c->fblokptr - s->bloktab ) ) < 0;
    }
    for ( b = 0; b < s->bloknbr && !o; b++ ) {
        const SolverBlok *bl = s->bloktab + b;
        o |= fprintf( f, "%ld\t%ld\t%ld\n", (long)bl->frownum, (long)bl->lrownum, (long)bl->fcblknm ) < 0;
    }
    o |= fprintf( f, "\n%ld\t%ld\t%ld\t%ld\t%ld\t%ld\t%ld\t%ld\t%ld\n", (long)s->coefnbr, (long)s->gemmmax, 0L, 0L,
                  (long)s->clustnum, (long)s->clustnbr, (long)s->tasknbr, (long)s->procnbr, (long)s->thrdnbr ) < 0;
    for ( k = 0; k < s->cblknbr && !o; k++ ) o |= fprintf( f, "%ld\n", (long)s->cblktab[k].stride ) < 0;
    for ( b = 0; b < s->bloknbr && !o; b++ ) o |= fprintf( f, "%ld\n", (long)s->bloktab[b].coefind ) < 0;
    o |= fprintf( f, "\n\n" ) < 0;
    for ( t = 0; t < s->tasknbr && !o; t++ ) {
        const Task *tk = s->tasktab + t;
        o |= fprintf( f, "%ld\t%ld\t%ld\t%ld\t%ld\n", (long)tk->taskid, (long)tk->prionum, (long)tk->cblknum,
                      (long)tk->bloknum, (long)tk->ctrbcnt ) < 0;
    }
    for ( t = 0; t < s->thrdnbr && !o; t++ ) {
        pastix_int_t nt = s->ttsknbr ? s->ttsknbr[t] : 0;
        o |= fprintf( f, "%ld\n", (long)nt ) < 0;
        for ( i = 0; i < nt && !o; i++ ) o |= fprintf( f, "%ld\n", (long)s->ttsktab[t][i] ) < 0;
    }
    return o ? PASTIX_ERR_FILE : PASTIX_SUCCESS;
}

static int
pa_load_fail( SolverMatrix *s, const char *what )
{
    pa_error( "solverLoad: bad input (%s)", what );
    pa_solver_exit( s );
    return PASTIX_ERR_FILE;
}

int
solverLoad( SolverMatrix *s, FILE *f )
{
    pastix_int_t vers = -1, cblknbr = 0, bloknbr = 0, nodenbr = 0, baseval = -1, k, b, t, i, first = 0, tmp = 0;
    pastix_int_t clustnum = 0, clustnbr = 0;
    if ( s == NULL || f == NULL ) return PASTIX_ERR_BADPARAMETER;
    memset( s, 0, sizeof( *s ) );
    if ( pa_int_load( f, &vers ) + pa_int_load( f, &cblknbr ) + pa_int_load( f, &bloknbr ) +
             pa_int_load( f, &nodenbr ) + pa_int_load( f, &baseval ) != 5 ||
         vers != 2 || cblknbr < 0 || bloknbr < cblknbr || nodenbr < cblknbr || baseval != 0 ) {
        pa_error( "solverLoad: bad header (version 2, 0-based numbering expected)" );
        return PASTIX_ERR_FILE;
    }
    s->cblknbr = s->gcblknbr = cblknbr;
    s->bloknbr = s->gbloknbr = bloknbr;
    s->nodenbr = nodenbr;
    s->cblktab = calloc( cblknbr + 1, sizeof( SolverCblk ) );
    s->bloktab = calloc( bloknbr + 1, sizeof( SolverBlok ) );
    if ( !s->cblktab || !s->bloktab ) return pa_load_fail( s, "out of memory" );
    for ( k = 0; k < cblknbr; k++ ) {
        SolverCblk *c = s->cblktab + k;
        if ( pa_int_load( f, &c->fcolnum ) + pa_int_load( f, &c->lcolnum ) + pa_int_load( f, &first ) != 3 ||
             c->fcolnum > c->lcolnum || first < 0 || first >= bloknbr )
            return pa_load_fail( s, "column blocks" );
        c->fblokptr = s->bloktab + first;
    }
    s->cblktab[cblknbr].fcolnum = s->cblktab[cblknbr].lcolnum = nodenbr;
    s->cblktab[cblknbr].fblokptr = s->bloktab + bloknbr;
    s->cblktab[cblknbr].lcolidx  = nodenbr;
    for ( b = 0; b < bloknbr; b++ ) {
        SolverBlok *bl = s->bloktab + b;
        if ( pa_int_load( f, &bl->frownum ) + pa_int_load( f, &bl->lrownum ) + pa_int_load( f, &bl->fcblknm ) != 3 ||
             bl->frownum > bl->lrownum || bl->fcblknm < 0 || bl->fcblknm >= cblknbr )
            return pa_load_fail( s, "blocks" );
    }
    if ( pa_int_load( f, &s->coefnbr ) + pa_int_load( f, &s->gemmmax ) + pa_int_load( f, &tmp ) +
             pa_int_load( f, &tmp ) + pa_int_load( f, &clustnum ) + pa_int_load( f, &clustnbr ) +
             pa_int_load( f, &s->tasknbr ) + pa_int_load( f, &s->procnbr ) + pa_int_load( f, &s->thrdnbr ) != 9 ||
         s->tasknbr < 0 || s->thrdnbr < 0 )
        return pa_load_fail( s, "counts" );
    s->clustnum = (int)clustnum;
    s->clustnbr = (int)clustnbr;
    for ( k = 0; k < cblknbr; k++ )
        if ( !pa_int_load( f, &s->cblktab[k].stride ) ) return pa_load_fail( s, "strides" );
    for ( b = 0; b < bloknbr; b++ )
        if ( !pa_int_load( f, &s->bloktab[b].coefind ) ) return pa_load_fail( s, "coefficient offsets" );
    s->tasktab = calloc( s->tasknbr + 1, sizeof( Task ) );
    s->ttsknbr = calloc( s->thrdnbr + 1, sizeof( pastix_int_t ) );
    s->ttsktab = calloc( s->thrdnbr + 1, sizeof( pastix_int_t * ) );
    if ( !s->tasktab || !s->ttsknbr || !s->ttsktab ) return pa_load_fail( s, "out of memory" );
    for ( t = 0; t < s->tasknbr; t++ ) {
        Task *tk = s->tasktab + t;
        if ( pa_int_load( f, &tk->taskid ) + pa_int_load( f, &tk->prionum ) + pa_int_load( f, &tk->cblknum ) +
                 pa_int_load( f, &tk->bloknum ) + pa_int_load( f, &tmp ) != 5 ||
             tk->cblknum < 0 || tk->cblknum >= cblknbr )
            return pa_load_fail( s, "tasks" );
        tk->ctrbcnt = tmp;
    }
    for ( t = 0; t < s->thrdnbr; t++ ) {
        if ( !pa_int_load( f, &s->ttsknbr[t] ) || s->ttsknbr[t] < 0 || s->ttsknbr[t] > s->tasknbr )
            return pa_load_fail( s, "thread task counts" );
        s->ttsktab[t] = malloc( ( s->ttsknbr[t] + 1 ) * sizeof( pastix_int_t ) );
        if ( !s->ttsktab[t] ) return pa_load_fail( s, "out of memory" );
        for ( i = 0; i < s->ttsknbr[t]; i++ )
            if ( !pa_int_load( f, &s->ttsktab[t][i] ) || s->ttsktab[t][i] < 0 || s->ttsktab[t][i] >= s->tasknbr )
                return pa_load_fail( s, "thread task lists" );
    }

    /* fields the file does not hold: the cblk of each block, the 2D layout,
     * the maxima of the cblks, browtab; the blocks of a cblk must be
     * contiguous, start with the diagonal block and tile the coefficients
     * of the cblk in the 2D layout */
    for ( k = 0; k < cblknbr; k++ ) {
        SolverCblk  *c  = s->cblktab + k;
        pastix_int_t nc = c->lcolnum - c->fcolnum + 1, rows = 0, nblk = c[1].fblokptr - c->fblokptr;
        if ( nblk < 1 || ( k == 0 && c->fblokptr != s->bloktab ) || c->fblokptr->frownum != c->fcolnum ||
             c->fblokptr->lrownum != c->lcolnum || c->fblokptr->fcblknm != k )
            return pa_load_fail( s, "diagonal blocks" );
        c->cblktype  = CBLK_LAYOUT_2D | CBLK_TASKS_2D;
        c->lcolidx   = c->fcolnum;
        c->gcblknum  = k;
        c->bcscnum   = k;
        c->sndeidx   = k;
        c->gfaninnum = -1;
        for ( b = c->fblokptr - s->bloktab; b < c[1].fblokptr - s->bloktab; b++ ) {
            SolverBlok  *bl = s->bloktab + b;
            pastix_int_t m  = bl->lrownum - bl->frownum + 1;
            if ( bl->coefind != rows * nc ) return pa_load_fail( s, "2D tile offsets" );
            bl->lcblknm  = k;
            bl->gbloknm  = b;
            bl->gfaninnm = -1;
            rows += m;
            if ( b > c->fblokptr - s->bloktab ) s->blokmax = pa_imax( s->blokmax, m * nc );
        }
        if ( rows != c->stride ) return pa_load_fail( s, "strides" );
        s->cblkmaxblk = pa_imax( s->cblkmaxblk, nblk );
        s->colmax     = pa_imax( s->colmax, nc );
    }
    return pa_solver_build_browtab( s );
}
