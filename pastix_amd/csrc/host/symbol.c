/*
 * symbol.c -- adjacency graph, supernodal symbolic factorization, browtab,
 * structural check and the flop model of the block structure.
 */
#include <assert.h>
#include <math.h>
#include "internal.h"

static int
cmp_int( const void *a, const void *b )
{
    pastix_int_t x = *(const pastix_int_t *)a, y = *(const pastix_int_t *)b;
    return ( x > y ) - ( x < y );
}

/*
 * Symmetrised pattern of the user matrix without the diagonal, 0-based CSC in
 * the user numbering (the role of reference graph/graph_prepare.c).  Accepts a
 * CSC or CSR spm storing either triangle or both.
 */
int
pa_graph_from_spm( pa_graph_t *g, const spmatrix_t *spm )
{
    pastix_int_t n = spm->n, b = spm->baseval, i, e, *cnt, *pos;
    const pastix_int_t *ptr = spm->colptr, *idx = spm->rowptr;

    if ( spm->fmttype == SpmCSR ) {
        ptr = spm->rowptr;
        idx = spm->colptr;
    }
    else if ( spm->fmttype != SpmCSC ) {
        pa_error( "graph: only CSC and CSR matrices are supported" );
        return PASTIX_ERR_BADPARAMETER;
    }
    cnt = calloc( n + 1, sizeof( pastix_int_t ) );
    for ( i = 0; i < n; i++ ) {
        for ( e = ptr[i] - b; e < ptr[i + 1] - b; e++ ) {
            pastix_int_t r = idx[e] - b;
            if ( r == i ) continue;
            cnt[i]++;
            cnt[r]++;
        }
    }
    g->n      = n;
    g->colptr = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    g->colptr[0] = 0;
    for ( i = 0; i < n; i++ ) g->colptr[i + 1] = g->colptr[i] + cnt[i];
    g->rows = malloc( ( g->colptr[n] + 1 ) * sizeof( pastix_int_t ) );
    pos = cnt;
    memcpy( pos, g->colptr, n * sizeof( pastix_int_t ) );
    for ( i = 0; i < n; i++ ) {
        for ( e = ptr[i] - b; e < ptr[i + 1] - b; e++ ) {
            pastix_int_t r = idx[e] - b;
            if ( r == i ) continue;
            g->rows[pos[i]++] = r;
            g->rows[pos[r]++] = i;
        }
    }
    /* sort and drop duplicates (a matrix storing both triangles contributes
     * every edge twice) */
    {
        pastix_int_t w = 0, start = 0;
        for ( i = 0; i < n; i++ ) {
            pastix_int_t s = g->colptr[i], t = g->colptr[i + 1], last = -1;
            qsort( g->rows + s, t - s, sizeof( pastix_int_t ), cmp_int );
            g->colptr[i] = start;
            for ( e = s; e < t; e++ ) {
                if ( g->rows[e] != last ) {
                    last = g->rows[w++] = g->rows[e];
                }
            }
            start = w;
        }
        g->colptr[n] = w;
    }
    free( cnt );
    return PASTIX_SUCCESS;
}

void
pa_graph_exit( pa_graph_t *g )
{
    if ( g ) {
        free( g->colptr );
        free( g->rows );
        memset( g, 0, sizeof( *g ) );
    }
}

void
pa_symbol_exit( symbol_matrix_t *s )
{
    if ( s ) {
        free( s->cblktab );
        free( s->bloktab );
        free( s->browtab );
        memset( s, 0, sizeof( *s ) );
    }
}

/*
 * Block symbolic factorization for a given supernode partition.
 *
 * The off-diagonal row set of supernode k is the union of the rows, below the
 * supernode, of the adjacency of its columns and of the off-diagonal row sets
 * of the supernodes whose first off-diagonal row falls into k.  The diagonal
 * block is dense; the rows are cut into blocks that are maximal runs of
 * consecutive rows inside one supernode.  This is the structure built by the
 * reference's direct fax (symbol/symbol_fax.c:300-490): same diagonal block,
 * same row sets, same block boundaries.
 */
int
pa_symbol_fax( symbol_matrix_t *symb, const pa_graph_t *g, const pastix_order_t *o )
{
    pastix_int_t  n = o->vertnbr, cblknbr = o->cblknbr, k, c, e;
    pastix_int_t *col2cblk, *mark, *buf;
    pastix_int_t **rows, *nrows, *child_head, *child_next;
    pastix_int_t  bloknbr = 0, blokcap;
    symbol_blok_t *blok;

    col2cblk   = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    mark       = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    buf        = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    rows       = calloc( cblknbr, sizeof( pastix_int_t * ) );
    nrows      = calloc( cblknbr, sizeof( pastix_int_t ) );
    child_head = malloc( cblknbr * sizeof( pastix_int_t ) );
    child_next = malloc( cblknbr * sizeof( pastix_int_t ) );

    for ( k = 0; k < cblknbr; k++ ) {
        child_head[k] = -1;
        for ( c = o->rangtab[k]; c < o->rangtab[k + 1]; c++ ) col2cblk[c] = k;
    }
    for ( c = 0; c < n; c++ ) mark[c] = -1;

    memset( symb, 0, sizeof( *symb ) );
    symb->cblknbr = cblknbr;
    symb->nodenbr = n;
    symb->cblktab = calloc( cblknbr + 1, sizeof( symbol_cblk_t ) );
    blokcap       = 2 * cblknbr + 16;
    blok          = malloc( blokcap * sizeof( symbol_blok_t ) );

    for ( k = 0; k < cblknbr; k++ ) {
        pastix_int_t fcol = o->rangtab[k], lcol = o->rangtab[k + 1] - 1, cnt = 0, ch, i;
        for ( c = fcol; c <= lcol; c++ ) {
            pastix_int_t old = o->peritab[c];
            for ( e = g->colptr[old]; e < g->colptr[old + 1]; e++ ) {
                pastix_int_t r = o->permtab[g->rows[e]];
                if ( r > lcol && mark[r] != k ) {
                    mark[r]    = k;
                    buf[cnt++] = r;
                }
            }
        }
        for ( ch = child_head[k]; ch != -1; ch = child_next[ch] ) {
            for ( i = 0; i < nrows[ch]; i++ ) {
                pastix_int_t r = rows[ch][i];
                if ( r > lcol && mark[r] != k ) {
                    mark[r]    = k;
                    buf[cnt++] = r;
                }
            }
            free( rows[ch] );
            rows[ch] = NULL;
        }
        qsort( buf, cnt, sizeof( pastix_int_t ), cmp_int );
        nrows[k] = cnt;
        if ( cnt > 0 ) {
            pastix_int_t fa = col2cblk[buf[0]];
            rows[k] = malloc( cnt * sizeof( pastix_int_t ) );
            memcpy( rows[k], buf, cnt * sizeof( pastix_int_t ) );
            child_next[k]  = child_head[fa];
            child_head[fa] = k;
        }

        if ( bloknbr + cnt + 2 > blokcap ) {
            blokcap = 2 * ( bloknbr + cnt + 2 );
            blok    = realloc( blok, blokcap * sizeof( symbol_blok_t ) );
        }
        symb->cblktab[k].fcolnum = fcol;
        symb->cblktab[k].lcolnum = lcol;
        symb->cblktab[k].bloknum = bloknbr;
        blok[bloknbr].frownum = fcol;
        blok[bloknbr].lrownum = lcol;
        blok[bloknbr].lcblknm = k;
        blok[bloknbr].fcblknm = k;
        bloknbr++;
        for ( i = 0; i < cnt; i++ ) {
            if ( i == 0 || buf[i] != buf[i - 1] + 1 || col2cblk[buf[i]] != col2cblk[buf[i - 1]] ) {
                blok[bloknbr].frownum = buf[i];
                blok[bloknbr].lcblknm = k;
                blok[bloknbr].fcblknm = col2cblk[buf[i]];
                bloknbr++;
            }
            blok[bloknbr - 1].lrownum = buf[i];
        }
    }
    symb->cblktab[cblknbr].fcolnum = n;
    symb->cblktab[cblknbr].lcolnum = n;
    symb->cblktab[cblknbr].bloknum = bloknbr;
    symb->bloknbr = bloknbr;
    symb->bloktab = realloc( blok, ( bloknbr + 1 ) * sizeof( symbol_blok_t ) );
    memset( symb->bloktab + bloknbr, 0, sizeof( symbol_blok_t ) );

    for ( k = 0; k < cblknbr; k++ ) free( rows[k] );
    free( rows );
    free( nrows );
    free( child_head );
    free( child_next );
    free( col2cblk );
    free( mark );
    free( buf );
    return PASTIX_SUCCESS;
}

/*
 * browtab: for each cblk, the off-diagonal blocks of other cblks that face it,
 * ordered by source cblk (reference symbol/symbol.c pastixSymbolBuildRowtab).
 */
void
pa_symbol_build_browtab( symbol_matrix_t *symb )
{
    pastix_int_t k, b, *cnt, *pos;
    free( symb->browtab );
    cnt = calloc( symb->cblknbr + 1, sizeof( pastix_int_t ) );
    for ( k = 0; k < symb->cblknbr; k++ ) {
        for ( b = symb->cblktab[k].bloknum + 1; b < symb->cblktab[k + 1].bloknum; b++ ) {
            cnt[symb->bloktab[b].fcblknm]++;
        }
    }
    symb->browmax = 0;
    pos = malloc( ( symb->cblknbr + 1 ) * sizeof( pastix_int_t ) );
    symb->cblktab[0].brownum = 0;
    for ( k = 0; k < symb->cblknbr; k++ ) {
        symb->cblktab[k + 1].brownum = symb->cblktab[k].brownum + cnt[k];
        pos[k] = symb->cblktab[k].brownum;
        symb->browmax = pa_imax( symb->browmax, cnt[k] );
    }
    symb->browtab = malloc( ( symb->cblktab[symb->cblknbr].brownum + 1 ) * sizeof( pastix_int_t ) );
    for ( k = 0; k < symb->cblknbr; k++ ) {
        for ( b = symb->cblktab[k].bloknum + 1; b < symb->cblktab[k + 1].bloknum; b++ ) {
            symb->browtab[pos[symb->bloktab[b].fcblknm]++] = b;
        }
    }
    free( pos );
    free( cnt );
}

/* Invariants of reference symbol/symbol_check.c: blocks sorted and disjoint,
 * each block inside the cblk it faces, diagonal block first. */
int
pa_symbol_check( const symbol_matrix_t *s )
{
    pastix_int_t k, b;
    for ( k = 0; k < s->cblknbr; k++ ) {
        const symbol_cblk_t *c = s->cblktab + k;
        const symbol_blok_t *d = s->bloktab + c->bloknum;
        if ( c->fcolnum > c->lcolnum || d->frownum != c->fcolnum || d->lrownum != c->lcolnum ||
             d->fcblknm != k ) {
            return PASTIX_ERR_INTERNAL;
        }
        for ( b = c->bloknum + 1; b < s->cblktab[k + 1].bloknum; b++ ) {
            const symbol_blok_t *bl = s->bloktab + b;
            const symbol_cblk_t *f  = s->cblktab + bl->fcblknm;
            if ( bl->lcblknm != k || bl->fcblknm <= k || bl->frownum > bl->lrownum ||
                 bl->frownum < f->fcolnum || bl->lrownum > f->lcolnum ||
                 bl->frownum <= s->bloktab[b - 1].lrownum ) {
                return PASTIX_ERR_INTERNAL;
            }
        }
    }
    return PASTIX_SUCCESS;
}

pastix_int_t
pa_symbol_nnz( const symbol_matrix_t *s )
{
    pastix_int_t k, b, nnz = 0;
    for ( k = 0; k < s->cblknbr; k++ ) {
        pastix_int_t nc = s->cblktab[k].lcolnum - s->cblktab[k].fcolnum + 1;
        nnz += ( nc * ( nc + 1 ) ) / 2;
        for ( b = s->cblktab[k].bloknum + 1; b < s->cblktab[k + 1].bloknum; b++ ) {
            nnz += ( s->bloktab[b].lrownum - s->bloktab[b].frownum + 1 ) * nc;
        }
    }
    return nnz;
}

/* Flop formulas of reference common/flops.h:49-255 (real arithmetic). */
#define PA_FMULS_POTRF( n ) ( (double)( n ) * ( ( ( 1. / 6. ) * (double)( n ) + 0.5 ) * (double)( n ) + ( 1. / 3. ) ) )
#define PA_FADDS_POTRF( n ) ( (double)( n ) * ( ( ( 1. / 6. ) * (double)( n ) ) * (double)( n ) - ( 1. / 6. ) ) )
#define PA_FMULS_GETRF( m, n )                                                                         \
    ( ( ( m ) < ( n ) ) ? ( 0.5 * (double)( m ) * ( (double)( m ) * ( (double)( n ) - ( 1. / 3. ) * ( m ) - 1. ) + (double)( n ) ) + ( 2. / 3. ) * ( m ) ) \
                        : ( 0.5 * (double)( n ) * ( (double)( n ) * ( (double)( m ) - ( 1. / 3. ) * ( n ) - 1. ) + (double)( m ) ) + ( 2. / 3. ) * ( n ) ) )
#define PA_FADDS_GETRF( m, n )                                                                         \
    ( ( ( m ) < ( n ) ) ? ( 0.5 * (double)( m ) * ( (double)( m ) * ( (double)( n ) - ( 1. / 3. ) * ( m ) ) - (double)( n ) ) + ( 1. / 6. ) * ( m ) ) \
                        : ( 0.5 * (double)( n ) * ( (double)( n ) * ( (double)( m ) - ( 1. / 3. ) * ( n ) ) - (double)( m ) ) + ( 1. / 6. ) * ( n ) ) )

/* FMULS_SYTRF of the reference (common/flops.h:103): FMULS_POTRF with
 * 10/3 n in place of n/3 (the scaling by D), FADDS as POTRF */
#define PA_FMULS_SYTRF( n ) ( (double)( n ) * ( ( ( 1. / 6. ) * (double)( n ) + 0.5 ) * (double)( n ) + ( 10. / 3. ) ) )

double pa_flops_potrf( double n ) { return PA_FMULS_POTRF( n ) + PA_FADDS_POTRF( n ); }
double pa_flops_sytrf( double n ) { return PA_FMULS_SYTRF( n ) + PA_FADDS_POTRF( n ); }
double pa_flops_getrf( double n ) { return PA_FMULS_GETRF( n, n ) + PA_FADDS_GETRF( n, n ); }
/* right-side TRSM of an M x N block: M*N*(N+1) (FMULS_TRSM counted twice,
 * as in the reference's FLOPS_DTRSM) */
double pa_flops_trsm( double m, double n ) { return m * n * ( n + 1. ); }
double pa_flops_gemm( double m, double n, double k ) { return 2. * m * n * k; }
double pa_flops_syrk( double k, double n ) { return k * n * ( n + 1. ); }

/*
 * Theoretical flops of the factorization (reference symbol/symbol_cost.c:53-95
 * sum1d with symbol/symbol_cost_flops.c flopstable): per cblk, the diagonal
 * factorization, the panel TRSM and the update (SYRK for LL^t, GEMM for LU).
 */
double
pa_symbol_flops( const symbol_matrix_t *s, int factotype )
{
    double total = 0.;
    pastix_int_t k, b;
    for ( k = 0; k < s->cblknbr; k++ ) {
        double N = (double)( s->cblktab[k].lcolnum - s->cblktab[k].fcolnum + 1 ), M = 0.;
        for ( b = s->cblktab[k].bloknum + 1; b < s->cblktab[k + 1].bloknum; b++ ) {
            M += (double)( s->bloktab[b].lrownum - s->bloktab[b].frownum + 1 );
        }
        if ( factotype == PastixFactLU ) {
            total += pa_flops_getrf( N );
            if ( M > 0 ) total += 2. * pa_flops_trsm( M, N ) + pa_flops_gemm( M, M, N );
        }
        else if ( factotype == PastixFactLDLT || factotype == PastixFactLDLH ) {
            /* reference flops_dsytrf_{diag,trsm,update}
             * (symbol/symbol_cost_flops.c:56-158): the D scaling adds N M
             * to the TRSM and M^2 to the update */
            total += pa_flops_sytrf( N );
            if ( M > 0 ) total += pa_flops_trsm( M, N ) + N * M + pa_flops_syrk( N, M ) + M * M;
        }
        else {
            total += pa_flops_potrf( N );
            if ( M > 0 ) total += pa_flops_trsm( M, N ) + pa_flops_syrk( N, M );
        }
    }
    return total;
}
