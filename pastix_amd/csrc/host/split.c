/*
 * split.c -- splitting of wide supernodes into column blocks of at most
 * IPARM_MAX_BLOCKSIZE columns (reference blend/splitsymbol.c:79-420 and the
 * extra-cblk merge of blend/extracblk.c).
 */
#include <assert.h>
#include "internal.h"

/* Number of pieces for a cblk of `width` columns
 * (reference blend/splitsymbol.c:79-101). */
static pastix_int_t
split_count( pastix_int_t width, pastix_int_t blmin, pastix_int_t blmax )
{
    pastix_int_t nseq;
    if ( width <= blmax ) {
        return 1;
    }
    nseq = pa_iceil( width, blmax );
    if ( nseq > 1 && ( width / nseq ) < blmin ) {
        nseq--;
    }
    return nseq;
}

/*
 * Choose the last column (relative to the piece start) of the next piece:
 * around step-1, within +-percent, the row boundary crossed by the fewest
 * blocks (reference blend/splitsymbol.c:148-185).  `cut[i]` is the number of
 * blocks containing both rows start+i and start+i+1.
 */
static pastix_int_t
split_smallest( const pastix_int_t *cut, pastix_int_t step, pastix_int_t min,
                pastix_int_t max, pastix_int_t percent )
{
    pastix_int_t limit = pa_iceil( step * percent, 100 );
    pastix_int_t lavg, lmin, lmax, best, nb, i;
    if ( step >= max ) {
        return max - 1;
    }
    lavg = step - 1;
    lmin = pa_imax( lavg - limit - 1, min );
    lmax = pa_imin( lavg + limit + 1, max );
    best = lavg;
    nb   = cut[lavg];
    for ( i = lavg + 1; i < lmax; i++ ) {
        if ( cut[i] < nb ) { best = i; nb = cut[i]; }
    }
    for ( i = lavg - 1; i > lmin; i-- ) {
        if ( cut[i] < nb ) { best = i; nb = cut[i]; }
    }
    return best;
}

/* Variant selected by PASTIX_BLEND_SPLIT=UPPER (reference :205-240). */
static pastix_int_t
split_smallest_upper( const pastix_int_t *cut, pastix_int_t step, pastix_int_t min,
                      pastix_int_t max, pastix_int_t percent )
{
    pastix_int_t limit = pa_iceil( step * percent, 100 );
    pastix_int_t lavg, lmin, lmax, best, nb, i;
    if ( step >= max ) {
        return max - 1;
    }
    lavg = step - 1;
    lmin = pa_imax( lavg - limit, min );
    lmax = pa_imin( lavg + limit + 1, max );
    best = lmin;
    nb   = cut[best];
    for ( i = lmin; i < lmax; i++ ) {
        if ( cut[i] <= nb ) { best = i; nb = cut[i]; }
    }
    return best;
}

/* Variant selected by PASTIX_BLEND_SPLIT=CONSTANT (reference :266-280). */
static pastix_int_t
split_constant( const pastix_int_t *cut, pastix_int_t step, pastix_int_t min,
                pastix_int_t max, pastix_int_t percent )
{
    (void)cut; (void)min; (void)percent;
    return ( step >= max ) ? max - 1 : step - 1;
}

typedef pastix_int_t ( *split_fn_t )( const pastix_int_t *, pastix_int_t, pastix_int_t,
                                      pastix_int_t, pastix_int_t );

int
pa_split_symbol( symbol_matrix_t *symb, pastix_int_t blmin, pastix_int_t blmax )
{
    pastix_int_t  cblknbr = symb->cblknbr, n = symb->nodenbr, k, b, r;
    pastix_int_t *npieces, *pstart, *pfcol, *plcol, totpieces = 0, cap;
    pastix_int_t *cut = NULL, percent = 10;
    split_fn_t    fn = split_smallest;
    const char   *env;

    env = getenv( "PASTIX_BLEND_SPLIT" );
    if ( env && !strcmp( env, "CONSTANT" ) ) fn = split_constant;
    else if ( env && !strcmp( env, "UPPER" ) ) fn = split_smallest_upper;
    env = getenv( "PASTIX_BLEND_SPLIT_AUTORIZED_PERCENTAGE" );
    if ( env ) percent = atol( env );

    npieces = malloc( cblknbr * sizeof( pastix_int_t ) );
    pstart  = malloc( ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    cap     = cblknbr + 16;
    pfcol   = malloc( cap * sizeof( pastix_int_t ) );
    plcol   = malloc( cap * sizeof( pastix_int_t ) );

    /* 1. choose the pieces of every cblk, in cblk order */
    for ( k = 0; k < cblknbr; k++ ) {
        pastix_int_t fcol = symb->cblktab[k].fcolnum, lcol = symb->cblktab[k].lcolnum;
        pastix_int_t width = lcol - fcol + 1, nseq = split_count( width, blmin, blmax ), np = 0;

        pstart[k] = totpieces;
        if ( totpieces + width / ( blmin > 0 ? blmin : 1 ) + 2 > cap ) {
            cap   = 2 * ( totpieces + width / ( blmin > 0 ? blmin : 1 ) + 2 );
            pfcol = realloc( pfcol, cap * sizeof( pastix_int_t ) );
            plcol = realloc( plcol, cap * sizeof( pastix_int_t ) );
        }
        if ( nseq <= 1 ) {
            pfcol[totpieces] = fcol;
            plcol[totpieces] = lcol;
            totpieces++;
            npieces[k] = 1;
            continue;
        }
        if ( cut == NULL ) {
            /* blocks crossing each row boundary (line, line+1), counted from
             * the first split cblk on (reference :103-141) */
            cut = calloc( n + 1, sizeof( pastix_int_t ) );
            for ( b = 0; b < symb->bloknbr; b++ ) {
                const symbol_blok_t *bl = symb->bloktab + b;
                if ( bl->lrownum < fcol ) continue;
                for ( r = pa_imax( bl->frownum, fcol ); r < bl->lrownum; r++ ) cut[r]++;
            }
        }
        while ( fcol <= lcol ) {
            pastix_int_t step, lwidth, last;
            nseq = split_count( width, blmin, blmax );
            if ( nseq == 1 ) {
                pfcol[totpieces] = fcol;
                plcol[totpieces] = lcol;
                totpieces++;
                np++;
                break;
            }
            step   = pa_iceil( width, nseq );
            lwidth = width - blmin;
            last   = fcol + fn( cut + fcol, step, blmin, lwidth, percent );
            pfcol[totpieces] = fcol;
            plcol[totpieces] = last;
            totpieces++;
            np++;
            width -= last - fcol + 1;
            fcol = last + 1;
        }
        npieces[k] = np;
        /* every off-diagonal block is now replicated np times */
        for ( b = symb->cblktab[k].bloknum + 1; b < symb->cblktab[k + 1].bloknum; b++ ) {
            for ( r = symb->bloktab[b].frownum; r < symb->bloktab[b].lrownum; r++ ) {
                cut[r] += np - 1;
            }
        }
    }
    pstart[cblknbr] = totpieces;
    free( cut );

    if ( totpieces == cblknbr ) {
        free( npieces ); free( pstart ); free( pfcol ); free( plcol );
        return PASTIX_SUCCESS;
    }

    /* 2. rebuild the symbol matrix on the finer partition */
    {
        pastix_int_t  *col2piece = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
        symbol_cblk_t *ncblk     = calloc( totpieces + 1, sizeof( symbol_cblk_t ) );
        pastix_int_t   nbcap     = symb->bloknbr * 2 + totpieces * 4 + 16, nb = 0, p, q;
        symbol_blok_t *nblok     = malloc( nbcap * sizeof( symbol_blok_t ) );

        for ( p = 0; p < totpieces; p++ ) {
            for ( r = pfcol[p]; r <= plcol[p]; r++ ) col2piece[r] = p;
        }
#define PA_PUSH_BLOK( fr, lr, lc, fc )                                           \
        do {                                                                     \
            if ( nb + 1 >= nbcap ) {                                             \
                nbcap = 2 * nbcap;                                               \
                nblok = realloc( nblok, nbcap * sizeof( symbol_blok_t ) );       \
            }                                                                    \
            nblok[nb].frownum = ( fr ); nblok[nb].lrownum = ( lr );              \
            nblok[nb].lcblknm = ( lc ); nblok[nb].fcblknm = ( fc ); nb++;        \
        } while ( 0 )

        for ( k = 0; k < cblknbr; k++ ) {
            for ( p = pstart[k]; p < pstart[k + 1]; p++ ) {
                ncblk[p].fcolnum = pfcol[p];
                ncblk[p].lcolnum = plcol[p];
                ncblk[p].bloknum = nb;
                ncblk[p].selevtx = symb->cblktab[k].selevtx;
                /* diagonal block, then the later pieces of the same cblk */
                for ( q = p; q < pstart[k + 1]; q++ ) {
                    PA_PUSH_BLOK( pfcol[q], plcol[q], p, q );
                }
                /* original off-diagonal blocks, cut at the pieces they face */
                for ( b = symb->cblktab[k].bloknum + 1; b < symb->cblktab[k + 1].bloknum; b++ ) {
                    pastix_int_t fr = symb->bloktab[b].frownum, lr = symb->bloktab[b].lrownum;
                    while ( fr <= lr ) {
                        pastix_int_t fp = col2piece[fr];
                        pastix_int_t end = pa_imin( lr, plcol[fp] );
                        PA_PUSH_BLOK( fr, end, p, fp );
                        fr = end + 1;
                    }
                }
            }
        }
#undef PA_PUSH_BLOK
        ncblk[totpieces].fcolnum = n;
        ncblk[totpieces].lcolnum = n;
        ncblk[totpieces].bloknum = nb;
        free( symb->cblktab );
        free( symb->bloktab );
        free( symb->browtab );
        symb->browtab = NULL;
        symb->cblktab = ncblk;
        symb->bloktab = realloc( nblok, ( nb + 1 ) * sizeof( symbol_blok_t ) );
        memset( symb->bloktab + nb, 0, sizeof( symbol_blok_t ) );
        symb->cblknbr = totpieces;
        symb->bloknbr = nb;
        free( col2piece );
    }
    free( npieces );
    free( pstart );
    free( pfcol );
    free( plcol );
    return PASTIX_SUCCESS;
}
