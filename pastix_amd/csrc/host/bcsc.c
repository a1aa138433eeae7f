/*
 * bcsc.c -- the user matrix permuted to the factorization numbering, fully
 * expanded (both triangles), columns grouped by cblk in increasing order
 * (reference bcsc/bcsc.c:1899 bcscInit, bcsc/bcsc_zinit.c).
 *
 * The pattern is the symmetric union of A and A^T (explicit zeros where only
 * one side is stored), so Lvalues = A(row,col) and Uvalues = A(col,row) share
 * one row index array; Uvalues is what the reference fills U^T from
 * (kernels/cpucblk_zinit.c:395-480).
 */
#include <math.h>
#include "internal.h"

/* one coefficient of the expanded pattern before the per-column sort: row,
 * A(row,col), A(col,row), and the user entries they come from (-1: none) */
typedef struct { pastix_int_t r; double v, vt; pastix_int_t sv, svt; } pa_entry_t;

static int
cmp_entry( const void *a, const void *b )
{
    pastix_int_t x = ( (const pa_entry_t *)a )->r, y = ( (const pa_entry_t *)b )->r;
    return ( x > y ) - ( x < y );
}

int
pa_bcsc_init( pa_bcsc_t *bcsc, const spmatrix_t *spm, const pastix_order_t *order )
{
    return pa_bcsc_build( bcsc, spm, order, NULL );
}

void
pa_valmap_exit( pa_valmap_t *map )
{
    if ( map ) {
        free( map->Lptr ); free( map->Lsrc ); free( map->Uptr ); free( map->Usrc );
        memset( map, 0, sizeof( *map ) );
    }
}

int
pa_bcsc_build( pa_bcsc_t *bcsc, const spmatrix_t *spm, const pastix_order_t *order, pa_valmap_t *map )
{
    pastix_int_t n = spm->n, b = spm->baseval, i, e, total, *cnt, *pos;
    const pastix_int_t *ptr = spm->colptr, *idx = spm->rowptr;
    const double *val = (const double *)spm->values;
    const float  *valf = (const float *)spm->values;
    int is_float = ( spm->flttype == SpmFloat );
    int sym      = ( spm->mtxtype != SpmGeneral );
    int csr      = ( spm->fmttype == SpmCSR );
    pa_entry_t *ent;

    if ( spm->fmttype != SpmCSC && spm->fmttype != SpmCSR ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( spm->flttype != SpmDouble && spm->flttype != SpmFloat ) {
        pa_error( "bcsc: only real (float/double) matrices are supported by this build" );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    if ( csr ) {
        ptr = spm->rowptr;
        idx = spm->colptr;
    }

    /* Each stored entry (i,j,v) lands in column j (value v) and in column i
     * (as the transpose); both are needed for the union pattern. */
    cnt = calloc( n + 1, sizeof( pastix_int_t ) );
    for ( i = 0; i < n; i++ ) {
        for ( e = ptr[i] - b; e < ptr[i + 1] - b; e++ ) {
            pastix_int_t r = idx[e] - b;
            cnt[order->permtab[i]]++;
            if ( r != i ) cnt[order->permtab[r]]++;
        }
    }
    pos = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    pos[0] = 0;
    for ( i = 0; i < n; i++ ) pos[i + 1] = pos[i] + cnt[i];
    total = pos[n];
    ent = malloc( ( total + 1 ) * sizeof( pa_entry_t ) );
    memcpy( cnt, pos, n * sizeof( pastix_int_t ) );
    for ( i = 0; i < n; i++ ) {
        for ( e = ptr[i] - b; e < ptr[i + 1] - b; e++ ) {
            pastix_int_t r = idx[e] - b;
            double v = is_float ? (double)valf[e] : val[e];
            /* stored coordinate: CSC -> (row r, col i); CSR -> (row i, col r) */
            pastix_int_t row = csr ? i : r, col = csr ? r : i;
            pastix_int_t prow = order->permtab[row], pcol = order->permtab[col];
            pa_entry_t *x = ent + cnt[pcol]++;
            x->r   = prow;
            x->v   = v;   /* A(row, col)                    */
            x->vt  = 0.;  /* A(col, row) unknown from here  */
            x->sv  = e;
            x->svt = -1;
            if ( row != col ) {
                pa_entry_t *y = ent + cnt[prow]++;
                y->r   = pcol;
                y->v   = sym ? v : 0.; /* A(col,row) = v if symmetric      */
                y->vt  = v;            /* transpose view of column prow    */
                y->sv  = sym ? e : -1;
                y->svt = e;
            }
            else {
                x->vt  = v;
                x->svt = e;
            }
            if ( sym && row != col ) {
                x->vt  = v;
                x->svt = e;
            }
        }
    }

    /* sort rows in each column, merge duplicates by summation */
    bcsc->n       = n;
    bcsc->mtxtype = spm->mtxtype;
    bcsc->colptr  = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    bcsc->rowtab  = malloc( ( total + 1 ) * sizeof( pastix_int_t ) );
    bcsc->Lvalues = malloc( ( total + 1 ) * sizeof( double ) );
    bcsc->Uvalues = malloc( ( total + 1 ) * sizeof( double ) );
    if ( map ) {
        memset( map, 0, sizeof( *map ) );
        map->nuser = ptr[n] - b;
        map->Lptr  = malloc( ( total + 1 ) * sizeof( pastix_int_t ) );
        map->Uptr  = malloc( ( total + 1 ) * sizeof( pastix_int_t ) );
        map->Lsrc  = malloc( ( total + 1 ) * sizeof( pastix_int_t ) );
        map->Usrc  = malloc( ( total + 1 ) * sizeof( pastix_int_t ) );
    }
    {
        pastix_int_t w = 0, nl = 0, nu = 0;
        double *rowsum = calloc( n + 1, sizeof( double ) );
        double  frob = 0.;
        for ( i = 0; i < n; i++ ) {
            pastix_int_t s = pos[i], t = pos[i + 1];
            qsort( ent + s, t - s, sizeof( pa_entry_t ), cmp_entry );
            bcsc->colptr[i] = w;
            for ( e = s; e < t; e++ ) {
                if ( w > bcsc->colptr[i] && bcsc->rowtab[w - 1] == ent[e].r ) {
                    if ( sym ) {
                        /* symmetric input storing both triangles: keep one copy */
                        continue;
                    }
                    bcsc->Lvalues[w - 1] += ent[e].v;
                    bcsc->Uvalues[w - 1] += ent[e].vt;
                }
                else {
                    if ( map ) {
                        map->Lptr[w] = nl;
                        map->Uptr[w] = nu;
                    }
                    bcsc->rowtab[w]  = ent[e].r;
                    bcsc->Lvalues[w] = ent[e].v;
                    bcsc->Uvalues[w] = ent[e].vt;
                    w++;
                }
                if ( map ) {
                    /* the entries summed into position w-1, by source */
                    if ( ent[e].sv >= 0 ) map->Lsrc[nl++] = ent[e].sv;
                    if ( ent[e].svt >= 0 ) map->Usrc[nu++] = ent[e].svt;
                }
            }
        }
        bcsc->colptr[n] = w;
        if ( map ) {
            map->nnz     = w;
            map->Lptr[w] = nl;
            map->Uptr[w] = nu;
        }
        bcsc->nnz       = w;
        for ( i = 0; i < n; i++ ) {
            for ( e = bcsc->colptr[i]; e < bcsc->colptr[i + 1]; e++ ) {
                frob += bcsc->Lvalues[e] * bcsc->Lvalues[e];
                rowsum[bcsc->rowtab[e]] += fabs( bcsc->Lvalues[e] );
            }
        }
        bcsc->normF   = sqrt( frob );
        bcsc->normInf = 0.;
        for ( i = 0; i < n; i++ ) bcsc->normInf = fmax( bcsc->normInf, rowsum[i] );
        free( rowsum );
    }
    free( ent );
    free( pos );
    free( cnt );
    return PASTIX_SUCCESS;
}

void
pa_bcsc_exit( pa_bcsc_t *bcsc )
{
    if ( bcsc ) {
        free( bcsc->colptr );
        free( bcsc->rowtab );
        free( bcsc->Lvalues );
        free( bcsc->Uvalues );
        memset( bcsc, 0, sizeof( *bcsc ) );
    }
}
