/*
 * order.c -- ordering structure, closed-form grid nested dissection, level
 * ordering of supernodes, and the elimination-tree fallback for a personal
 * permutation given without a supernode partition.
 */
#include <assert.h>
#include "internal.h"

int
pastixOrderAlloc( pastix_order_t *ordeptr, pastix_int_t vertnbr, pastix_int_t cblknbr )
{
    if ( ordeptr == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    memset( ordeptr, 0, sizeof( pastix_order_t ) );
    ordeptr->vertnbr = vertnbr;
    ordeptr->cblknbr = cblknbr;
    if ( vertnbr > 0 ) {
        ordeptr->permtab = malloc( vertnbr * sizeof( pastix_int_t ) );
        ordeptr->peritab = malloc( vertnbr * sizeof( pastix_int_t ) );
    }
    if ( cblknbr > 0 ) {
        ordeptr->rangtab = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
        ordeptr->treetab = calloc( cblknbr + 1, sizeof( pastix_int_t ) );
    }
    return PASTIX_SUCCESS;
}

int
pastixOrderInit( pastix_order_t *ordeptr, pastix_int_t baseval, pastix_int_t vertnbr,
                 pastix_int_t cblknbr, pastix_int_t *perm, pastix_int_t *invp,
                 pastix_int_t *rang, pastix_int_t *tree )
{
    if ( ordeptr == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    memset( ordeptr, 0, sizeof( pastix_order_t ) );
    ordeptr->baseval = baseval;
    ordeptr->vertnbr = vertnbr;
    ordeptr->cblknbr = cblknbr;
    ordeptr->permtab = perm;
    ordeptr->peritab = invp;
    ordeptr->rangtab = rang;
    ordeptr->treetab = tree;
    return PASTIX_SUCCESS;
}

void
pastixOrderExit( pastix_order_t *ordeptr )
{
    if ( ordeptr == NULL ) {
        return;
    }
    free( ordeptr->permtab );
    free( ordeptr->peritab );
    free( ordeptr->rangtab );
    free( ordeptr->treetab );
    free( ordeptr->selevtx );
    free( ordeptr->sndetab );
    free( ordeptr->peritab_exp );
    memset( ordeptr, 0, sizeof( pastix_order_t ) );
}

void
pastixOrderBase( pastix_order_t *ordeptr, pastix_int_t baseval )
{
    pastix_int_t i, shift;
    if ( ordeptr == NULL ) {
        return;
    }
    shift = baseval - ordeptr->baseval;
    if ( shift == 0 ) {
        return;
    }
    for ( i = 0; i < ordeptr->vertnbr; i++ ) {
        if ( ordeptr->permtab ) ordeptr->permtab[i] += shift;
        if ( ordeptr->peritab ) ordeptr->peritab[i] += shift;
    }
    if ( ordeptr->rangtab ) {
        for ( i = 0; i <= ordeptr->cblknbr; i++ ) ordeptr->rangtab[i] += shift;
    }
    if ( ordeptr->treetab ) {
        for ( i = 0; i < ordeptr->cblknbr; i++ ) {
            if ( ordeptr->treetab[i] >= ordeptr->baseval ) ordeptr->treetab[i] += shift;
        }
    }
    ordeptr->baseval = baseval;
}

int
pastixOrderCopy( pastix_order_t *dst, const pastix_order_t *src )
{
    pastixOrderAlloc( dst, src->vertnbr, src->cblknbr );
    dst->baseval = src->baseval;
    if ( src->permtab ) memcpy( dst->permtab, src->permtab, src->vertnbr * sizeof( pastix_int_t ) );
    if ( src->peritab ) memcpy( dst->peritab, src->peritab, src->vertnbr * sizeof( pastix_int_t ) );
    if ( src->rangtab ) memcpy( dst->rangtab, src->rangtab, ( src->cblknbr + 1 ) * sizeof( pastix_int_t ) );
    if ( src->treetab ) memcpy( dst->treetab, src->treetab, src->cblknbr * sizeof( pastix_int_t ) );
    return PASTIX_SUCCESS;
}

/*
 * Structural check (the invariants of reference order/order_check.c):
 * permtab and peritab are inverse permutations, rangtab is increasing from
 * baseval to vertnbr+baseval, every father index is larger than its son.
 */
int
pastixOrderCheck( const pastix_order_t *o )
{
    pastix_int_t i, n = o->vertnbr, b = o->baseval;
    char *seen;
    if ( n < 0 || o->permtab == NULL || o->peritab == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    seen = calloc( n > 0 ? n : 1, 1 );
    for ( i = 0; i < n; i++ ) {
        pastix_int_t p = o->permtab[i] - b;
        if ( p < 0 || p >= n || seen[p] || ( o->peritab[p] - b ) != i ) {
            free( seen );
            return PASTIX_ERR_BADPARAMETER;
        }
        seen[p] = 1;
    }
    free( seen );
    if ( o->rangtab ) {
        if ( o->rangtab[0] != b || o->rangtab[o->cblknbr] != n + b ) {
            return PASTIX_ERR_BADPARAMETER;
        }
        for ( i = 0; i < o->cblknbr; i++ ) {
            if ( o->rangtab[i + 1] <= o->rangtab[i] ) return PASTIX_ERR_BADPARAMETER;
        }
    }
    if ( o->treetab ) {
        for ( i = 0; i < o->cblknbr; i++ ) {
            pastix_int_t f = o->treetab[i];
            if ( f != -1 && ( f - b <= i || f - b >= o->cblknbr ) ) return PASTIX_ERR_BADPARAMETER;
        }
    }
    return PASTIX_SUCCESS;
}

/* -------------------------------------------------------------------------- */
/* Grid nested dissection                                                     */
/* -------------------------------------------------------------------------- */

/* Number every point of the rectangle [a0,an) x [b0,bn) of a grid plane with
 * decreasing labels.  `base` points at the plane's origin in the numbering
 * array, `sa` and `sb` are the strides of the two in-plane directions
 * (reference order/order_grids.c:58-76). */
static void
grid_number_plane( pastix_int_t *label, pastix_int_t a0, pastix_int_t an,
                   pastix_int_t b0, pastix_int_t bn, pastix_int_t *next,
                   pastix_int_t sa, pastix_int_t sb )
{
    pastix_int_t a, b;
    for ( a = a0; a < an; a++ ) {
        for ( b = b0; b < bn; b++ ) {
            label[a * sa + b * sb] = ( *next )--;
        }
    }
}

typedef struct grid_nd_s {
    pastix_int_t *label;    /* label[vertex] = new number                     */
    pastix_int_t *top;      /* per supernode (creation order): largest label  */
    pastix_int_t *depth;    /* per supernode (creation order): tree depth     */
    pastix_int_t  nsn;      /* supernodes created so far                      */
    pastix_int_t  next;     /* next (largest unused) label                    */
    pastix_int_t  sx, sy, sz;
} grid_nd_t;

/*
 * Recursive dissection of the box [x0,xn) x [y0,yn) x [z0,zn) following
 * reference order/order_grids.c:138-290: a box of fewer than 15 points is a
 * leaf supernode; otherwise the middle plane orthogonal to the longest
 * direction (x on ties, then y) is a separator that takes the largest labels,
 * followed by the lower half, then the upper half.
 */
static void
grid_nd_box( grid_nd_t *g, pastix_int_t x0, pastix_int_t xn, pastix_int_t y0, pastix_int_t yn,
             pastix_int_t z0, pastix_int_t zn, pastix_int_t depth )
{
    pastix_int_t nx = xn - x0, ny = yn - y0, nz = zn - z0;
    pastix_int_t sn = g->nsn++;

    g->top[sn]   = g->next;
    g->depth[sn] = depth;

    if ( nx * ny * nz < 15 ) {
        pastix_int_t i, j, k, c = 0;
        for ( i = x0; i < xn; i++ ) {
            for ( j = y0; j < yn; j++ ) {
                for ( k = z0; k < zn; k++ ) {
                    g->label[i * g->sx + j * g->sy + k * g->sz] = g->next - c;
                    c++;
                }
            }
        }
        g->next -= c;
        return;
    }

    if ( ( nz > nx ) && ( nz > ny ) ) {
        pastix_int_t zc = z0 + nz / 2;
        grid_number_plane( g->label + zc * g->sz, x0, xn, y0, yn, &g->next, g->sx, g->sy );
        grid_nd_box( g, x0, xn, y0, yn, z0, zc, depth + 1 );
        grid_nd_box( g, x0, xn, y0, yn, zc + 1, zn, depth + 1 );
    }
    else if ( ny > nx ) {
        pastix_int_t yc = y0 + ny / 2;
        grid_number_plane( g->label + yc * g->sy, x0, xn, z0, zn, &g->next, g->sx, g->sz );
        grid_nd_box( g, x0, xn, y0, yc, z0, zn, depth + 1 );
        grid_nd_box( g, x0, xn, yc + 1, yn, z0, zn, depth + 1 );
    }
    else {
        pastix_int_t xc = x0 + nx / 2;
        grid_number_plane( g->label + xc * g->sx, y0, yn, z0, zn, &g->next, g->sy, g->sz );
        grid_nd_box( g, x0, xc, y0, yn, z0, zn, depth + 1 );
        grid_nd_box( g, xc + 1, xn, y0, yn, z0, zn, depth + 1 );
    }
}

int
pastixOrderGrid( pastix_order_t **myorder, pastix_int_t nx, pastix_int_t ny, pastix_int_t nz )
{
    pastix_order_t *o;
    pastix_int_t    n = nx * ny * nz, i, c, cblknbr;
    grid_nd_t       g;

    if ( myorder == NULL || *myorder == NULL || nx < 1 || ny < 1 || nz < 1 ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    o = *myorder;
    pastixOrderAlloc( o, n, n );

    g.label = o->permtab;
    g.top   = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    g.depth = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    g.nsn   = 0;
    g.next  = n - 1;
    g.sx    = 1;
    g.sy    = nx;
    g.sz    = nx * ny;
    grid_nd_box( &g, 0, nx, 0, ny, 0, nz, 1 );
    cblknbr = g.nsn;

    for ( i = 0; i < n; i++ ) {
        o->peritab[o->permtab[i]] = i;
    }

    /* Supernodes were created from the largest labels down: reverse them so
     * that supernode c covers [rangtab[c], rangtab[c+1]) in increasing order. */
    o->cblknbr    = cblknbr;
    o->rangtab[0] = 0;
    for ( c = 0; c < cblknbr; c++ ) {
        o->rangtab[c + 1] = g.top[cblknbr - 1 - c] + 1;
        o->treetab[c]     = g.depth[cblknbr - 1 - c];
    }
    /* Father of c: the first later supernode that is shallower
     * (reference order/order_grids.c:371-380). */
    for ( c = 0; c + 1 < cblknbr; c++ ) {
        pastix_int_t d;
        for ( d = c + 1; d < cblknbr; d++ ) {
            if ( o->treetab[d] < o->treetab[c] ) {
                o->treetab[c] = d;
                break;
            }
        }
    }
    o->treetab[cblknbr - 1] = -1;

    free( g.top );
    free( g.depth );
    o->rangtab = realloc( o->rangtab, ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    o->treetab = realloc( o->treetab, ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    return PASTIX_SUCCESS;
}

/* -------------------------------------------------------------------------- */
/* Level ordering of the supernodes                                           */
/* -------------------------------------------------------------------------- */

/* Children lists of a forest given by father[] (-1 = root), children in
 * increasing index order; index `n` collects the roots. */
typedef struct pa_forest_s {
    pastix_int_t  n;
    pastix_int_t *father;   /* n, may be rewritten                      */
    pastix_int_t *nsons;    /* n+1 (last = virtual root)                */
    pastix_int_t *fson;     /* n+2 prefix offsets                       */
    pastix_int_t *sons;     /* n                                        */
} pa_forest_t;

static void
forest_init( pa_forest_t *f, pastix_int_t n, const pastix_int_t *father )
{
    pastix_int_t i, *pos;
    f->n      = n;
    f->father = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    f->nsons  = calloc( n + 1, sizeof( pastix_int_t ) );
    f->fson   = calloc( n + 2, sizeof( pastix_int_t ) );
    f->sons   = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    for ( i = 0; i < n; i++ ) {
        pastix_int_t fa = father[i];
        f->father[i] = fa;
        f->nsons[( fa < 0 ) ? n : fa]++;
    }
    for ( i = 0; i <= n; i++ ) {
        f->fson[i + 1] = f->fson[i] + f->nsons[i];
    }
    pos = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    memcpy( pos, f->fson, ( n + 1 ) * sizeof( pastix_int_t ) );
    for ( i = 0; i < n; i++ ) {
        pastix_int_t fa = ( f->father[i] < 0 ) ? n : f->father[i];
        f->sons[pos[fa]++] = i;
    }
    free( pos );
}

static void
forest_exit( pa_forest_t *f )
{
    free( f->father );
    free( f->nsons );
    free( f->fson );
    free( f->sons );
}

#define FSON( f, node, i ) ( ( f )->sons[( f )->fson[( node ) < 0 ? ( f )->n : ( node )] + ( i )] )
#define NSONS( f, node )   ( ( f )->nsons[( node ) < 0 ? ( f )->n : ( node )] )

/*
 * Reorder the supernodes breadth-first from the roots, 2D candidates first,
 * and renumber so that deeper levels come first (reference
 * order/order_apply_level_order.c:48-330, with the default
 * PASTIX_BLEND_DEEPEST_DISTRIB marking of 2D nodes, CMakeLists.txt:205).
 */
int
pa_order_apply_level_order( pastix_order_t *order, pastix_int_t level_tasks2d,
                            pastix_int_t width_tasks2d )
{
    pastix_int_t  cblknbr = order->cblknbr, n = order->vertnbr;
    pastix_int_t *old_rang, *old_peri, *sorted, *newfather;
    pastix_int_t  i, s, node;
    pa_forest_t   et;

    if ( cblknbr <= 1 ) {
        return PASTIX_SUCCESS;
    }
    forest_init( &et, cblknbr, order->treetab );
    sorted    = malloc( cblknbr * sizeof( pastix_int_t ) );
    newfather = malloc( cblknbr * sizeof( pastix_int_t ) );
    for ( i = 0; i < cblknbr; i++ ) newfather[i] = et.father[i];

    if ( level_tasks2d < 0 ) {
        int8_t      *is2d = calloc( cblknbr, 1 );
        pastix_int_t tot2d = 0, pos2d, pos1d;

        for ( node = 0; node < cblknbr; node++ ) {
            pastix_int_t size = order->rangtab[node + 1] - order->rangtab[node];
            pastix_int_t fa   = et.father[node];
            if ( is2d[node] ) {
                continue;
            }
            if ( size >= width_tasks2d ) {
                for ( i = 0; i < NSONS( &et, fa ); i++ ) {
                    s = FSON( &et, fa, i );
                    if ( !is2d[s] ) { is2d[s] = 1; tot2d++; }
                }
                while ( ( fa != -1 ) && !is2d[fa] ) {
                    fa = et.father[fa];
                    for ( i = 0; i < NSONS( &et, fa ); i++ ) {
                        s = FSON( &et, fa, i );
                        if ( !is2d[s] ) { is2d[s] = 1; tot2d++; }
                    }
                }
            }
        }

        pos2d = 0;
        pos1d = tot2d;
        for ( i = 0; i < NSONS( &et, -1 ); i++ ) {
            node = FSON( &et, -1, i );
            if ( is2d[node] ) sorted[pos2d++] = node;
            else              sorted[pos1d++] = node;
        }
        for ( i = 0; i < cblknbr; i++ ) {
            pastix_int_t ns, n2 = 0, n1, c2 = 0, c1 = 0;
            node = sorted[i];
            ns   = NSONS( &et, node );
            for ( s = 0; s < ns; s++ ) {
                if ( is2d[FSON( &et, node, s )] ) n2++;
            }
            n1 = ns - n2;
            /* sons are placed in reverse order to keep the original order
             * between brothers once the list is reversed below */
            for ( s = 0; s < ns; s++ ) {
                pastix_int_t son = FSON( &et, node, s );
                if ( is2d[son] ) { sorted[pos2d + n2 - c2 - 1] = son; c2++; }
                else             { sorted[pos1d + n1 - c1 - 1] = son; c1++; }
                newfather[son] = cblknbr - i - 1;
            }
            pos2d += n2;
            pos1d += n1;
        }
        free( is2d );
    }
    else {
        pastix_int_t pos = NSONS( &et, -1 );
        for ( i = 0; i < pos; i++ ) sorted[i] = FSON( &et, -1, i );
        for ( i = 0; i < cblknbr; i++ ) {
            pastix_int_t ns;
            node = sorted[i];
            ns   = NSONS( &et, node );
            for ( s = 0; s < ns; s++ ) {
                pastix_int_t son = FSON( &et, node, s );
                sorted[pos + ns - 1 - s] = son;
                newfather[son] = cblknbr - i - 1;
            }
            pos += ns;
        }
    }

    old_rang = malloc( ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    old_peri = malloc( n * sizeof( pastix_int_t ) );
    memcpy( old_rang, order->rangtab, ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    memcpy( old_peri, order->peritab, n * sizeof( pastix_int_t ) );

    order->rangtab[0] = 0;
    for ( i = 0; i < cblknbr; i++ ) {
        pastix_int_t size, nf, of;
        node = sorted[cblknbr - i - 1];
        size = old_rang[node + 1] - old_rang[node];
        of   = old_rang[node];
        nf   = order->rangtab[i];
        order->rangtab[i + 1] = nf + size;
        order->treetab[i]     = newfather[node];
        memcpy( order->peritab + nf, old_peri + of, size * sizeof( pastix_int_t ) );
    }
    for ( i = 0; i < n; i++ ) {
        order->permtab[order->peritab[i]] = i;
    }

    free( old_rang );
    free( old_peri );
    free( sorted );
    free( newfather );
    forest_exit( &et );
    return PASTIX_SUCCESS;
}

/* -------------------------------------------------------------------------- */
/* Ordering files (the reference's text format, order/order_io.c:30-255):      */
/*   version (0 or 1), cblknbr, vertnbr, rangtab[cblknbr+1], permtab[vertnbr], */
/*   and with version 1 treetab[cblknbr]; integers separated by white space.   */
/* The file is named by PASTIX_FILE_ORDER, else "ordername" (load) or         */
/* "ordergen" (save) in the current directory.                                 */
/* -------------------------------------------------------------------------- */

static int
pa_read_int( FILE *f, pastix_int_t *v )
{
    long long x;
    if ( fscanf( f, "%lld", &x ) != 1 ) return 0;
    *v = (pastix_int_t)x;
    return 1;
}

int
pastixOrderLoad( const pastix_data_t *pastix_data, pastix_order_t *ordeptr )
{
    const char  *name = getenv( "PASTIX_FILE_ORDER" );
    FILE        *f;
    pastix_int_t vers, cblknbr, vertnbr, i;
    int          ok;
    (void)pastix_data;
    if ( ordeptr == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( name == NULL ) name = "ordername";
    f = fopen( name, "r" );
    if ( f == NULL ) {
        pa_error( "pastixOrderLoad: cannot open %s", name );
        return PASTIX_ERR_FILE;
    }
    if ( pa_read_int( f, &vers ) + pa_read_int( f, &cblknbr ) + pa_read_int( f, &vertnbr ) != 3 ||
         ( vers != 0 && vers != 1 ) || cblknbr > vertnbr || cblknbr < 0 || vertnbr < 0 ) {
        pa_error( "pastixOrderLoad: bad header in %s", name );
        fclose( f );
        return PASTIX_ERR_FILE;
    }
    pastixOrderAlloc( ordeptr, vertnbr, cblknbr );
    ok = 1;
    for ( i = 0; ok && i <= cblknbr; i++ ) ok = pa_read_int( f, &ordeptr->rangtab[i] );
    for ( i = 0; ok && i < vertnbr; i++ ) ok = pa_read_int( f, &ordeptr->permtab[i] );
    if ( vers == 1 ) {
        for ( i = 0; ok && i < cblknbr; i++ ) ok = pa_read_int( f, &ordeptr->treetab[i] );
    }
    else {
        free( ordeptr->treetab );
        ordeptr->treetab = NULL;
    }
    fclose( f );
    if ( !ok ) {
        pa_error( "pastixOrderLoad: truncated file %s", name );
        pastixOrderExit( ordeptr );
        return PASTIX_ERR_FILE;
    }
    ordeptr->baseval = ordeptr->rangtab[0];
    for ( i = 0; i < vertnbr; i++ ) {
        pastix_int_t p = ordeptr->permtab[i] - ordeptr->baseval;
        if ( p < 0 || p >= vertnbr ) {
            pa_error( "pastixOrderLoad: permutation out of range in %s", name );
            pastixOrderExit( ordeptr );
            return PASTIX_ERR_FILE;
        }
        ordeptr->peritab[p] = i + ordeptr->baseval;
    }
    return PASTIX_SUCCESS;
}

int
pastixOrderSave( pastix_data_t *pastix_data, const pastix_order_t *ordeptr )
{
    const char  *name = getenv( "PASTIX_FILE_ORDER" );
    FILE        *f;
    pastix_int_t i;
    int          ok = 1;
    (void)pastix_data;
    if ( ordeptr == NULL ) return PASTIX_ERR_BADPARAMETER;
    if ( ordeptr->permtab == NULL || ordeptr->rangtab == NULL || ordeptr->treetab == NULL ) {
        pa_error( "pastixOrderSave: the ordering needs permtab, rangtab and treetab" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( name == NULL ) name = "ordergen";
    f = fopen( name, "w" );
    if ( f == NULL ) {
        pa_error( "pastixOrderSave: cannot open %s", name );
        return PASTIX_ERR_FILE;
    }
    ok = fprintf( f, "1\n%ld\t%ld\n", (long)ordeptr->cblknbr, (long)ordeptr->vertnbr ) > 0;
    for ( i = 0; ok && i <= ordeptr->cblknbr; i++ )
        ok = fprintf( f, "%ld%c", (long)ordeptr->rangtab[i], ( ( i & 7 ) == 7 || i == ordeptr->cblknbr ) ? '\n' : '\t' ) > 0;
    for ( i = 0; ok && i < ordeptr->vertnbr; i++ )
        ok = fprintf( f, "%ld%c", (long)ordeptr->permtab[i], ( ( i & 7 ) == 7 || i + 1 == ordeptr->vertnbr ) ? '\n' : '\t' ) > 0;
    for ( i = 0; ok && i < ordeptr->cblknbr; i++ )
        ok = fprintf( f, "%ld%c", (long)ordeptr->treetab[i], ( ( i & 7 ) == 7 || i + 1 == ordeptr->cblknbr ) ? '\n' : '\t' ) > 0;
    if ( fclose( f ) != 0 ) ok = 0;
    if ( !ok ) {
        pa_error( "pastixOrderSave: write error on %s", name );
        return PASTIX_ERR_FILE;
    }
    return PASTIX_SUCCESS;
}
