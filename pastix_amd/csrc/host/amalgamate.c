/*
 * amalgamate.c -- supernode partition of an ordering given as a bare
 * permutation (a personal ordering, an ordering file without its tree, the
 * legacy interface's perm/invp):
 *
 *   1. fundamental supernodes of the elimination tree, after a postorder of
 *      the tree (reference order/order_find_supernodes.c:360-560);
 *   2. supernodal symbolic factorization of the partition and greedy
 *      amalgamation of son/father supernodes under a fill budget, first by
 *      least added fill, then (when the BLAS ratio exceeds the structural
 *      one) by the gain of the reference's BLAS time model (reference
 *      order/order_amalgamate.c:77-190, symbol/fax_csr.c:70-330,
 *      symbol/fax_csr_direct.c, symbol/fax_csr_amalgamate.c:160-860).
 *
 * The elimination tree is unique and the supernode sets follow from it, but
 * the postorder, the order in which candidate merges leave the priority
 * queue and the renumbering of the merged partition decide the final
 * permutation; those steps follow the reference's so that the same input
 * gives the same partition.  Row sets are kept sorted and merged linearly.
 */
#include <math.h>
#include "internal.h"

/* ------------------------------------------------------------------------- */
/* elimination tree, postorder, fundamental supernodes                        */
/* ------------------------------------------------------------------------- */

/* father[v] of every vertex v (original numbering) in the elimination tree
 * of the symmetric graph permuted by perm; roots are their own father.
 * Liu's algorithm with path compression over the vertices in elimination
 * order. */
static void
pa_etree( pastix_int_t n, const pastix_int_t *colptr, const pastix_int_t *rows, const pastix_int_t *perm,
          const pastix_int_t *invp, pastix_int_t *father )
{
    pastix_int_t *anc = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t  k, e;
    for ( k = 0; k < n; k++ ) {
        const pastix_int_t v = invp[k];
        father[v] = -1;
        anc[v]    = -1;
        for ( e = colptr[v]; e < colptr[v + 1]; e++ ) {
            pastix_int_t u = rows[e];
            if ( perm[u] >= k ) continue;
            /* climb from u to the root of its current subtree, pointing the
             * path at v */
            while ( u != -1 && u != v ) {
                pastix_int_t next = anc[u];
                anc[u] = v;
                if ( next == -1 ) father[u] = v;
                u = next;
            }
        }
    }
    for ( k = 0; k < n; k++ )
        if ( father[k] == -1 ) father[k] = k;
    free( anc );
}

/* size[v] = vertices of the subtree rooted at v; the vertices are visited in
 * elimination order (children before their father) */
static void
pa_subtree_sizes( pastix_int_t n, const pastix_int_t *father, const pastix_int_t *invp, pastix_int_t *size )
{
    pastix_int_t k;
    memset( size, 0, n * sizeof( pastix_int_t ) );
    for ( k = 0; k < n; k++ ) {
        const pastix_int_t v = invp[k];
        size[v] += 1;
        if ( father[v] != v ) size[father[v]] += size[v];
    }
}

/*
 * Postorder of the elimination tree (reference compute_post_order,
 * order/order_find_supernodes.c:99-170): the trees follow each other in the
 * elimination order of their roots; inside a subtree the root comes last and
 * its children's subtrees are laid out below it from the top down in
 * decreasing elimination order.  perm/invp are replaced.
 */
static void
pa_postorder( pastix_int_t n, const pastix_int_t *father, pastix_int_t *perm, pastix_int_t *invp )
{
    pastix_int_t *size = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t *top  = malloc( ( n + 1 ) * sizeof( pastix_int_t ) ); /* next free slot (1-based) below v */
    pastix_int_t *pos  = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t  k, offset = 0;
    pa_subtree_sizes( n, father, invp, size );
    /* each tree occupies a range after the trees of earlier roots */
    for ( k = 0; k < n; k++ ) {
        const pastix_int_t v = invp[k];
        if ( father[v] == v ) {
            top[v] = offset + size[v];
            offset += size[v];
        }
    }
    /* from the last eliminated vertex down: a vertex takes the highest free
     * slot of its father's range, its own subtree the slots below it */
    for ( k = n - 1; k >= 0; k-- ) {
        const pastix_int_t v = invp[k], f = father[v];
        if ( f == v ) {
            pos[v] = top[v];
        }
        else {
            pos[v] = top[f];
            top[f] -= size[v];
        }
        top[v] = pos[v] - 1;
    }
    for ( k = 0; k < n; k++ ) perm[k] = pos[k] - 1;
    for ( k = 0; k < n; k++ ) invp[perm[k]] = k;
    free( size );
    free( top );
    free( pos );
}

/*
 * Fundamental supernodes (reference orderFindSupernodes): after the
 * postorder, column j starts a new supernode when its vertex does not have
 * exactly one child, or when j is a leaf of some row subtree (the previous
 * column with a nonzero in that row is not a descendant of j).  The
 * supernode tree links each supernode to the one holding the father of its
 * last column.  perm/invp are postordered in place; rangtab/treetab are
 * (re)allocated.
 */
int
pa_order_find_supernodes( pastix_order_t *o, pastix_int_t n, const pastix_int_t *colptr, const pastix_int_t *rows )
{
    pastix_int_t *perm = o->permtab, *invp = o->peritab;
    pastix_int_t *father, *size, *nson, *start, *prev, *snode;
    pastix_int_t  j, k, e, nsn;

    father = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    size   = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    nson   = calloc( n + 1, sizeof( pastix_int_t ) );
    start  = calloc( n + 1, sizeof( pastix_int_t ) );
    prev   = calloc( n + 1, sizeof( pastix_int_t ) );
    snode  = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    if ( !father || !size || !nson || !start || !prev || !snode ) {
        free( father ); free( size ); free( nson ); free( start ); free( prev ); free( snode );
        return PASTIX_ERR_ALLOC;
    }

    pa_etree( n, colptr, rows, perm, invp, father );
    pa_postorder( n, father, perm, invp );
    pa_subtree_sizes( n, father, invp, size );

    /* row-subtree leaves: column j has a nonzero in row r > j whose previous
     * nonzero column lies outside j's subtree [j - size + 1, j] */
    for ( j = 0; j < n; j++ ) {
        const pastix_int_t v = invp[j];
        for ( e = colptr[v]; e < colptr[v + 1]; e++ ) {
            const pastix_int_t r = perm[rows[e]];
            if ( r <= j ) continue;
            if ( prev[r] < j - size[v] + 1 ) start[j] = 1;
            prev[r] = j;
        }
    }
    for ( k = 0; k < n; k++ )
        if ( father[k] != k ) nson[father[k]]++;
    for ( k = 0; k < n; k++ )
        if ( nson[k] != 1 ) start[perm[k]] = 1;

    free( o->rangtab );
    free( o->treetab );
    nsn = 0;
    for ( j = 0; j < n; j++ ) nsn += start[j] ? 1 : 0;
    o->rangtab = malloc( ( nsn + 1 ) * sizeof( pastix_int_t ) );
    o->treetab = malloc( ( nsn + 1 ) * sizeof( pastix_int_t ) );
    o->cblknbr = nsn;
    k = 0;
    for ( j = 0; j < n; j++ )
        if ( start[j] ) o->rangtab[k++] = j;
    o->rangtab[nsn] = n;
    for ( k = 0; k < nsn; k++ )
        for ( j = o->rangtab[k]; j < o->rangtab[k + 1]; j++ ) snode[j] = k;
    /* father supernode: the nearest later supernode holding the father of
     * one of k's columns */
    for ( k = 0; k < nsn; k++ ) {
        pastix_int_t best = nsn;
        for ( j = o->rangtab[k]; j < o->rangtab[k + 1]; j++ ) {
            const pastix_int_t d = snode[perm[father[invp[j]]]];
            if ( d > k && d < best ) best = d;
        }
        o->treetab[k] = ( best == nsn ) ? -1 : best;
    }
    free( father ); free( size ); free( nson ); free( start ); free( prev ); free( snode );
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* supernodal row sets                                                        */
/* ------------------------------------------------------------------------- */

/* a growable sorted set of row indices per supernode */
typedef struct {
    pastix_int_t  nnz;
    pastix_int_t *rows;
} pa_rowset_t;

/* out = a U b for sorted sets; returns |out| */
static pastix_int_t
pa_set_union( pastix_int_t na, const pastix_int_t *a, pastix_int_t nb, const pastix_int_t *b, pastix_int_t *out )
{
    pastix_int_t i = 0, j = 0, m = 0;
    while ( i < na && j < nb ) {
        if ( a[i] < b[j] )      out[m++] = a[i++];
        else if ( b[j] < a[i] ) out[m++] = b[j++];
        else { out[m++] = a[i++]; j++; }
    }
    while ( i < na ) out[m++] = a[i++];
    while ( j < nb ) out[m++] = b[j++];
    return m;
}

static int
pa_rowset_assign( pa_rowset_t *s, const pastix_int_t *src, pastix_int_t m )
{
    pastix_int_t *r = malloc( ( m > 0 ? m : 1 ) * sizeof( pastix_int_t ) );
    if ( !r ) return PASTIX_ERR_ALLOC;
    memcpy( r, src, m * sizeof( pastix_int_t ) );
    free( s->rows );
    s->rows = r;
    s->nnz  = m;
    return PASTIX_SUCCESS;
}

static int
pa_cmp_int( const void *a, const void *b )
{
    const pastix_int_t x = *(const pastix_int_t *)a, y = *(const pastix_int_t *)b;
    return ( x > y ) - ( x < y );
}

/*
 * Row sets of the supernodal factor L (reference faxCSRGenPA +
 * faxCSRCblkCompress + faxCSRFactDirect): each supernode starts with its own
 * columns and the rows of A (new numbering) below its columns; the rows below
 * a supernode are then added to its father's set, in supernode order.
 * Returns nnz(L) of the partition, or -1 on allocation failure.
 */
static pastix_int_t
pa_supernodal_fill( const pastix_order_t *o, pastix_int_t n, const pastix_int_t *colptr, const pastix_int_t *rows,
                    pa_rowset_t *L, pastix_int_t *work )
{
    const pastix_int_t  cnbr = o->cblknbr;
    const pastix_int_t *rang = o->rangtab, *tree = o->treetab;
    pastix_int_t       *mark = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t        k, j, e, m, nnz = 0;
    if ( !mark ) return -1;
    for ( j = 0; j < n; j++ ) mark[j] = -1;
    for ( k = 0; k < cnbr; k++ ) {
        m = 0;
        for ( j = rang[k]; j < rang[k + 1]; j++ ) {
            work[m++] = j;
            mark[j]   = k;
        }
        for ( j = rang[k]; j < rang[k + 1]; j++ ) {
            const pastix_int_t v = o->peritab[j];
            for ( e = colptr[v]; e < colptr[v + 1]; e++ ) {
                const pastix_int_t r = o->permtab[rows[e]];
                if ( r > j && mark[r] != k ) {
                    mark[r]   = k;
                    work[m++] = r;
                }
            }
        }
        qsort( work, m, sizeof( pastix_int_t ), pa_cmp_int );
        L[k].rows = NULL;
        if ( pa_rowset_assign( L + k, work, m ) ) { free( mark ); return -1; }
    }
    free( mark );
    /* the rows below each supernode move up to its father */
    for ( k = 0; k < cnbr; k++ ) {
        const pastix_int_t f = tree[k];
        pastix_int_t       i = 0;
        if ( f == k || f <= 0 ) continue;
        while ( i < L[k].nnz && L[k].rows[i] < rang[k + 1] ) i++;
        m = pa_set_union( L[k].nnz - i, L[k].rows + i, L[f].nnz, L[f].rows, work );
        if ( pa_rowset_assign( L + f, work, m ) ) return -1;
    }
    for ( k = 0; k < cnbr; k++ ) {
        const pastix_int_t nc = rang[k + 1] - rang[k], nr = L[k].nnz;
        nnz += nc * ( nc + 1 ) / 2 + nc * ( nr - nc );
    }
    return nnz;
}

/* ------------------------------------------------------------------------- */
/* amalgamation                                                               */
/* ------------------------------------------------------------------------- */

/* Fitted BLAS time model of the reference (blend/perf.h:28-84, "AMD 6180
 * MKL"): the same coefficients, so that the BLAS-driven merges choose as the
 * reference's do. */
static inline double pa_perf_gemm( double i, double j, double k )
{
    return 2.429169e-10 * i * j * k + 2.724804e-10 * i * j + 1.328900e-09 * j * k + 1.148989e-07 * i -
           2.704179e-10 * j + 1.216278e-06;
}
static inline double pa_perf_trsm( double i, double j )
{
    return 2.626177e-10 * i * i * j + 3.976198e-08 * i + 3.255168e-06;
}
static inline double pa_perf_potrf( double i )
{
    return 2.439599e-11 * i * i * i + 1.707006e-08 * i * i - 1.469893e-07 * i + 4.071507e-07;
}
static inline double pa_perf_trsv( double i )
{
    return 3.871789e-10 * i * i + 2.295025e-08 * i + 1.282744e-06;
}
static inline double pa_perf_gemv( double i, double j )
{
    return 6.192657e-10 * i * j - 2.884799e-09 * i + 7.594831e-10 * j + 3.575035e-07;
}

/* modelled time of a supernode of `ncol` columns and row set rows[0..nr)
 * (its own columns first): factorization, or the solve for ILU(k) */
static double
pa_cblk_time( int ilu, pastix_int_t nr, const pastix_int_t *rows, pastix_int_t ncol )
{
    double       cost;
    pastix_int_t i, G, H;
    if ( ilu ) return pa_perf_trsv( (double)ncol ) + pa_perf_gemv( (double)ncol, (double)( nr - ncol ) );
    G    = nr - ncol;
    cost = pa_perf_potrf( (double)ncol ) + pa_perf_trsm( (double)ncol, (double)G );
    /* one GEMM per run of consecutive rows below the diagonal block */
    i = ncol;
    while ( i < nr ) {
        H = 1;
        i++;
        while ( i < nr && rows[i] == rows[i - 1] + 1 ) {
            i++;
            H++;
        }
        cost += pa_perf_gemm( (double)G, (double)H, (double)ncol );
        G -= H;
    }
    return cost;
}

/* fill added by merging son a into father b: each row of one set missing
 * from the other costs the other's column count (a's own diagonal rows,
 * those before b's first row, add nothing) */
static pastix_int_t
pa_merge_fill( const pa_rowset_t *L, pastix_int_t a, pastix_int_t b, const pastix_int_t *w )
{
    const pastix_int_t *ra = L[a].rows, *rb = L[b].rows;
    const pastix_int_t  na = L[a].nnz, nb = L[b].nnz;
    pastix_int_t        ia = 0, ib = 0, fill = 0;
    while ( ia < na && ra[ia] < rb[0] ) ia++;
    while ( ia < na && ib < nb ) {
        if ( ra[ia] < rb[ib] )      { fill += w[b]; ia++; }
        else if ( ra[ia] > rb[ib] ) { fill += w[a]; ib++; }
        else                        { ia++; ib++; }
    }
    fill += w[b] * ( na - ia ) + w[a] * ( nb - ib );
    return fill;
}

/* change of the modelled time when a merges into b (negative: faster) */
static double
pa_merge_gain( int ilu, const pa_rowset_t *L, pastix_int_t a, pastix_int_t b, const pastix_int_t *w,
               pastix_int_t *tmp )
{
    const double       ta = pa_cblk_time( ilu, L[a].nnz, L[a].rows, w[a] );
    const double       tb = pa_cblk_time( ilu, L[b].nnz, L[b].rows, w[b] );
    const pastix_int_t nm = pa_set_union( L[a].nnz, L[a].rows, L[b].nnz, L[b].rows, tmp );
    return pa_cblk_time( ilu, nm, tmp, w[a] + w[b] ) - ta - tb;
}

/* b takes the rows of a; a's set is emptied */
static int
pa_merge_rows( pa_rowset_t *L, pastix_int_t a, pastix_int_t b, pastix_int_t *tmp )
{
    const pastix_int_t m = pa_set_union( L[a].nnz, L[a].rows, L[b].nnz, L[b].rows, tmp );
    free( L[a].rows );
    L[a].rows = NULL;
    L[a].nnz  = 0;
    return pa_rowset_assign( L + b, tmp, m );
}

/*
 * Binary min-heap of (key, element), ties in key keeping the heap's own
 * order: the sift-up moves an entry above its parent only when its key is
 * strictly smaller, the sift-down takes the left child unless the right one
 * is strictly smaller, and moves down while the child is not larger.  These
 * are the reference's rules (kernels/queue.c:137-240), which fix the order
 * in which equal-cost merges happen.
 */
typedef struct {
    double        key;
    pastix_int_t  elt;
} pa_heap_item_t;

typedef struct {
    pa_heap_item_t *a;
    pastix_int_t    used, cap;
} pa_heap_t;

static int
pa_heap_push( pa_heap_t *h, pastix_int_t elt, double key )
{
    pastix_int_t i;
    if ( h->used == h->cap ) {
        pastix_int_t    cap = 2 * h->cap + 1;
        pa_heap_item_t *a   = realloc( h->a, cap * sizeof( pa_heap_item_t ) );
        if ( !a ) return PASTIX_ERR_ALLOC;
        h->a   = a;
        h->cap = cap;
    }
    i = h->used++;
    h->a[i].key = key;
    h->a[i].elt = elt;
    while ( i > 0 ) {
        const pastix_int_t p = ( i + 1 ) / 2 - 1;
        pa_heap_item_t     t;
        if ( !( h->a[i].key < h->a[p].key ) ) break;
        t       = h->a[i];
        h->a[i] = h->a[p];
        h->a[p] = t;
        i       = p;
    }
    return PASTIX_SUCCESS;
}

static pastix_int_t
pa_heap_pop( pa_heap_t *h, double *key )
{
    pastix_int_t i, c, top;
    if ( h->used == 0 ) return -1;
    top  = h->a[0].elt;
    *key = h->a[0].key;
    h->a[0] = h->a[--h->used];
    /* 1-based positions: children of i are 2i and 2i + 1 */
    i = 1;
    while ( i <= h->used / 2 ) {
        pa_heap_item_t t;
        c = ( 2 * i == h->used || h->a[2 * i - 1].key < h->a[2 * i].key ) ? 2 * i : 2 * i + 1;
        if ( h->a[i - 1].key < h->a[c - 1].key ) break;
        t            = h->a[i - 1];
        h->a[i - 1]  = h->a[c - 1];
        h->a[c - 1]  = t;
        i            = c;
    }
    return top;
}

/* the live (not merged) supernodes below `node` in the son lists: a merged
 * son is replaced by its own sons */
static pastix_int_t
pa_live_sons( pastix_int_t node, const pastix_int_t *sonidx, const pastix_int_t *sons, const pastix_int_t *w,
              pastix_int_t *list )
{
    pastix_int_t i, m = 0;
    for ( i = sonidx[node]; i < sonidx[node + 1]; i++ ) {
        const pastix_int_t s = sons[i];
        if ( w[s] <= 0 ) m += pa_live_sons( s, sonidx, sons, w, list + m );
        else             list[m++] = s;
    }
    return m;
}

#define PA_NOMERGE 1e9

/*
 * Amalgamation of the partition of `o` (reference orderAmalgamate with
 * faxCSRAmalgamate): rat_cblk and rat_blas are the fill budgets in percent
 * of nnz(L) (IPARM_AMALGAMATION_LVLCBLK / LVLBLAS); a negative rat_blas
 * asks for the structural phase up to min(|rat_blas|, 2 %) and the BLAS
 * phase up to |rat_blas|.  The partition, tree and permutation of `o` are
 * replaced.
 */
int
pa_order_amalgamate( pastix_order_t *o, pastix_int_t n, const pastix_int_t *colptr, const pastix_int_t *rows,
                     int ilu, pastix_int_t lvl_cblk, pastix_int_t lvl_blas )
{
    const pastix_int_t cn   = o->cblknbr;
    pastix_int_t      *tree = o->treetab;
    double             rat_cblk = (double)lvl_cblk / 100., rat_blas = (double)lvl_blas / 100.;
    pa_rowset_t       *L;
    pastix_int_t      *w, *sonidx, *sons, *tmp, *tmp2, *fill, *newnum, *nrang, *ntree, *nperi;
    double            *gain;
    pa_heap_t          heap = { NULL, 0, 0 };
    pastix_int_t       i, j, k, f, nnzL, nmerged = 0, budget, budget_blas, total, ncn;
    int                blas_phase = 0, rc = PASTIX_ERR_ALLOC;

    if ( cn <= 1 ) return PASTIX_SUCCESS;
    if ( rat_blas < 0 ) {
        rat_blas = -rat_blas;
        rat_cblk = rat_blas < 0.02 ? rat_blas : 0.02;
    }
    L      = calloc( cn, sizeof( pa_rowset_t ) );
    w      = malloc( cn * sizeof( pastix_int_t ) );
    sonidx = calloc( cn + 1, sizeof( pastix_int_t ) );
    sons   = malloc( cn * sizeof( pastix_int_t ) );
    tmp    = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    tmp2   = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    fill   = malloc( cn * sizeof( pastix_int_t ) );
    gain   = malloc( cn * sizeof( double ) );
    if ( !L || !w || !sonidx || !sons || !tmp || !tmp2 || !fill || !gain ) goto end;

    nnzL = pa_supernodal_fill( o, n, colptr, rows, L, tmp );
    if ( nnzL < 0 ) goto end;
    for ( k = 0; k < cn; k++ ) w[k] = o->rangtab[k + 1] - o->rangtab[k];
    budget      = (pastix_int_t)lround( (double)nnzL * rat_cblk );
    budget_blas = (pastix_int_t)lround( (double)nnzL * rat_blas );

    /* son lists of the supernode tree (the last supernode is a root) */
    for ( k = 0; k < cn - 1; k++ )
        if ( tree[k] >= 0 ) sonidx[tree[k] + 1]++;
    for ( k = 0; k < cn; k++ ) sonidx[k + 1] += sonidx[k];
    memcpy( tmp, sonidx, cn * sizeof( pastix_int_t ) );
    for ( k = 0; k < cn - 1; k++ )
        if ( tree[k] >= 0 ) sons[tmp[tree[k]]++] = k;

    /* fill of merging each supernode into its father */
    for ( k = 0; k < cn; k++ ) {
        f = tree[k];
        if ( f == -1 || f == k ) {
            fill[k] = (pastix_int_t)PA_NOMERGE;
            gain[k] = PA_NOMERGE;
            continue;
        }
        fill[k] = pa_merge_fill( L, k, f, w );
        gain[k] = (double)fill[k];
    }

/* merge supernode i into its father; the sons of i become the father's */
#define PA_MERGE( i )                                                                          \
    do {                                                                                       \
        f = tree[i];                                                                           \
        if ( pa_merge_rows( L, i, f, tmp ) ) goto end;                                         \
        w[f] += w[i];                                                                          \
        w[i] = 0;                                                                              \
        nmerged++;                                                                             \
        k = tree[f];                                                                           \
        if ( k != -1 && k != f ) {                                                             \
            fill[f] = pa_merge_fill( L, f, k, w );                                             \
            if ( blas_phase ) {                                                                \
                gain[f] = pa_merge_gain( ilu, L, f, k, w, tmp2 ) / fill[f];                    \
                if ( queued && gain[f] <= 0 && pa_heap_push( &heap, f, gain[f] ) ) goto end;   \
            }                                                                                  \
            else {                                                                             \
                gain[f] = (double)fill[f];                                                     \
                if ( queued && pa_heap_push( &heap, f, gain[f] ) ) goto end;                   \
            }                                                                                  \
        }                                                                                      \
        total = pa_live_sons( i, sonidx, sons, w, tmp );                                       \
        for ( j = 0; j < total; j++ ) {                                                        \
            k       = tmp[j];                                                                  \
            tree[k] = f;                                                                       \
            fill[k] = pa_merge_fill( L, k, f, w );                                             \
            if ( blas_phase ) {                                                                \
                gain[k] = pa_merge_gain( ilu, L, k, f, w, tmp2 ) / fill[k];                    \
                if ( queued && gain[k] <= 0 && pa_heap_push( &heap, k, gain[k] ) ) goto end;   \
            }                                                                                  \
            else {                                                                             \
                gain[k] = (double)fill[k];                                                     \
                if ( queued && pa_heap_push( &heap, k, gain[k] ) ) goto end;                   \
            }                                                                                  \
        }                                                                                      \
    } while ( 0 )

    {
        int queued = 0;
        /* merges that add no fill, in supernode order */
        for ( i = 0; i < cn; i++ ) {
            if ( w[i] != 0 && fill[i] == 0 ) PA_MERGE( i );
        }
        for ( i = 0; i < cn; i++ )
            if ( w[i] > 0 && tree[i] > 0 && tree[i] != i )
                if ( pa_heap_push( &heap, i, gain[i] ) ) goto end;

        queued = 1;
        for ( ;; ) {
            /* cheapest merges first until the fill budget is spent */
            pastix_int_t spent = 0;
            while ( heap.used > 0 && spent < budget ) {
                double key;
                i = pa_heap_pop( &heap, &key );
                if ( gain[i] != key || w[i] <= 0 ) continue; /* stale entry */
                if ( spent + fill[i] > budget ) continue;
                spent += fill[i];
                PA_MERGE( i );
            }
            heap.used = 0;
            if ( blas_phase || budget >= budget_blas ) break;
            /* second round: merges that lower the modelled BLAS time, up to
             * the larger budget */
            blas_phase = 1;
            budget     = budget_blas;
            for ( i = 0; i < cn; i++ ) {
                f = tree[i];
                if ( f == -1 || f == i ) {
                    gain[i] = PA_NOMERGE;
                    continue;
                }
                gain[i] = pa_merge_gain( ilu, L, i, f, w, tmp ) / fill[i];
                if ( w[i] > 0 && tree[i] > 0 && tree[i] != i && gain[i] <= 0. )
                    if ( pa_heap_push( &heap, i, gain[i] ) ) goto end;
            }
        }
    }
#undef PA_MERGE

    /* the merged partition: live supernodes keep their order; a merged
     * supernode's columns go to the live ancestor it was merged into */
    ncn    = cn - nmerged;
    newnum = tmp;
    nrang  = calloc( ncn + 1, sizeof( pastix_int_t ) );
    ntree  = malloc( ( ncn + 1 ) * sizeof( pastix_int_t ) );
    nperi  = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    if ( !nrang || !ntree || !nperi ) { free( nrang ); free( ntree ); free( nperi ); goto end; }
    k = 0;
    for ( i = 0; i < cn; i++ ) {
        if ( w[i] > 0 ) {
            newnum[i] = k++;
            nrang[k]  = w[i];
        }
    }
    for ( i = 0; i < ncn; i++ ) nrang[i + 1] += nrang[i];
    /* nrang[c] is the next free slot of new supernode c while its columns
     * are placed, in the order of the old supernodes */
    for ( i = 0; i < cn; i++ ) {
        f = i;
        while ( w[f] <= 0 ) f = tree[f];
        {
            const pastix_int_t c = newnum[f];
            for ( j = o->rangtab[i]; j < o->rangtab[i + 1]; j++ ) nperi[nrang[c]++] = o->peritab[j];
            ntree[c] = ( tree[f] != -1 ) ? newnum[tree[f]] : -1;
        }
    }
    for ( i = ncn; i > 0; i-- ) nrang[i] = nrang[i - 1];
    nrang[0] = 0;
    for ( i = 0; i < n; i++ ) o->permtab[nperi[i]] = i;
    free( o->rangtab );
    free( o->treetab );
    free( o->peritab );
    o->rangtab = nrang;
    o->treetab = ntree;
    o->peritab = nperi;
    o->cblknbr = ncn;
    tree       = NULL;
    rc         = PASTIX_SUCCESS;

end:
    if ( L ) {
        for ( k = 0; k < cn; k++ ) free( L[k].rows );
        free( L );
    }
    free( w ); free( sonidx ); free( sons ); free( tmp ); free( tmp2 ); free( fill ); free( gain );
    free( heap.a );
    return rc;
}
