/*
 * factor_io.c -- factor I/O: the reference's human-readable coefficient dump
 * (coeftab_zdump, reference sopalin/coeftab_z.c:55-90 with cpucblk_zdump,
 * sopalin/coeftab_zinit.c:54-160) and binary checkpoints of the factors held
 * in HBM, so that a factorization can be saved and re-used by another
 * process (no reference counterpart: the reference keeps its factors in
 * memory only).
 */
#include <stdint.h>
#include "internal.h"
#include "pa_device.h"

/* ---- human-readable dump ------------------------------------------------ */

/* one cblk of one side, lines "row col value [L]" (L) or "col row value [U]"
 * (U, from the stored U^T), nonzero coefficients only and the lower
 * triangle of the diagonal block (reference sopalin/coeftab_zinit.c:54-160) */
static void
pa_cblk_dump( int side, const SolverCblk *cblk, const double *coef, FILE *f )
{
    const SolverBlok *bl, *lb = cblk[1].fblokptr;
    pastix_int_t      col, row;
    for ( col = cblk->fcolnum; col <= cblk->lcolnum; col++ ) {
        for ( bl = cblk->fblokptr; bl < lb; bl++ ) {
            const pastix_int_t m  = blok_rownbr( bl );
            const double      *cv = coef + bl->coefind + ( col - cblk->fcolnum ) * m;
            for ( row = bl->frownum; row <= bl->lrownum; row++, cv++ ) {
                if ( !( *cv != 0. ) || ( bl == cblk->fblokptr && col > row ) ) continue;
                if ( side )
                    fprintf( f, "%ld %ld %13e [U]\n", (long)col, (long)row, *cv );
                else
                    fprintf( f, "%ld %ld %13e [L]\n", (long)row, (long)col, *cv );
            }
        }
    }
}

/* Every cblk of the factors in HBM into <prefix>_<cblk>.txt, L then (for
 * two-array factorizations) the second array */
static void
pa_coeftab_dump( pastix_data_t *pd, const SolverMatrix *solv, const char *prefix, int want_prec, const char *name )
{
    int          prec, two, fact, side;
    int64_t      coefnbr;
    double      *coef[2] = { NULL, NULL };
    pastix_int_t k, base;
    char         fname[4096];
    if ( pd == NULL || solv == NULL || prefix == NULL || pa_dev_factor_info( pd, &prec, &two, &fact, &coefnbr ) ) {
        pa_error( "%s: no factors to dump", name );
        return;
    }
    if ( prec != want_prec ) {
        pa_error( "%s: the factors are in %s precision", name, prec == PA_PREC_DOUBLE ? "double" : "single" );
        return;
    }
    for ( side = 0; side <= two; side++ ) {
        coef[side] = malloc( (size_t)pa_imax( coefnbr, 1 ) * sizeof( double ) );
        if ( !coef[side] || pa_dev_get_coeftab( pd, side, coef[side] ) ) {
            pa_error( "%s: cannot read the factors", name );
            free( coef[0] );
            free( coef[1] );
            return;
        }
    }
    for ( k = 0, base = 0; k < solv->cblknbr; base += solv->cblktab[k].stride * cblk_colnbr( solv->cblktab + k ), k++ ) {
        const SolverCblk *c = solv->cblktab + k;
        FILE             *f;
        if ( c->cblktype & ( CBLK_FANIN | CBLK_RECV ) ) continue;
        snprintf( fname, sizeof( fname ), "%s_%ld.txt", prefix, (long)c->gcblknum );
        f = fopen( fname, "w" );
        if ( !f ) {
            pa_error( "%s: cannot open %s", name, fname );
            continue;
        }
        /* the cblk's panel starts after those of the cblks before it */
        pa_cblk_dump( 0, c, coef[0] + base, f );
        if ( two ) pa_cblk_dump( 1, c, coef[1] + base, f );
        fclose( f );
    }
    free( coef[0] );
    free( coef[1] );
}

void
coeftab_ddump( pastix_data_t *pastix_data, const SolverMatrix *solvmtx, const char *prefix )
{
    pa_coeftab_dump( pastix_data, solvmtx, prefix, PA_PREC_DOUBLE, "coeftab_ddump" );
}

void
coeftab_sdump( pastix_data_t *pastix_data, const SolverMatrix *solvmtx, const char *prefix )
{
    pa_coeftab_dump( pastix_data, solvmtx, prefix, PA_PREC_FLOAT, "coeftab_sdump" );
}

/* ---- binary checkpoints --------------------------------------------------- */

#define PA_CKPT_MAGIC   "PAFACT01"
#define PA_CKPT_VERSION 1
#define PA_CKPT_CHUNK   ( (int64_t)1 << 23 )   /* coefficients per transfer */

typedef struct pa_ckpt_head_s {
    char     magic[8];
    int64_t  version;
    int64_t  factotype;
    int64_t  prec;        /* 2: FP32, 3: FP64                    */
    int64_t  two;         /* second factor array (LU, LDL^t)     */
    int64_t  nodenbr, cblknbr, bloknbr, coefnbr;
    uint64_t structure;   /* pa_solver_fingerprint                */
    int64_t  nbpivots;
    double   diagthreshold;
    double   normA;
} pa_ckpt_head_t;

/* FNV-1a hash of the block structure: the factors of one structure are
 * meaningless on another */
static uint64_t
pa_solver_fingerprint( const SolverMatrix *s )
{
    uint64_t     h = 1469598103934665603ULL;
    pastix_int_t k, b;
#define PA_MIX( v ) do { uint64_t _v = (uint64_t)( v ); int _i; \
        for ( _i = 0; _i < 8; _i++ ) { h ^= ( _v >> ( 8 * _i ) ) & 0xff; h *= 1099511628211ULL; } } while ( 0 )
    PA_MIX( s->nodenbr ); PA_MIX( s->cblknbr ); PA_MIX( s->bloknbr ); PA_MIX( s->coefnbr );
    for ( k = 0; k < s->cblknbr; k++ ) {
        const SolverCblk *c = s->cblktab + k;
        PA_MIX( c->fcolnum ); PA_MIX( c->lcolnum ); PA_MIX( c->stride ); PA_MIX( c->fblokptr - s->bloktab );
    }
    for ( b = 0; b < s->bloknbr; b++ ) {
        const SolverBlok *bl = s->bloktab + b;
        PA_MIX( bl->frownum ); PA_MIX( bl->lrownum ); PA_MIX( bl->fcblknm ); PA_MIX( bl->coefind );
    }
#undef PA_MIX
    return h;
}

int
pastix_amd_factors_save( pastix_data_t *pd, const char *path )
{
    pa_ckpt_head_t h;
    int            prec, two, fact, side, rc = PASTIX_SUCCESS;
    int64_t        coefnbr, off, es;
    void          *buf;
    FILE          *f;
    if ( pd == NULL || path == NULL || !( pd->steps & STEP_NUMFACT ) ||
         pa_dev_factor_info( pd, &prec, &two, &fact, &coefnbr ) ) {
        pa_error( "pastix_amd_factors_save: pastix_task_numfact() has to be called first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    memset( &h, 0, sizeof( h ) );
    memcpy( h.magic, PA_CKPT_MAGIC, 8 );
    h.version       = PA_CKPT_VERSION;
    h.factotype     = fact;
    h.prec          = prec;
    h.two           = two;
    h.nodenbr       = pd->solvmatr->nodenbr;
    h.cblknbr       = pd->solvmatr->cblknbr;
    h.bloknbr       = pd->solvmatr->bloknbr;
    h.coefnbr       = coefnbr;
    h.structure     = pa_solver_fingerprint( pd->solvmatr );
    h.nbpivots      = pd->solvmatr->nbpivots;
    h.diagthreshold = pd->solvmatr->diagthreshold;
    h.normA         = pd->dparm[DPARM_A_NORM];
    es  = prec == PA_PREC_DOUBLE ? 8 : 4;
    buf = malloc( (size_t)( PA_CKPT_CHUNK * es ) );
    f   = fopen( path, "wb" );
    if ( !buf || !f ) {
        pa_error( "pastix_amd_factors_save: cannot open %s", path );
        free( buf );
        if ( f ) fclose( f );
        return PASTIX_ERR_FILE;
    }
    if ( fwrite( &h, sizeof( h ), 1, f ) != 1 ) rc = PASTIX_ERR_FILE;
    for ( side = 0; side <= two && !rc; side++ ) {
        for ( off = 0; off < coefnbr && !rc; off += PA_CKPT_CHUNK ) {
            int64_t cnt = pa_imin( PA_CKPT_CHUNK, coefnbr - off );
            if ( ( rc = pa_dev_factor_get( pd, side, off, cnt, buf ) ) ) break;
            if ( fwrite( buf, (size_t)es, (size_t)cnt, f ) != (size_t)cnt ) rc = PASTIX_ERR_FILE;
        }
    }
    free( buf );
    if ( fclose( f ) != 0 && !rc ) rc = PASTIX_ERR_FILE;
    if ( rc ) pa_error( "pastix_amd_factors_save: writing %s failed", path );
    return rc;
}

int
pastix_amd_factors_load( pastix_data_t *pd, const char *path )
{
    pa_ckpt_head_t h;
    int            prec, side, rc = PASTIX_SUCCESS;
    int64_t        off, es;
    void          *buf;
    FILE          *f;
    if ( pd == NULL || path == NULL || !( pd->steps & STEP_ANALYSE ) || !( pd->steps & STEP_CSC2BCSC ) ) {
        pa_error( "pastix_amd_factors_load: pastix_task_analyze() and pastix_subtask_spm2bcsc() have to be called "
                  "first" );
        return PASTIX_ERR_BADPARAMETER;
    }
    f = fopen( path, "rb" );
    if ( !f ) {
        pa_error( "pastix_amd_factors_load: cannot open %s", path );
        return PASTIX_ERR_FILE;
    }
    if ( fread( &h, sizeof( h ), 1, f ) != 1 || memcmp( h.magic, PA_CKPT_MAGIC, 8 ) != 0 ||
         h.version != PA_CKPT_VERSION ) {
        fclose( f );
        pa_error( "pastix_amd_factors_load: %s is not a factor checkpoint", path );
        return PASTIX_ERR_FILE;
    }
    prec = pd->iparm[IPARM_MIXED] ? PA_PREC_FLOAT : PA_PREC_DOUBLE;
    if ( pd->iparm[IPARM_FLOAT] == SpmFloat ) prec = PA_PREC_FLOAT;
    if ( h.structure != pa_solver_fingerprint( pd->solvmatr ) || h.coefnbr != pd->solvmatr->coefnbr ||
         h.nodenbr != pd->solvmatr->nodenbr || h.factotype != pd->iparm[IPARM_FACTORIZATION] || h.prec != prec ||
         h.two != ( h.factotype == PastixFactLU || h.factotype == PastixFactLDLT || h.factotype == PastixFactLDLH ) ) {
        fclose( f );
        pa_error( "pastix_amd_factors_load: %s holds factors of another structure, factorization kind or precision",
                  path );
        return PASTIX_ERR_BADPARAMETER;
    }
    pd->solvmatr->flttype = ( prec == PA_PREC_FLOAT ) ? PastixFloat : PastixDouble;
    if ( pd->solvmatr->factotype != h.factotype ) {
        pd->solvmatr->factotype       = (int)h.factotype;
        pd->dparm[DPARM_FACT_THFLOPS] = pa_symbol_flops( pd->symbmtx, (int)h.factotype );
    }
    pd->steps &= ~( STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE );
    if ( ( rc = pa_dev_factor_begin( pd, prec, (int)h.two ) ) ) {
        fclose( f );
        return rc;
    }
    es  = prec == PA_PREC_DOUBLE ? 8 : 4;
    buf = malloc( (size_t)( PA_CKPT_CHUNK * es ) );
    if ( !buf ) rc = PASTIX_ERR_OUTOFMEMORY;
    for ( side = 0; side <= h.two && !rc; side++ ) {
        for ( off = 0; off < h.coefnbr && !rc; off += PA_CKPT_CHUNK ) {
            int64_t cnt = pa_imin( PA_CKPT_CHUNK, h.coefnbr - off );
            if ( fread( buf, (size_t)es, (size_t)cnt, f ) != (size_t)cnt ) {
                pa_error( "pastix_amd_factors_load: %s is truncated", path );
                rc = PASTIX_ERR_FILE;
                break;
            }
            rc = pa_dev_factor_put( pd, side, off, cnt, buf );
        }
    }
    free( buf );
    fclose( f );
    if ( !rc ) rc = pa_dev_factor_end( pd );
    if ( rc ) return rc;
    pd->solvmatr->nbpivots           = (int32_t)h.nbpivots;
    pd->solvmatr->diagthreshold      = h.diagthreshold;
    pd->iparm[IPARM_STATIC_PIVOTING] = h.nbpivots;
    pd->steps |= STEP_BCSC2CTAB | STEP_NUMFACT;
    return PASTIX_SUCCESS;
}
