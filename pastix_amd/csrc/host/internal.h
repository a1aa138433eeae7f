/*
 * internal.h -- host-side data structures shared by the analysis, the API and
 * the device driver of the MI355X PaStiX engine.
 */
#ifndef PASTIX_AMD_INTERNAL_H
#define PASTIX_AMD_INTERNAL_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pastix.h"
#include "pastix/solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Steps bitmask (reference common/pastixdata.h:28-36). */
#define STEP_INIT      (1 << 0)
#define STEP_ORDERING  (1 << 1)
#define STEP_SYMBFACT  (1 << 2)
#define STEP_ANALYSE   (1 << 3)
#define STEP_CSC2BCSC  (1 << 4)
#define STEP_BCSC2CTAB (1 << 5)
#define STEP_NUMFACT   (1 << 6)
#define STEP_SOLVE     (1 << 7)
#define STEP_REFINE    (1 << 8)

#define pa_imin(a, b) ((a) < (b) ? (a) : (b))
#define pa_imax(a, b) ((a) > (b) ? (a) : (b))
#define pa_iceil(a, b) (((a) + (b) - 1) / (b))

void pa_error( const char *fmt, ... );
void pa_warning( const char *fmt, ... );
double pa_wtime( void );

/* -------------------------------------------------------------------------- */
/* Symbol matrix: the block structure of L (reference symbol/symbol.h).       */
/* -------------------------------------------------------------------------- */
typedef struct symbol_cblk_s {
    pastix_int_t fcolnum;
    pastix_int_t lcolnum;
    pastix_int_t bloknum;   /* first (diagonal) block                        */
    pastix_int_t brownum;   /* first entry of browtab for this cblk          */
    int8_t       selevtx;
} symbol_cblk_t;

typedef struct symbol_blok_s {
    pastix_int_t frownum;
    pastix_int_t lrownum;
    pastix_int_t lcblknm;
    pastix_int_t fcblknm;
} symbol_blok_t;

typedef struct symbol_matrix_s {
    pastix_int_t   baseval;
    pastix_int_t   cblknbr;
    pastix_int_t   bloknbr;
    pastix_int_t   nodenbr;
    pastix_int_t   schurfcol;
    symbol_cblk_t *cblktab;   /* cblknbr+1 (sentinel)                         */
    symbol_blok_t *bloktab;   /* bloknbr                                      */
    pastix_int_t  *browtab;   /* blocks grouped by facing cblk                */
    pastix_int_t   browmax;
} symbol_matrix_t;

void pa_symbol_exit( symbol_matrix_t *symb );

/* Grid / personal orderings (order.c) */
int  pa_order_apply_level_order( pastix_order_t *order, pastix_int_t level_tasks2d,
                                 pastix_int_t width_tasks2d );
/* Supernode partition of a bare permutation (amalgamate.c): fundamental
 * supernodes of the postordered elimination tree (perm/invp postordered in
 * place), then amalgamation under the fill budgets lvl_cblk / lvl_blas (in
 * percent of nnz(L)); colptr/rows: the symmetric graph without diagonal. */
int  pa_order_find_supernodes( pastix_order_t *order, pastix_int_t n, const pastix_int_t *colptr,
                               const pastix_int_t *rows );
int  pa_order_amalgamate( pastix_order_t *order, pastix_int_t n, const pastix_int_t *colptr,
                          const pastix_int_t *rows, int ilu, pastix_int_t lvl_cblk, pastix_int_t lvl_blas );

/* Symmetrised adjacency graph without diagonal, 0-based CSC. */
typedef struct pa_graph_s {
    pastix_int_t  n;
    pastix_int_t *colptr;   /* n+1 */
    pastix_int_t *rows;     /* colptr[n] */
} pa_graph_t;

int  pa_graph_from_spm( pa_graph_t *g, const spmatrix_t *spm );
void pa_graph_exit( pa_graph_t *g );

/* Supernodal symbolic factorization (symbol.c) */
int  pa_symbol_fax( symbol_matrix_t *symb, const pa_graph_t *g, const pastix_order_t *order );
void pa_symbol_build_browtab( symbol_matrix_t *symb );
int  pa_symbol_check( const symbol_matrix_t *symb );
double pa_symbol_flops( const symbol_matrix_t *symb, int factotype );
pastix_int_t pa_symbol_nnz( const symbol_matrix_t *symb );

/* Supernode splitting (split.c) */
int pa_split_symbol( symbol_matrix_t *symb, pastix_int_t blcolmin, pastix_int_t blcolmax );

/* SolverMatrix generation (solver.c) */
int  pa_solver_gen( SolverMatrix *solv, const symbol_matrix_t *symb, int factotype, int flttype );
void pa_solver_exit( SolverMatrix *solv );
int  pa_solver_build_browtab( SolverMatrix *solv );
/* Static schedule of the CPU factorization over thrdnbr threads (simu.c):
 * tasktab, ttsktab / ttsknbr and each cblk's thread; *pred_time gets the
 * simulated factorization time (seconds of the fitted kernel model) */
int  pa_blend_static( SolverMatrix *solv, int fact, pastix_int_t thrdnbr, int allcand, double *pred_time );
/* Level (height in the elimination tree of cblks) of each cblk: leaves 0. */
pastix_int_t pa_solver_levels( const SolverMatrix *solv, pastix_int_t *level );

/* -------------------------------------------------------------------------- */
/* Block CSC: the user matrix permuted to the new numbering (reference        */
/* bcsc/bcsc.h).  Values stay FP64 on the host and are mirrored on the GPU.   */
/* -------------------------------------------------------------------------- */
typedef struct pa_bcsc_s {
    pastix_int_t  n;
    pastix_int_t  nnz;          /* entries of the fully expanded matrix        */
    int           mtxtype;
    pastix_int_t *colptr;       /* n+1, permuted columns                       */
    pastix_int_t *rowtab;       /* nnz, permuted rows sorted per column        */
    double       *Lvalues;      /* A(row, col)                                 */
    double       *Uvalues;      /* A(col, row): the transpose, same pattern    */
    double        normF;        /* ||A||_F                                     */
    double        normInf;      /* ||A||_inf                                   */
} pa_bcsc_t;

/* Where the values of the expanded pattern come from: position w of
 * Lvalues is the sum of the user's stored values vals[Lsrc[Lptr[w] ..
 * Lptr[w+1]-1]], likewise for Uvalues (duplicates are summed, a symmetric
 * matrix storing both triangles keeps one copy). */
typedef struct pa_valmap_s {
    pastix_int_t  nnz;          /* positions of the expanded pattern           */
    pastix_int_t  nuser;        /* stored entries of the user matrix           */
    pastix_int_t *Lptr, *Lsrc;
    pastix_int_t *Uptr, *Usrc;
} pa_valmap_t;

int  pa_bcsc_init( pa_bcsc_t *bcsc, const spmatrix_t *spm, const pastix_order_t *order );
/* pa_bcsc_init that also records the value map (map may be NULL) */
int  pa_bcsc_build( pa_bcsc_t *bcsc, const spmatrix_t *spm, const pastix_order_t *order, pa_valmap_t *map );
void pa_bcsc_exit( pa_bcsc_t *bcsc );
void pa_valmap_exit( pa_valmap_t *map );

/* -------------------------------------------------------------------------- */
/* Main instance structure.                                                   */
/* -------------------------------------------------------------------------- */
/* Right-hand sides (reference common/pastixdata.h:155-168): allocated is -1
 * until the vector is set, 0 when b is the caller's array (the permutation
 * is applied in place), 1 when the library owns it. */
struct pastix_rhs_s {
    int           allocated;
    int           flttype;
    pastix_int_t  m;
    pastix_int_t  n;
    pastix_int_t  ld;
    void         *b;
};

struct pa_device_s;   /* GPU state, defined by the device layer */

struct pastix_data_s {
    pastix_int_t       id;
    pastix_int_t      *iparm;
    double            *dparm;
    pastix_int_t       steps;
    PASTIX_Comm        pastix_comm;
    const spmatrix_t  *csc;
    pa_graph_t        *graph;
    pastix_order_t    *ordemesh;
    symbol_matrix_t   *symbmtx;
    pa_bcsc_t         *bcsc;
    SolverMatrix      *solvmatr;
    struct pa_device_s *dev;
    int                valmap_ready;   /* the device holds the value map of csc */
    /* op(A) of the running refinement: IPARM_TRANSPOSE_SOLVE for the API
     * entry points (reference bcsc/bcsc_zspmv.c:638), or the transposition
     * given to pastix_amd_refine_dev_adv (the adjoint solve) */
    pastix_trans_t     refine_trans;
    /* legacy pastix() interface: the spm built from the caller's arrays and
     * the backups of b / x0 kept between its solve and refinement steps
     * (reference sopalin/pastix.c:193-250, 363-420) */
    spmatrix_t        *legacy_spm;
    void              *b;
    void              *x0;
};

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_INTERNAL_H */
