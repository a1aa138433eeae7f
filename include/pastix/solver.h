/*
 * pastix/solver.h -- the supernodal structure (SolverMatrix) that drives the
 * numerical factorization and the solve.
 *
 * Field names and their order follow the reference's SolverBlok, SolverCblk
 * and SolverMatrix (reference blend/solver.h:141-156, :161-184, :203-279): the
 * GPU entry points receive host pointers to these structures exactly as the
 * reference's gpucblk_zgemmsp does (reference kernels/pastix_zcuda.h:32-35),
 * so the layout is part of the drop-in boundary.  Runtime-only members of the
 * reference (handlers, MPI requests, thread queues) are kept as opaque
 * pointers of the same size.
 */
#ifndef PASTIX_AMD_SOLVER_H
#define PASTIX_AMD_SOLVER_H

#include <stdio.h>
#include "pastix/datatypes.h"
#include "pastix/api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cblktype flags (reference blend/solver.h:51-57) */
#define CBLK_FANIN      (1 << 0)
#define CBLK_LAYOUT_2D  (1 << 1)
#define CBLK_TASKS_2D   (1 << 2)
#define CBLK_COMPRESSED (1 << 3)
#define CBLK_IN_SCHUR   (1 << 4)
#define CBLK_IN_LAST    (1 << 5)
#define CBLK_RECV       (1 << 6)

typedef volatile int32_t pastix_atomic_lock_t;

typedef struct solver_blok_s {
    void        *handler[2];
    pastix_int_t lcblknm;    /* column block owning the block                        */
    pastix_int_t fcblknm;    /* column block the block's rows face                   */
    pastix_int_t gbloknm;
    pastix_int_t gfaninnm;
    pastix_int_t frownum;    /* first row (global numbering, inclusive)              */
    pastix_int_t lrownum;    /* last row (inclusive)                                 */
    pastix_int_t coefind;    /* offset of the block tile inside its cblk's coeftab   */
    pastix_int_t browind;    /* position of the block in browtab                     */
    int8_t       inlast;
    int          iluklvl;
    void        *LRblock[2];
} SolverBlok;

typedef struct solver_cblk_s {
    pastix_atomic_lock_t lock;
    volatile int32_t     ctrbcnt;     /* number of incoming block updates        */
    int8_t               cblktype;    /* CBLK_* flags                            */
    int8_t               partitioned;
    pastix_int_t         fcolnum;     /* first column (inclusive)                */
    pastix_int_t         lcolnum;     /* last column (inclusive)                 */
    SolverBlok          *fblokptr;    /* diagonal block, first block of the cblk */
    pastix_int_t         stride;      /* total number of rows of the cblk        */
    pastix_int_t         lcolidx;     /* first column in the rhs vectors         */
    pastix_int_t         brownum;     /* first entry in browtab facing this cblk */
    pastix_int_t         brown2d;
    pastix_int_t         sndeidx;
    pastix_int_t         gcblknum;
    pastix_int_t         bcscnum;
    pastix_int_t         gfaninnum;
    void                *lcoeftab;    /* L panel (device memory on the GPU path) */
    void                *ucoeftab;    /* U^T panel for LU                        */
    void                *handler[2];
    pastix_int_t         selevtx;
    int                  ownerid;
    int                  threadid;
    pastix_int_t         priority;
} SolverCblk;

/* A task of the static scheduler: one cblk factorization with its updates
 * (reference blend/solver.h:121-130, task type COMP_1D). */
typedef struct task_s {
    pastix_int_t          taskid;   /* COMP_1D (0)                                */
    pastix_int_t          prionum;  /* rank of the task in the simulated execution */
    pastix_int_t          cblknum;  /* attached column block                       */
    pastix_int_t          bloknum;  /* its diagonal block                          */
    pastix_int_t volatile ctrbcnt;  /* number of incoming contributions            */
} Task;

typedef struct solver_matrix_s {
    int          restore;
    pastix_int_t baseval;
    pastix_int_t nodenbr;      /* number of columns                               */
    pastix_int_t coefnbr;      /* number of stored coefficients per factor        */
    pastix_int_t gcblknbr;
    pastix_int_t cblknbr;
    pastix_int_t gfanincblknbr;
    pastix_int_t faninnbr;
    pastix_int_t fanincnt;
    pastix_int_t maxrecv;
    pastix_int_t recvnbr;
    pastix_int_t recvcnt;
    pastix_int_t cblkmax1d;
    pastix_int_t cblkmin2d;
    pastix_int_t cblkmaxblk;   /* maximum number of blocks of a cblk               */
    pastix_int_t cblkschur;
    pastix_int_t nb2dcblk;
    pastix_int_t nb2dblok;
    pastix_int_t bloknbr;
    pastix_int_t gbloknbr;
    pastix_int_t gfaninbloknbr;
    pastix_int_t brownbr;
    SolverCblk  *cblktab;      /* cblknbr+1 entries, the last one is a sentinel    */
    SolverBlok  *bloktab;      /* bloknbr+1 entries                                */
    pastix_int_t *browtab;     /* blocks facing each cblk, grouped by target       */
    pastix_coeftype_t flttype;
    int          globalalloc;
    pastix_int_t *gcbl2loc;
    void        *lowrank_unused[8];
    pastix_factotype_t factotype;
    double       diagthreshold;
    volatile int32_t nbpivots;
    pastix_int_t offdmax;
    pastix_int_t gemmmax;
    pastix_int_t blokmax;
    pastix_int_t colmax;
    /* static scheduling of the CPU factorization (reference blend/solver.h:
     * 251-265): the proportional mapping and the simulation of the analysis
     * give each of the thrdnbr threads the ordered list of its tasks */
    int           clustnum;
    int           clustnbr;
    pastix_int_t  procnbr;
    pastix_int_t  thrdnbr;
    pastix_int_t  bublnbr;
    Task         *tasktab;     /* tasknbr tasks                                 */
    pastix_int_t  tasknbr;
    pastix_int_t  tasknbr_1dp;
    pastix_int_t **ttsktab;    /* thrdnbr lists of task indices                 */
    pastix_int_t *ttsknbr;     /* their lengths                                 */
} SolverMatrix;

/* Structure I/O in the reference's text format (reference
 * blend/solver_io.c:42-355).  solverSave writes version 2; solverLoad reads
 * what solverSave writes and rebuilds the derived fields (coefficient
 * counts, browtab, the cblks of each block). */
int solverLoad( SolverMatrix *solvptr, FILE *stream );
int solverSave( const SolverMatrix *solvptr, FILE *stream );
void solverExit( SolverMatrix *solvptr );

static inline pastix_int_t cblk_colnbr( const SolverCblk *cblk ) {
    return cblk->lcolnum - cblk->fcolnum + 1;
}
static inline pastix_int_t blok_rownbr( const SolverBlok *blok ) {
    return blok->lrownum - blok->frownum + 1;
}
static inline pastix_int_t cblk_bloknbr( const SolverCblk *cblk ) {
    return (cblk + 1)->fblokptr - cblk->fblokptr + 1;
}
/* Row range of `blok` contained in the row range of `fblok`
 * (reference blend/solver.h:504). */
static inline int is_block_inside_fblock( const SolverBlok *blok, const SolverBlok *fblok ) {
    return (blok->frownum >= fblok->frownum) && (blok->lrownum <= fblok->lrownum);
}

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_SOLVER_H */
