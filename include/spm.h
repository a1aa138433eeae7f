/*
 * spm.h -- the sparse-matrix input type consumed by the PaStiX API.
 *
 * The reference takes its input through the SpM library (the `spm` git
 * submodule, SpM >= 1.2.4, reference CMakeLists.txt:502), which is not part of
 * the reference tree.  This header restates the subset of SpM that the
 * factorization path reads: the spmatrix_t descriptor of a CSC/CSR/IJV matrix
 * and its enumerations.  Enumerator values follow SpM's published constants
 * (coefficient types 0/2/3/4/5, matrix types 111/112/113, formats 0/1/2) so that
 * `flttype - 2` indexes precision tables the way the reference does
 * (reference sopalin/pastix_task_sopalin.c:509).
 */
#ifndef PASTIX_AMD_SPM_H
#define PASTIX_AMD_SPM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t spm_int_t;

typedef enum spm_coeftype_e {
    SpmPattern   = 0,
    SpmFloat     = 2,
    SpmDouble    = 3,
    SpmComplex32 = 4,
    SpmComplex64 = 5
} spm_coeftype_t;

typedef enum spm_mtxtype_e {
    SpmGeneral   = 111,
    SpmSymmetric = 112,
    SpmHermitian = 113
} spm_mtxtype_t;

typedef enum spm_fmttype_e {
    SpmCSC = 0,
    SpmCSR = 1,
    SpmIJV = 2
} spm_fmttype_t;

typedef enum spm_layout_e {
    SpmRowMajor = 101,
    SpmColMajor = 102
} spm_layout_t;

typedef int SPM_Comm;

/*
 * Sparse matrix descriptor.  Only single-process, dof == 1 matrices are
 * accepted by this build; the distributed fields are kept so that a caller's
 * initialiser lines up.
 */
typedef struct spmatrix_s {
    spm_mtxtype_t  mtxtype;  /* SpmGeneral, SpmSymmetric or SpmHermitian            */
    spm_coeftype_t flttype;  /* arithmetic of `values`                               */
    spm_fmttype_t  fmttype;  /* SpmCSC (colptr/rowptr), SpmCSR or SpmIJV             */
    spm_int_t      baseval;  /* 0 or 1 based indices                                */
    spm_int_t      gN;       /* global number of unknowns                            */
    spm_int_t      n;        /* local number of unknowns                             */
    spm_int_t      gnnz;     /* global number of stored entries                      */
    spm_int_t      nnz;      /* local number of stored entries                       */
    spm_int_t      gNexp;    /* gN after dof expansion (== gN here)                  */
    spm_int_t      nexp;     /* n after dof expansion (== n here)                    */
    spm_int_t      gnnzexp;  /* gnnz after dof expansion                             */
    spm_int_t      nnzexp;   /* nnz after dof expansion                              */
    spm_int_t      dof;      /* degrees of freedom per unknown (must be 1)           */
    spm_int_t     *dofs;     /* variable dofs (must be NULL)                         */
    spm_layout_t   layout;   /* layout of dof blocks (unused when dof == 1)          */
    spm_int_t     *colptr;   /* CSC: column pointers [n+1]; IJV: column indices     */
    spm_int_t     *rowptr;   /* CSC: row indices [nnz]                               */
    spm_int_t     *loc2glob; /* distributed numbering (must be NULL)                 */
    void          *values;   /* nnz coefficients of type flttype                     */
    spm_int_t     *glob2loc;
    int            clustnum;
    int            clustnbr;
    SPM_Comm       comm;
    int            replicated;
} spmatrix_t;

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_SPM_H */
