/*
 * pastix/order.h -- the ordering structure and its helpers.
 *
 * pastix_order_t keeps the reference's field list (reference
 * include/pastix/order.h:47-59); permtab/peritab/rangtab/treetab have the
 * reference's meaning: permtab[old] = new, peritab[new] = old, rangtab is the
 * supernode partition [cblknbr+1] in the new numbering and treetab[c] the
 * father supernode of c (-1 for a root).
 */
#ifndef PASTIX_AMD_ORDER_H
#define PASTIX_AMD_ORDER_H

#include "pastix/datatypes.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pastix_order_s {
    pastix_int_t  baseval;
    pastix_int_t  vertnbr;
    pastix_int_t  cblknbr;
    pastix_int_t *permtab;
    pastix_int_t *peritab;
    pastix_int_t *rangtab;
    pastix_int_t *treetab;
    int8_t       *selevtx;
    pastix_int_t  sndenbr;
    pastix_int_t *sndetab;
    pastix_int_t *peritab_exp;
} pastix_order_t;

int  pastixOrderInit( pastix_order_t *ordeptr, pastix_int_t baseval,
                      pastix_int_t vertnbr, pastix_int_t cblknbr,
                      pastix_int_t *perm, pastix_int_t *invp,
                      pastix_int_t *rang, pastix_int_t *tree );
int  pastixOrderAlloc( pastix_order_t *ordeptr, pastix_int_t vertnbr, pastix_int_t cblknbr );
void pastixOrderExit( pastix_order_t *ordeptr );
void pastixOrderBase( pastix_order_t *ordeptr, pastix_int_t baseval );
int  pastixOrderCheck( const pastix_order_t *ordeptr );
int  pastixOrderCopy( pastix_order_t *ordedst, const pastix_order_t *ordesrc );
pastix_order_t *pastixOrderGet( const pastix_data_t *pastix_data );

/* Closed-form nested dissection of an nx*ny*nz grid whose vertex (i,j,k) is
 * numbered i + nx*j + nx*ny*k (reference order/order_grids.c:307-384). */
int  pastixOrderGrid( pastix_order_t **myorder, pastix_int_t nx, pastix_int_t ny, pastix_int_t nz );

/* Ordering files in the reference's text format (reference
 * include/pastix/order.h:110-114, order/order_io.c): the file is named by the
 * PASTIX_FILE_ORDER environment variable, else "ordername" (load) /
 * "ordergen" (save) in the current directory.  pastix_subtask_order loads
 * with IPARM_IO_STRATEGY & PastixIOLoad and saves with & PastixIOSave. */
int  pastixOrderLoad( const pastix_data_t *pastix_data, pastix_order_t *ordeptr );
int  pastixOrderSave( pastix_data_t *pastix_data, const pastix_order_t *ordeptr );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_ORDER_H */
