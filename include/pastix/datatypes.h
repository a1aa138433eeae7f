/*
 * pastix/datatypes.h -- scalar and opaque types of the PaStiX API.
 *
 * pastix_int_t is 64-bit, the reference's default build option
 * (reference CMakeLists.txt:178 PASTIX_INT64=ON, include/pastix/datatypes.h:28-33).
 */
#ifndef PASTIX_AMD_DATATYPES_H
#define PASTIX_AMD_DATATYPES_H

#include <stdint.h>
#include <stddef.h>
#include "spm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t  pastix_int_t;
typedef uint64_t pastix_uint_t;
#define PASTIX_INT_MAX INT64_MAX

typedef double pastix_fixdbl_t;

/* No MPI in this build: a communicator is an opaque handle, as in the
 * reference's build without MPI (reference include/pastix.h:36-44). */
typedef uintptr_t PASTIX_Comm;
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD 0
#endif

struct pastix_data_s;
typedef struct pastix_data_s pastix_data_t;

struct pastix_rhs_s;
typedef struct pastix_rhs_s *pastix_rhs_t;

typedef spmatrix_t pastix_graph_t;

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_DATATYPES_H */
