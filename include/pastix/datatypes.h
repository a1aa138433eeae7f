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

/* No MPI in this build: a communicator is a plain integer handle. */
typedef int PASTIX_Comm;

struct pastix_data_s;
typedef struct pastix_data_s pastix_data_t;

struct pastix_rhs_s;
typedef struct pastix_rhs_s *pastix_rhs_t;

typedef spmatrix_t pastix_graph_t;

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_DATATYPES_H */
