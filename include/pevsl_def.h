/*
 * pevsl_def.h -- constants and macros of the pEVSL C API
 * (same names and values as INC/pevsl_def.h:10-112).
 */
#ifndef PEVSL_DEF_H
#define PEVSL_DEF_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "pevsl_mpi.h"

#define PEVSL_NOT_DEFINED -2   /* INC/pevsl_def.h:10 */
#define PEVSL_NOT_IMPLEMENT 11 /* INC/pevsl_def.h:11 */
#define PEVSL_MAX_LINE 2048

#define PEVSL_MAX(a, b) ((a) > (b) ? (a) : (b))
#define PEVSL_MIN(a, b) ((a) < (b) ? (a) : (b))

/* relative threshold of the lucky-breakdown test (INC/pevsl_def.h:18) */
#define orthTol 1e-14

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef PI
#define PI M_PI
#endif

/* fatal error: print and abort the process (INC/pevsl_def.h:22-27) */
#define PEVSL_ABORT(comm, errcode, msg)                                   \
  {                                                                       \
    (void)(comm);                                                         \
    printf("PEVSL error (processor 0): %s \n", msg);                      \
    fflush(stdout);                                                       \
    exit(errcode);                                                        \
  }

#ifdef PEVSL_DEBUG
#define PEVSL_CHKERR(ierr) \
  if (ierr) { PEVSL_ABORT(MPI_COMM_WORLD, 999, "Assertion failed") }
#else
#define PEVSL_CHKERR(ierr)
#endif

#define PEVSL_MALLOC(base, nmem, type)                                        \
  {                                                                           \
    size_t nbytes__ = (size_t)(nmem) * sizeof(type);                          \
    (base) = (type *)malloc(nbytes__);                                        \
    if ((base) == NULL && nbytes__ > 0) {                                     \
      fprintf(stdout, "EVSL Error: out of memory [%zu bytes asked]\n",        \
              nbytes__);                                                      \
      exit(-1);                                                               \
    }                                                                         \
  }

#define PEVSL_CALLOC(base, nmem, type)                                        \
  {                                                                           \
    (base) = (type *)calloc((size_t)(nmem), sizeof(type));                    \
    if ((base) == NULL && (nmem) > 0) {                                       \
      fprintf(stdout, "EVSL Error: out of memory [%zu bytes asked]\n",        \
              (size_t)(nmem) * sizeof(type));                                 \
      exit(-1);                                                               \
    }                                                                         \
  }

#define PEVSL_REALLOC(base, nmem, type)                                       \
  {                                                                           \
    size_t nbytes__ = (size_t)(nmem) * sizeof(type);                          \
    (base) = (type *)realloc((base), nbytes__);                               \
    if ((base) == NULL && nbytes__ > 0) {                                     \
      fprintf(stdout, "EVSL Error: out of memory [%zu bytes asked]\n",        \
              nbytes__);                                                      \
      exit(-1);                                                               \
    }                                                                         \
  }

#define PEVSL_FREE(base) \
  {                      \
    if (base) {          \
      free(base);        \
    }                    \
    base = NULL;         \
  }

/* ordered execution over a communicator (INC/pevsl_def.h:79-92); one rank */
#define PEVSL_SEQ_BEGIN(MPI_COMM) { int size = 1, rank = 0; (void)(MPI_COMM); (void)size; {
#define PEVSL_SEQ_END(MPI_COMM) } }

#define PEVSL_FORT(name) name##_f90_

/* max number of Gram-Schmidt passes in reorthogonalization (:112) */
#define NGS_MAX 2

#endif
