/*
 * pevsl_mpi.h -- single-rank stand-in for the MPI names that appear in the
 * pEVSL public signatures (INC/pevsl.h:18,40,55,67 carry MPI_Comm and
 * INC/pevsl_def.h:8 includes <mpi.h>) and in drivers written against them
 * (MPI_Init/Finalize, communicator splits, reductions of one rank).
 *
 * This library runs one process on one GPU: the reference's row partition
 * (diag/offd + halo) collapses into one device-resident CSR and every
 * communicator has exactly one rank, so every collective below is a copy.
 * Callers that compile against a real MPI define PEVSL_USE_SYSTEM_MPI before
 * including pevsl.h and run with one rank.
 */
#ifndef PEVSL_MPI_H
#define PEVSL_MPI_H

#ifdef PEVSL_USE_SYSTEM_MPI
#include <mpi.h>
#else

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef int MPI_Comm;
typedef int MPI_Fint;  /* Fortran handle of a communicator */
typedef int MPI_Datatype;
typedef int MPI_Op;
typedef int MPI_Request;
typedef struct { int MPI_SOURCE, MPI_TAG, MPI_ERROR; } MPI_Status;

#define MPI_COMM_WORLD ((MPI_Comm)0)
#define MPI_COMM_SELF ((MPI_Comm)1)
#define MPI_COMM_NULL ((MPI_Comm)-1)
#define MPI_SUCCESS 0
#define MPI_UNDEFINED (-32766)
#define MPI_INT ((MPI_Datatype)4)
#define MPI_DOUBLE ((MPI_Datatype)8)
#define MPI_UNSIGNED_LONG ((MPI_Datatype)9)
#define MPI_SUM ((MPI_Op)1)
#define MPI_MAX ((MPI_Op)2)
#define MPI_MIN ((MPI_Op)3)

static inline size_t pevsl_mpi_size_of(MPI_Datatype t) {
  return t == MPI_INT ? sizeof(int) : (t == MPI_UNSIGNED_LONG ? sizeof(unsigned long) : sizeof(double));
}

static inline MPI_Comm MPI_Comm_f2c(MPI_Fint f) { return (MPI_Comm)f; }
static inline int MPI_Init(int *argc, char ***argv) { (void)argc; (void)argv; return 0; }
static inline int MPI_Finalize(void) { return 0; }
static inline int MPI_Abort(MPI_Comm c, int code) { (void)c; exit(code); return code; }
static inline int MPI_Comm_rank(MPI_Comm c, int *r) { (void)c; *r = 0; return 0; }
static inline int MPI_Comm_size(MPI_Comm c, int *s) { (void)c; *s = 1; return 0; }
static inline int MPI_Barrier(MPI_Comm c) { (void)c; return 0; }
static inline int MPI_Comm_split(MPI_Comm c, int color, int key, MPI_Comm *out) {
  (void)c; (void)key;
  *out = color == MPI_UNDEFINED ? MPI_COMM_NULL : MPI_COMM_SELF;
  return 0;
}
static inline int MPI_Comm_free(MPI_Comm *c) { *c = MPI_COMM_NULL; return 0; }
static inline int MPI_Allreduce(const void *in, void *out, int n, MPI_Datatype t, MPI_Op op,
                                MPI_Comm c) {
  (void)op; (void)c;
  if (in != out) memcpy(out, in, (size_t)n * pevsl_mpi_size_of(t));
  return 0;
}
static inline int MPI_Reduce(const void *in, void *out, int n, MPI_Datatype t, MPI_Op op,
                             int root, MPI_Comm c) {
  (void)root;
  return MPI_Allreduce(in, out, n, t, op, c);
}
static inline int MPI_Bcast(void *buf, int n, MPI_Datatype t, int root, MPI_Comm c) {
  (void)buf; (void)n; (void)t; (void)root; (void)c;
  return 0;
}
static inline double MPI_Wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

#endif /* PEVSL_USE_SYSTEM_MPI */
#endif /* PEVSL_MPI_H */
