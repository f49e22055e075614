/*
 * pevsl_mpi.h -- single-rank stand-in for the MPI names that appear in the
 * pEVSL public signatures (INC/pevsl.h:18,40,55,67 carry MPI_Comm and
 * INC/pevsl_def.h:8 includes <mpi.h>).
 *
 * This library runs one process on one GPU: the reference's row partition
 * (diag/offd + halo) collapses into one device-resident CSR and every
 * communicator has exactly one rank.  Callers that compile against a real MPI
 * define PEVSL_USE_SYSTEM_MPI before including pevsl.h; the communicator is
 * then only checked to have size 1.
 */
#ifndef PEVSL_MPI_H
#define PEVSL_MPI_H

#ifdef PEVSL_USE_SYSTEM_MPI
#include <mpi.h>
#else

#include <time.h>

typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0)
#define MPI_COMM_SELF ((MPI_Comm)1)
#define MPI_COMM_NULL ((MPI_Comm)-1)
#define MPI_SUCCESS 0

static inline int MPI_Comm_rank(MPI_Comm c, int *r) { (void)c; *r = 0; return 0; }
static inline int MPI_Comm_size(MPI_Comm c, int *s) { (void)c; *s = 1; return 0; }
static inline int MPI_Barrier(MPI_Comm c) { (void)c; return 0; }
static inline double MPI_Wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

#endif /* PEVSL_USE_SYSTEM_MPI */
#endif /* PEVSL_MPI_H */
