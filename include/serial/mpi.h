/*
 * serial/mpi.h -- for drivers that include <mpi.h> directly (e.g. the
 * reference's TESTS/Lap/LapPLanN.c:3) and run on one process: add
 * -I include/serial to use the single-rank names of pevsl_mpi.h instead of a
 * system MPI.  Not needed when a real MPI is on the include path.
 */
#ifndef PEVSL_SERIAL_MPI_H
#define PEVSL_SERIAL_MPI_H
#include "../pevsl_mpi.h"
#endif
