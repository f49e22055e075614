/*
 * pevsl_struct.h -- public types of the pEVSL C API.
 *
 * Field names, types and order follow INC/pevsl_struct.h so that drivers
 * written against the reference (e.g. TESTS/Lap/LapPLanN.c, which touches
 * A.ncol_global, A.ncol_local, A.first_col, pol.deg/bar/gam/type, Y->data)
 * compile unchanged.  Two trailing fields are additions of this library:
 * pevsl_Parcsr.dev and pevsl_Data.dev hold the device-resident state (CSR
 * arrays in HBM, HIP stream, workspaces).  They are owned by the library and
 * must be left alone by callers.
 */
#ifndef PEVSL_STRUCT_H
#define PEVSL_STRUCT_H

#include <stddef.h>
#include <stdlib.h>

#include "pevsl_mpi.h"

/* COO, 0-based (INC/pevsl_struct.h:12-15) */
typedef struct _pevsl_cooMat {
  int nrows, ncols, nnz, *ir, *jc;
  double *vv;
} pevsl_Coo;

/* CSR, 0-based, nnz == ia[nrows] (INC/pevsl_struct.h:23-26) */
typedef struct _pevsl_csrMat {
  int nrows, ncols, *ia, *ja;
  double *a;
} pevsl_Csr;

#define PEVSL_CSRNNZ(A) A->ia[A->nrows]

/* halo-exchange plan of the parallel matvec (INC/pevsl_struct.h:37-52).
 * With one rank there is no halo: every count is zero and every pointer is
 * NULL.  Kept so that code inspecting A->comm_handle still compiles. */
typedef struct _commHandle {
  int num_proc_send_to;
  int *proc_send_to;
  int *send_elmts_ids;
  int *send_elmts_ptr;
  int num_proc_recv_from;
  int *proc_recv_from;
  int *recv_elmts_ptr;
  double *send_buf;
  double *recv_buf;
  void *send_recv_requests;
  void *send_recv_status;
} commHandle;

/* parallel CSR (INC/pevsl_struct.h:59-103) */
typedef struct _pevsl_parcsr {
  MPI_Comm comm;
  commHandle *comm_handle;
  int nrow_global;
  int ncol_global;
  int nrow_local;
  int ncol_local;
  int *row_starts;
  int *col_starts;
  int first_row;
  int first_col;
  pevsl_Csr *diag;
  pevsl_Csr *offd;
  int *col_map_offd;
  /* addition: device copy of the matrix (library-owned) */
  void *dev;
} pevsl_Parcsr;

/* parallel vector, host data (INC/pevsl_struct.h:109-116) */
typedef struct _pevsl_parvec {
  MPI_Comm comm;
  int n_global;
  int n_local;
  int n_first;
  double *data;
} pevsl_Parvec;

/* parallel multi-vector, column-major, ld >= n_local
 * (INC/pevsl_struct.h:126-135) */
typedef struct _pevsl_parvecs {
  MPI_Comm comm;
  int n_global;
  int n_local;
  int n_first;
  int n_vecs;
  int ld;
  double *data;
} pevsl_Parvecs;

/* polynomial filter parameters (INC/pevsl_struct.h:143-173) */
typedef struct _pevsl_polparams {
  int max_deg;
  int min_deg;
  int damping;       /* 0 none, 1 Jackson, 2 Lanczos sigma */
  double thresh_ext;
  double thresh_int;
  double intvtol;
  int type;          /* 0 middle, 1 left, 2 right interval */
  double *mu;        /* Chebyshev coefficients, malloc'd by FindPol */
  double cc;         /* centre of [lmin, lmax] */
  double dd;         /* half-width of [lmin, lmax] */
  double gam;        /* centre of the delta function */
  double bar;        /* p(theta) >= bar marks a wanted Ritz value */
  int deg;           /* 0 on input: degree chosen by FindPol */
} pevsl_Polparams;

/* callback prototypes (INC/pevsl_struct.h:190-208, the host-array forms) and
 * their holders (INC/pevsl_struct.h:214-239) */
typedef void (*SVFuncC)(double *br, double *bz, double *xr, double *xz,
                        void *data);
typedef void (*SVFunc)(double *b, double *x, void *data);
typedef void (*MVFunc)(double *x, double *y, void *data);

typedef struct _pevsl_Matvec {
  MVFunc func;
  void *data;
} pevsl_Matvec;

typedef struct _pevsl_Bsol {
  SVFunc func;
  void *data;
} pevsl_Bsol;

typedef struct _pevsl_LtSol {
  SVFunc func;
  void *data;
} pevsl_Ltsol;

/* timers and counters (INC/pevsl_struct.h:241-272) */
typedef struct _pevsl_Stat {
  double t_setBsv;
  double t_setASigBsv;
  double t_eigbounds;
  double t_dos;
  double t_iter;
  double t_mvA;
  double t_mvB;
  double t_svB;
  double t_svLT;
  double t_svASigB;
  double t_reorth;
  double t_eig;
  double t_blas;
  double t_ritz;
  double t_polAv;
  double t_ratAv;
  double t_sth;
  size_t n_mvA;
  size_t n_mvB;
  size_t n_svB;
  size_t n_svLT;
  size_t n_svASigB;
  size_t n_polAv;
  size_t n_ratAv;
  size_t alloced;
  size_t alloced_total;
  size_t alloced_max;
} pevsl_Stat;

/* instance (INC/pevsl_struct.h:278-308) */
typedef struct _pevsl_Data {
  MPI_Comm comm;
  int N;
  int n;
  int nfirst;
  int ifGenEv;
  pevsl_Matvec *Amv;
  pevsl_Matvec *Bmv;
  pevsl_Bsol *Bsol;
  pevsl_Ltsol *LTsol;
  pevsl_Stat *stats;
  int nev_computed;
  double *eval_computed;
  pevsl_Parvecs *evec_computed;
  double sigma_mult;
  /* addition: device context (stream, matrix handle, options) */
  void *dev;
} pevsl_Data;

#endif
