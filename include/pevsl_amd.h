/*
 * pevsl_amd.h -- extensions of the MI355X library beyond the reference API.
 *
 * None of these entry points has a counterpart in INC/pevsl.h; each says what
 * it adds.  All are optional: a driver written against pevsl.h alone gets the
 * reference behaviour (FP64 filter, libc rand() stream, one slice per call).
 */
#ifndef PEVSL_AMD_H
#define PEVSL_AMD_H

#include <stddef.h>
#include <stdio.h>

#include "pevsl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Diagnostics of one filtered-Lanczos solve (one slice). */
typedef struct {
  int its;          /* Lanczos steps taken                                  */
  int kdim;         /* dimension of T at the final eigen-solve              */
  int nbreak;       /* lucky breakdowns (fresh filtered random vectors)     */
  int nconv;        /* Ritz values of p(A) >= bar at the last test           */
  int nev;          /* accepted eigenpairs                                  */
  int deg;          /* filter degree                                        */
  long long nmatvec;/* products with A (filter + Ritz residuals)            */
  double tr;        /* restricted trace at the last test                    */
  double t_total;   /* wall seconds of the solve                            */
  double tol_accept;/* residual threshold applied to the Ritz pairs: the
                       caller's tol, or on the FP32-filter path
                       max(tol, mixed_tol * rho) (pEVSL_SetMixedTol)        */
} pevsl_LanInfo;

/* Accumulated device time (seconds, HIP events) by phase, and algorithmic
 * bytes moved by the filter kernels (reset by pEVSL_StatsReset). */
typedef struct {
  double t_filter;        /* Chebyshev filter SpMV/SpMM steps          */
  double t_reorth;        /* Lanczos updates + CGS-DGKS                 */
  double t_eig;           /* tridiagonal eigen-solves                   */
  double t_ritz;          /* Ritz vectors, residuals                    */
  double t_kpm;           /* KPM moments                                */
  double filter_bytes;    /* algorithmic bytes of the filter steps      */
  double reorth_bytes;    /* algorithmic bytes of the reorthogonalization */
  long long filter_steps; /* column-degree steps                        */
  long long filter_launches;
  double t_cgs;           /* Gram-Schmidt passes only (GEMV-T + GEMV-N)  */
  double cgs_bytes;       /* algorithmic bytes of those passes           */
  long long cgs_passes;   /* passes run (first + DGKS second passes)     */
} pevsl_KernelStats;

/* Filter precision: FP64 (reference) or FP32 vectors and matrix values with
 * FP32 accumulation.  Lanczos coefficients, reorthogonalization, the
 * tridiagonal eigen-solve and Ritz residuals stay FP64 either way. */
#define PEVSL_PREC_FP64 0
#define PEVSL_PREC_FP32 1
int pEVSL_SetFilterPrecision(pevsl_Data *pevsl, int prec);

/* Storage of the Lanczos basis on the FP32-filter path: PEVSL_PREC_FP64
 * (default) or PEVSL_PREC_FP32.  With FP32 storage the reorthogonalization
 * reads half the bytes; its dots and updates still accumulate in FP64, the
 * tridiagonal problem, the Rayleigh-Ritz refinement and the residuals stay
 * FP64, and the basis is orthonormal to FP32 rounding.  Ignored when the
 * filter runs in FP64 or A is a host callback.  No counterpart in the
 * reference. */
int pEVSL_SetBasisPrecision(pevsl_Data *pevsl, int prec);

/* Relative residual threshold of the FP32-filter path (default 1e-6): a
 * Ritz pair is accepted when ||A u - theta u|| < max(tol, mixed_tol * rho),
 * rho = max(|lmin|, |lmax|) of the spectrum frame (1.2e-5 for the 3-D
 * Laplacian's [0, 12]).  Ritz vectors built from an FP32-filtered Krylov
 * basis carry residuals of about 1e-7 rho, while their FP64 Rayleigh-Ritz
 * values stay accurate to ~1e-12 relative; eigenvalue errors are reported
 * separately.  On the FP32-filter path this threshold therefore replaces a
 * smaller tol passed to pEVSL_ChebLanNr / pEVSL_ChebLanNrBatch; the value
 * applied is reported per slice in pevsl_LanInfo.tol_accept.  The FP64 path
 * always applies the caller's tol. */
int pEVSL_SetMixedTol(pevsl_Data *pevsl, double tol);

/* Ritz stage: 0 = the reference's per-vector Ritz pairs (SRC/cheblanNr.c:
 * 339-413); 1 = FP64 Rayleigh-Ritz refinement of the selected Ritz vectors
 * with A ((U^T A U) c = theta (U^T U) c, Gram products on FP64 MFMA + SYGVD) before
 * the same acceptance test; -1 (default) = refinement only with the FP32
 * filter.  The refinement separates eigenvectors that the filter resolves
 * only to its own precision. */
int pEVSL_SetRitzRefine(pevsl_Data *pevsl, int mode);

/* Random vectors: PEVSL_RAND_LIBC draws (rand()-RAND_MAX/2)/(RAND_MAX/2) from
 * libc on the host in the reference order (SRC/parvec.c:65-72);
 * PEVSL_RAND_RADEMACHER / PEVSL_RAND_UNIFORM draw +-1 / uniform[-1,1] from a
 * counter-based generator on the device (seeded). */
#define PEVSL_RAND_LIBC 0
#define PEVSL_RAND_RADEMACHER 1
#define PEVSL_RAND_UNIFORM 2
int pEVSL_SetRandMode(pevsl_Data *pevsl, int mode, unsigned long long seed);

/* Whether ChebLanNr copies the Ritz vectors back (*Wo); 1 by default. */
int pEVSL_SetKeepVectors(pevsl_Data *pevsl, int keep);

/* Least-squares Chebyshev approximations for generalized problems
 * (SRC/lspolapprox.c:63-210): of B^{-1} (SetupLSPolRec, a B-solve), and of
 * any positive function ffun of B (SetupLSPol).  pEVSL_LSPolInfo returns the
 * chosen degree, the deg + 1 coefficients (mu may be NULL) and the interval
 * map x -> (x - cc)/dd.  Free with pEVSL_LSPolFree. */
void pEVSL_SetupLSPolRec(int max_deg, double tol, double lmin, double lmax,
                         pevsl_Parcsr *B, void **vdata);
void pEVSL_SetupLSPol(int max_deg, double tol, double lmin, double lmax,
                      pevsl_Parcsr *B, double (*ffun)(double), void **vdata);
int pEVSL_LSPolInfo(void *vdata, int *deg, double *mu, double *cc, double *dd);

/* Seeds the library's rand() stream.  Random vectors in PEVSL_RAND_LIBC
 * mode (pEVSL_ParvecRand, the KPM and Lanczos-DOS vectors) follow glibc's
 * rand() algorithm after srand(seed), in a generator private to the library,
 * so other users of libc rand() in the process do not shift them.
 * pEVSL_Start seeds it with 0, as the reference's srand(rank) on one rank
 * (SRC/pevsl.c:36). */
void pEVSL_RandSeed(unsigned seed);

/* Kernel of the one-column filter step: lanes per row of the vector-CSR
 * kernel (1..64, power of two), PEVSL_SPMV_STREAM for the CSR-stream kernel
 * (row blocks staged in LDS), or 0 = automatic; and the block width of
 * batched solves (1..64, power of two; 0 = 8 columns per block for an
 * FP32 filter and 4 for FP64, the widths pEVSL_Autotune picks on the
 * 7-point stencil). */
#define PEVSL_SPMV_STREAM 128
int pEVSL_SetSpmvTuning(pevsl_Data *pevsl, int lanes_per_row, int block_width);

/* Block SpMM kernel of batched solves: 1 = one thread per row with B-wide
 * row gathers (B <= 32) and the reference's three-term recurrence with an
 * accumulator updated every other degree (4 vector passes per degree on
 * average, the sums of one update per degree bit for bit), 2 = the same
 * kernel in Clenshaw form (4 passes), 3 / 4 = the nnz-balanced forms of
 * 1 / 2 (products formed in parallel over nonzeros and staged in LDS, long
 * rows reduced by a whole workgroup: for uneven row lengths), 0 = one lane
 * per (row, column), 5 = the Clenshaw form with one thread per (row, 16-byte
 * piece of columns): up to 64 FP32 columns share one pass over the matrix,
 * each piece active from its own top degree (FP32 filter only; 1 on FP64).
 * Default (-1): 2 for an FP32 filter, 1 for FP64, or 4 / 3 when the
 * matrix has a row longer than 64 entries.  A group of slices whose top
 * degree exceeds 400 runs the three-term form of its kernel (2 -> 1,
 * 4 -> 3, 5 -> 1 or 3): Clenshaw sums lose accuracy at high degree. */
int pEVSL_SetBlockKernel(pevsl_Data *pevsl, int rowkernel);

/* Eigenvalues >= thr of the symmetric tridiagonal matrix (diag[0..n-1],
 * sdiag[0..n-2]) by Sturm multisection on the GPU (thr = -HUGE_VAL: all), in
 * ascending order in eigval[0..*count-1], and their sum in *trace.  The
 * restricted-trace test of ChebLanNr uses the same kernels
 * (SRC/cheblanNr.c:298-321 computes it with LAPACK DSTEV). */
int pEVSL_SymmTridEigvals(int n, const double *diag, const double *sdiag, double thr,
                          double *eigval, int *count, double *trace);

/* Batched filtered Lanczos: nslices slices [intvs[4*s], intvs[4*s+1]] (with
 * bounds intvs[4*s+2..3]) advance together, one column of a block SpMM each,
 * with their own filters pols[s] (from FindPol), maxits[s], Lanczos state,
 * breakdown handling and stopping test.  Outputs per slice as ChebLanNr:
 * nevOut[s], lamo[s], reso[s] (malloc'd) and Wo[s] (NULL when vectors are
 * not kept).  A one-slice call is pEVSL_ChebLanNr. */
int pEVSL_ChebLanNrBatch(pevsl_Data *pevsl, int nslices, double *intvs,
                         int *maxits, double tol, pevsl_Parvec *vinit,
                         pevsl_Polparams *pols, int *nevOut, double **lamo,
                         pevsl_Parvecs **Wo, double **reso, FILE *fstats);

/* Slices per wave of pEVSL_ChebLanNrBatch (0, the default: all slices in
 * one batch).  With wave > 0 a call with more slices solves them in
 * ceil(nslices / wave) batches in order of filter degree, one after the
 * other: first the nslices % wave lowest-degree slices (if any), then full
 * batches of `wave` slices.  Peak device memory is
 * then that of the largest wave
 * (each slice's Lanczos basis grows to n x its step count, in 8-byte FP64
 * or 4-byte FP32 elements).  Outputs are per slice as for one batch. */
int pEVSL_SetBatchWave(pevsl_Data *pevsl, int wave);

/* Filter groups of pEVSL_ChebLanNrBatch replayed from hipGraphs (on = 1;
 * default 0: eager launches).  The first step of each group captures the
 * kernels of one filter application (pack, one launch per degree, unpack)
 * with the group's own coefficients and column pointers on the device; later
 * steps restage the pointers and replay the graph.  Results are bitwise those
 * of eager launches. */
int pEVSL_SetGraphFilter(pevsl_Data *pevsl, int on);

/* 16-bit column offsets (no counterpart in the reference).  At
 * pEVSL_ParcsrSetup each nonzero's column is also stored as a 16-bit offset
 * from the first row of its row block (at most 256 rows and 2,048 nonzeros),
 * when every offset fits (banded orderings: the natural order of a grid
 * Laplacian; random sparsity does not fit).  The CSR-stream and row-per-
 * thread block kernels then read 2 index bytes per nonzero instead of 4;
 * results are bitwise the same.  on = 1 (default) uses them where the
 * matrix has them, 0 keeps the 32-bit indices for this instance.
 * pEVSL_ColumnOffsets16 returns 1 when the instance's matrix is read
 * through them. */
int pEVSL_SetColumnOffsets16(pevsl_Data *pevsl, int on);
int pEVSL_ColumnOffsets16(pevsl_Data *pevsl);

/* Return code of pEVSL_ChebLanNrBatch / pEVSL_ChebLanNr when a Lanczos
 * basis does not fit in device memory; the call releases what it held and
 * returns no eigenpairs (nevOut 0, NULL outputs). */
#define PEVSL_ERR_DEVICE_OOM (-7)

/* y = p((A - cc)/dd) v for host vectors v, y (w unused: the work vectors
 * live on the device).  Same arguments as the reference's internal
 * pEVSL_ChebAv (INC/pevsl_int.h:25, SRC/chebpol.c:509-574); exported for
 * callers that apply a designed filter outside the Lanczos solve. */
int pEVSL_ChebAv(pevsl_Data *pevsl, pevsl_Polparams *pol, pevsl_Parvec *v,
                 pevsl_Parvec *y, pevsl_Parvec *w);

/* Classical Gram-Schmidt of v against Q(:, 0:k-1) with the DGKS test (at
 * most i_max passes; another pass while the new norm is at most 1/sqrt(2)
 * of the previous one), *nrmv = |v| after the last pass; w unused.  Same
 * arguments as the reference's internal CGS_DGKS (INC/pevsl_int.h:37,
 * SRC/miscla.c:211-240), on host vectors; the work runs in the device
 * kernels of the Lanczos solve's reorthogonalization (FP64). */
int pEVSL_CGS_DGKS(pevsl_Data *pevsl, int k, int i_max, pevsl_Parvecs *Q, pevsl_Parvec *v,
                   double *nrmv, double *w);

/* Gram products G = U^T U and H = U^T A U of the FP64 Rayleigh-Ritz stage
 * (mixed path): 0 = rocBLAS strided-batched DGEMM over row parts (whole
 * products), 1 = the library's own FP64 MFMA kernel (v_mfma_f64_16x16x4
 * tiles over row parts, summed in a fixed order; only the tiles on and above
 * the diagonal, the triangle the pencil's solver reads), -1 (default) = 1. */
int pEVSL_SetGramKernel(pevsl_Data *pevsl, int kernel);

/* Times both Gram kernels on C = X^T Y for random n x m FP64 X, Y (HIP
 * events, `reps` calls each) and returns the largest difference of the two
 * results relative to the largest entry.  upper = 1: the MFMA kernel forms
 * only the upper triangle (as for the Rayleigh-Ritz pencil) and returns it
 * in both triangles; the difference is taken over the upper triangle, and
 * is HUGE_VAL if the lower triangle does not mirror it exactly. */
int pEVSL_BenchGram(pevsl_Data *pevsl, int n, int m, int reps, int upper, double *sec_blas,
                    double *sec_mfma, double *maxdiff);

/* Diagnostics of slice s of the last ChebLanNr / ChebLanNrBatch call. */
int pEVSL_GetLanInfo(pevsl_Data *pevsl, int s, pevsl_LanInfo *info);
int pEVSL_GetKernelStats(pevsl_Data *pevsl, pevsl_KernelStats *st);

/* Live timing of the hot kernels: HIP events on the instance's stream around
 * each run of back-to-back launches of one kernel (the filter steps of a
 * block or column, a Gram-Schmidt pass), accumulated per kernel since the
 * last pEVSL_StatsReset: launches, seconds, algorithmic bytes.  Copies up
 * to maxn entries to out and sets *count to the number recorded. */
typedef struct {
  char name[48];      /* e.g. "k_clen_blkrow<float,8>"                     */
  long long launches;
  double seconds;
  double bytes;
} pevsl_KernelTime;
int pEVSL_GetKernelTimes(pevsl_Data *pevsl, int maxn, pevsl_KernelTime *out, int *count);

/* y = A x with device pointers (HBM), on the instance's stream; blocks
 * until done. */
int pEVSL_DeviceMatvec(pevsl_Data *pevsl, const double *d_x, double *d_y);

/* y = p((A - cc)/dd) v as pEVSL_ChebAv, on vectors already in HBM (d_v,
 * d_y: n doubles each), allocating nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The filter's work
 * vectors live in a library-owned device array of 3 * 2^21 doubles (48 MiB)
 * per GPU, held until the process ends, so n is limited to
 * pEVSL_ChebAvNoAllocMaxRows() = 2^21 rows; larger n returns -5 with a
 * message.  Needs a device matrix (pEVSL_SetAParcsr) and the standard
 * problem, else PEVSL_NOT_IMPLEMENT.  Same filter kernels and result as
 * pEVSL_ChebAv; blocks until done.  All calls on one GPU share the array:
 * calls that overlap on one GPU race. */
int pEVSL_ChebAvNoAlloc(pevsl_Data *pevsl, pevsl_Polparams *pol, const double *d_v,
                        double *d_y);
size_t pEVSL_ChebAvNoAllocMaxRows(void);

/* Device name and HBM size of the current device; 0 on success. */
int pEVSL_DeviceInfo(char *name, int len, size_t *hbm_bytes);

/* Autotuner of the fused filter step on the instance's matrix.  Times
 * `reps` back-to-back Chebyshev steps (HIP events) for every candidate and
 * records the fastest:
 *   one column: lanes per row in {1,2,4,8,16,32} or the CSR-stream
 *               kernel                                       -> lanes
 *   block SpMM: widths {2,4,8,16,32,64}, per-column cost     -> block_width
 * for the instance's filter precision, and installs the choice (as
 * pEVSL_SetSpmvTuning).  Results: seconds per step of each candidate
 * (t_lanes[i] for lanes 2^i, t_block[i] for width 2^(i+1)) and the
 * algorithmic bytes of one step (bytes_lanes, bytes_block[i]). */
typedef struct {
  int lanes;          /* chosen: lanes per row, or PEVSL_SPMV_STREAM */
  int block_width;
  double t_lanes[6];
  double bytes_lanes;
  double t_block[6];
  double bytes_block[6];
  double t_stream;    /* CSR-stream kernel, seconds per step */
  double t_blockrow[6]; /* row-per-thread block kernel, widths 2..64 (0: n/a) */
  int block_rowkernel;  /* chosen block kernel (pEVSL_SetBlockKernel codes) */
  double t_blockclen[6];  /* Clenshaw row kernel, widths 2..64 (FP32 only) */
  double bytes_clen[6];   /* its algorithmic bytes per step */
  double t_blockbal[6];   /* nnz-balanced three-term kernel, widths 2..64 */
  double t_blockbalclen[6];  /* nnz-balanced Clenshaw kernel (FP32 only) */
  double t_blockseg[6];   /* segmented Clenshaw kernel, widths 4..64 (FP32 only) */
} pevsl_TuneResult;
int pEVSL_Autotune(pevsl_Data *pevsl, int reps, pevsl_TuneResult *res);

/* Times `reps` back-to-back filter steps of width B (power of two <= 32)
 * with the block kernel a batched solve of the instance would use
 * (pEVSL_SetBlockKernel; Clenshaw forms need the FP32 filter): seconds per
 * step and its algorithmic bytes.  -1 on invalid arguments. */
int pEVSL_BenchBlock(pevsl_Data *pevsl, int B, int reps, double *sec, double *bytes);

/* Times `reps` steps of a batched filtered-Lanczos solve at a fixed basis
 * size k, after `warmup` untimed ones: every slice s (interval intvs[4s..],
 * filter pols[s]) holds k + 2 basis columns, and one step runs what one
 * iteration of pEVSL_ChebLanNrBatch runs -- the filter groups (block SpMM
 * steps at the slices' own degrees), the three-term update and the
 * Gram-Schmidt pass against k + 1 columns -- and its per-step
 * synchronization.  The instance's kernel statistics and per-kernel timers
 * (pEVSL_GetKernelTimes) are reset after the warm-up, so they cover the
 * timed steps.  With graph_reps > 0 the first filter group's launch sequence
 * is also captured in a HIP graph and replayed graph_reps times against as
 * many eager calls.  Needs the device matrix (pEVSL_SetAParcsr); -1 on
 * invalid arguments.  No counterpart in the reference. */
typedef struct {
  int nslices, k, reps;
  double sec_step;       /* HIP-event time of one step                      */
  double wall_step;      /* host wall time of one step, synchronization incl. */
  double sec_filter;     /* per step: filter groups (pack, steps, unpack)    */
  double sec_reorth;     /* per step: three-term update + Gram-Schmidt       */
  double sec_cgs;        /* per step: Gram-Schmidt passes                    */
  double filter_bytes;   /* algorithmic bytes per step                       */
  double cgs_bytes;
  long long filter_launches;  /* block/column filter launches per step       */
  double eager_group_sec;     /* first group: one eager call (wall)          */
  double graph_group_sec;     /* the same sequence as one graph replay (wall) */
  int graph_nodes, graph_group_width;
} pevsl_StepBench;
int pEVSL_BenchBatchStep(pevsl_Data *pevsl, int nslices, double *intvs, pevsl_Polparams *pols,
                         int k, int warmup, int reps, int graph_reps, pevsl_StepBench *out);

/* Times `reps` launches of a plain copy kernel over n doubles (read n,
 * write n; 16-byte accesses) on the current device: seconds per launch and
 * its bytes (16 n).  A bandwidth reference for the other kernels and the
 * known-size kernel used to calibrate hardware byte counters.  No
 * counterpart in the reference. */
int pEVSL_BenchCopy(size_t n, int reps, double *sec, double *bytes);

/* Times `reps` classical Gram-Schmidt passes of one vector against a basis
 * of k FP64 columns (the reorthogonalization kernels); returns seconds per
 * pass and its algorithmic bytes (2 reads of the basis + vector traffic). */
int pEVSL_BenchReorth(pevsl_Data *pevsl, int k, int reps, double *sec,
                      double *bytes);

/* ---- Learned slicer (no counterpart in the reference) -------------------
 * A fully connected network (GELU hidden layers) trained and evaluated by
 * this library's own HIP kernels proposes the density of states of an
 * interval from a cheap KPM run; slices and the eigenvalue count follow from
 * it, filter degrees from pEVSL_FindPol on the proposed slices.  Full KPM
 * (pEVSL_Kpmdos) produces the training targets and judges a proposal
 * (pEVSL_SlicerJudge); pEVSL_SpslicerKpm is the fallback.  It stands in for
 * the outputs of pEVSL_Kpmdos (INC/pevsl.h:17-18) and pEVSL_SpslicerKpm
 * (INC/pevsl.h:133). */
typedef struct _pevsl_Mlp pevsl_Mlp;

/* nlayers hidden layers of width nhidden between nin inputs and nout
 * outputs.  The output z is added to a per-sample prior q (NULL: zero); the
 * first nsoft values of y = z + q are a softmax group (cross-entropy
 * against target masses), the rest are regressions (squared error).  Hidden
 * weights are He-normal from `seed`, the output layer starts at zero. */
int pEVSL_MlpCreate(int nin, int nhidden, int nlayers, int nout, int nsoft,
                    unsigned long long seed, pevsl_Mlp **mlp);
int pEVSL_MlpFree(pevsl_Mlp *mlp);
int pEVSL_MlpShape(pevsl_Mlp *mlp, int *nin, int *nhidden, int *nlayers, int *nout,
                   int *nsoft, size_t *nweights);
/* Parameters packed layer by layer: W_l (out x in, row-major), then b_l. */
int pEVSL_MlpSetWeights(pevsl_Mlp *mlp, const float *w, size_t nw);
int pEVSL_MlpGetWeights(pevsl_Mlp *mlp, float *w, size_t nw);
/* Y[batch x nout]: softmax probabilities, then regression values.
 * X is batch x nin and Q batch x nout, row-major host arrays. */
int pEVSL_MlpForward(pevsl_Mlp *mlp, const float *X, const float *Q, int batch, float *Y);
/* Mean loss over the batch and its gradient (nweights floats, may be NULL). */
int pEVSL_MlpLossGrad(pevsl_Mlp *mlp, const float *X, const float *Q, const float *T,
                      int batch, float *loss, float *grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at learning rate lr. */
int pEVSL_MlpTrainStep(pevsl_Mlp *mlp, const float *X, const float *Q, const float *T,
                       int batch, float lr, float *loss);
/* `epochs` shuffled passes over a device-resident training set, minibatches
 * of `batch`, Adam with cosine decay of lr to 2%; losses[e] = mean loss of
 * epoch e (may be NULL).  verbose > 0 prints the loss every `verbose`
 * epochs to stdout. */
int pEVSL_MlpTrain(pevsl_Mlp *mlp, const float *X, const float *Q, const float *T,
                   int nsamples, int batch, int epochs, float lr, unsigned long long seed,
                   float *losses, int verbose);
int pEVSL_MlpSave(pevsl_Mlp *mlp, const char *path);
int pEVSL_MlpLoad(const char *path, pevsl_Mlp **mlp);

#define PEVSL_SLICER_MDEG 20   /* cheap KPM degree */
#define PEVSL_SLICER_NVEC 8    /* cheap KPM random vectors */
#define PEVSL_SLICER_NBINS 512 /* DOS bins over the interval */
#define PEVSL_SLICER_NFREQ 24  /* Fourier features per interval end */
#define PEVSL_SLICER_NIN (42 + 4 * PEVSL_SLICER_NFREQ) /* network inputs */
/* Network output layout: NBINS softmax logits, then log(fraction of the
 * spectrum inside the interval). */

/* Fractions of the spectrum in nbins equal bins of [intv[0], intv[1]]
 * (clipped to the frame [intv[2], intv[3]]) from KPM moments mu[0..Mdeg] in
 * the convention of pEVSL_Kpmdos; *frac = their sum. */
int pEVSL_KpmBinMasses(const double *mu, int Mdeg, const double *intv, int nbins,
                       double *masses, double *frac);
/* Equal-share cuts of a binned DOS over [intv[0], intv[1]] (linear inside a
 * bin): sli[0..n_int]. */
int pEVSL_SlicerFromBins(const double *masses, int nbins, const double *intv, int n_int,
                         double *sli);
/* Network inputs (PEVSL_SLICER_NIN) and prior (NBINS + 1) from cheap KPM
 * moments mu[0..Mdeg], Mdeg >= PEVSL_SLICER_MDEG (the first MDEG are used). */
int pEVSL_SlicerFeaturesFromMu(const double *mu, int Mdeg, int n, const double *intv,
                               float *feat, float *prior);
/* The same after running the cheap KPM on the instance's matrix. */
int pEVSL_SlicerFeatures(pevsl_Data *pevsl, const double *intv, float *feat, float *prior);
/* Proposed slices sli[0..n_int] of intv[0..3] and the eigenvalue count in
 * [intv[0], intv[1]]; returns -4 when mlp is not a slicer network. */
int pEVSL_SpslicerMlp(pevsl_Data *pevsl, pevsl_Mlp *mlp, const double *intv, int n_int,
                      double *sli, double *ecnt);
/* Judge: counts[j] = n x (KPM fraction of the spectrum in slice j) from the
 * moments of a full pEVSL_Kpmdos run, *imbalance = max count / mean - 1. */
int pEVSL_SlicerJudge(const double *mu, int Mdeg, const double *intv, double n, int n_int,
                      const double *sli, double *counts, double *imbalance);

#ifdef __cplusplus
}
#endif

#endif
