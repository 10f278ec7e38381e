/*
 * jadamilu_amd.h -- C-ABI of the MI355X Jacobi-Davidson eigensolver.
 *
 * Every entry point takes plain pointers (Fortran calling convention for the
 * legacy names: every argument by address, INTEGER = int32, REAL = double).
 * No HIP or torch types appear in any signature.
 *
 * Legacy drop-in symbols (global default context, HOST arrays).  Each replaces
 * the reference routine of the same name:
 *   dpjd_        DPJD        src/jd/PJD.F:5-259      (prototype jadamilu1.h:14-20)
 *   dpjdrevcom_  DPJDREVCOM  src/jd/PJD.F:268-362    (jadamilu1.h:21-27)
 *   djdrevcom_   DJDREVCOM   src/jd/PJD.F:369-427    (jadamilu1.h:29-35)
 *   dpjdcleanup_ DPJDCLEANUP src/jd/PJD.F:435-485    (jadamilu1.h:28)
 *   dpjd_gep_       DPJD_GEP       src/jd/PJD.F:492-570   (jadamilu1.h:36-43)
 *   dpjdrevcom_gep_ DPJDREVCOM_GEP src/jd/PJD.F:576-703   (jadamilu1.h:44-53)
 *   djdrevcom_gep_  DJDREVCOM_GEP  src/jd/PJD.F:709-798   (jadamilu1.h:54-59)
 * The generalized drivers solve A x = lambda B x with B symmetric positive
 * definite (all ISEARCH modes); their reverse
 * communication adds NDX3 and the codes IJOB = 1 (X(NDX2) = A X(NDX1) and
 * X(NDX3) = B X(NDX1)), 3 (A only) and 4 (X(NDX3) = B X(NDX1) only); a mass
 * matrix found not positive definite ends the run with INFO = -54322.
 * The un-underscored names (dpjd, dpjdrevcom, djdrevcom, dpjdcleanup) are
 * exported as aliases, matching a reference build without -D__UNDERSCORE__
 * (src/include/namesjadamilu.h:53-88).
 *
 * Semantics follow the reference: same argument meaning, same INFO codes,
 * same IJOB/NDX reverse-communication protocol (NDX are 1-based offsets into
 * the caller's X, X laid out as in JDaux.F:1534-1567), EIGS/RES/NEIG/ITER/
 * SHIFT/DROPTOL/GAP outputs.  The library never calls exit(): an allocation or
 * HIP failure returns INFO = JDHIP_INFO_DEVICE_ERROR (new code, no reference
 * counterpart).
 *
 * Extension entry points (jdhip_*) take an explicit context and keep the X
 * workspace in DEVICE memory: the reverse-communication NDX offsets then
 * index the device array, so a caller can service IJOB=1 with its own HIP
 * kernel (or with jdhip_spmv) without any host copy.
 */
#ifndef JADAMILU_AMD_H
#define JADAMILU_AMD_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define JDHIP_INFO_DEVICE_ERROR (-54330)

/* ---- legacy symbols: host arrays, global default context ---------------- */
void dpjd_(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
           int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
           int *iter, double *tol, double *shift, double *droptol, double *mem, int *icntl,
           int *iprint, int *info, double *gap);
void dpjdrevcom_(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
                 int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
                 int *iter, double *tol, double *shift, double *droptol, double *mem,
                 int *icntl, int *ijob, int *ndx1, int *ndx2, int *iprint, int *info,
                 double *gap);
void djdrevcom_(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                double *sigma, int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                int *ijob, int *ndx1, int *ndx2, int *iprint, int *info, double *gap);
void dpjdcleanup_(void);

void dpjd(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x, int *lx,
          int *neig, double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
          double *tol, double *shift, double *droptol, double *mem, int *icntl, int *iprint,
          int *info, double *gap);
void dpjdrevcom(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
                int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
                int *iter, double *tol, double *shift, double *droptol, double *mem,
                int *icntl, int *ijob, int *ndx1, int *ndx2, int *iprint, int *info,
                double *gap);
void djdrevcom(int *n, double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
               int *isearch, int *ninit, int *madspace, int *iter, double *tol, int *ijob,
               int *ndx1, int *ndx2, int *iprint, int *info, double *gap);
void dpjdcleanup(void);

void dpjd_gep_(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib, double *eigs,
               double *res, double *x, int *lx, int *neig, double *sigma, int *isearch,
               int *ninit, int *madspace, int *iter, double *tol, double *shift,
               double *droptol, double *mem, int *icntl, int *iprint, int *info, double *gap);
void dpjdrevcom_gep_(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib,
                     double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
                     int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                     double *shift, double *droptol, double *mem, int *icntl, int *ijob,
                     int *ndx1, int *ndx2, int *ndx3, int *iprint, int *info, double *gap);
void djdrevcom_gep_(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                    double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
                    double *tol, int *ijob, int *ndx1, int *ndx2, int *ndx3, int *iprint,
                    int *info, double *gap);
void dpjd_gep(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib, double *eigs,
              double *res, double *x, int *lx, int *neig, double *sigma, int *isearch,
              int *ninit, int *madspace, int *iter, double *tol, double *shift,
              double *droptol, double *mem, int *icntl, int *iprint, int *info, double *gap);
void dpjdrevcom_gep(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib,
                    double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
                    int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                    double *shift, double *droptol, double *mem, int *icntl, int *ijob,
                    int *ndx1, int *ndx2, int *ndx3, int *iprint, int *info, double *gap);
void djdrevcom_gep(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                   double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
                   double *tol, int *ijob, int *ndx1, int *ndx2, int *ndx3, int *iprint,
                   int *info, double *gap);

/* ---- complex Hermitian drivers (host arrays, their own global context) ---
 * Each replaces the reference routine of the same name (argument meaning
 * Doc/ZPJD_prototypes.f; prototypes jadamilu1.h:62-108):
 *   zpjd_           ZPJD            (jadamilu1.h:62-68)
 *   zpjdrevcom_     ZPJDREVCOM      (jadamilu1.h:69-77)
 *   zjdrevcom_      ZJDREVCOM       (jadamilu1.h:79-84)
 *   zpjdcleanup_    ZPJDCLEANUP     (jadamilu1.h:78)
 *   zpjd_gep_       ZPJD_GEP        (jadamilu1.h:85-92)
 *   zpjdrevcom_gep_ ZPJDREVCOM_GEP  (jadamilu1.h:93-103)
 *   zjdrevcom_gep_  ZJDREVCOM_GEP   (jadamilu1.h:104-109)
 * A and X (and B) are DOUBLE COMPLEX arrays passed as double *: each entry
 * is a (real, imaginary) pair of doubles, so X has 2*LX doubles.  The
 * matrices are Hermitian (B positive definite), given as 1-based CSR of one
 * triangle with every (real) diagonal entry, or JA(1) < 0 for a diagonal
 * matrix.  The solve runs on the real symmetric form of order 2N (each entry
 * a + ib as the block [[a, -b], [b, a]], the complex array's own memory
 * layout), for 2*NEIG eigenpairs, keeping one vector per complex direction;
 * reverse-communication requests use the slots X(1:N) (vector in),
 * X(N+1:2N) (A times it) and X(2N+1:3N) (B times it), and for IJOB = 2 the
 * right-hand side X(N+1:2N) and the solution X(1:N).  NINIT is capped at
 * NEIG.  The un-underscored names are exported as aliases. */
void zpjd_(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
           int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
           int *iter, double *tol, double *shift, double *droptol, double *mem, int *icntl,
           int *iprint, int *info, double *gap);
void zpjdrevcom_(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
                 int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
                 int *iter, double *tol, double *shift, double *droptol, double *mem,
                 int *icntl, int *ijob, int *ndx1, int *ndx2, int *iprint, int *info,
                 double *gap);
void zjdrevcom_(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                double *sigma, int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                int *ijob, int *ndx1, int *ndx2, int *iprint, int *info, double *gap);
void zpjdcleanup_(void);
void zpjd_gep_(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib, double *eigs,
               double *res, double *x, int *lx, int *neig, double *sigma, int *isearch,
               int *ninit, int *madspace, int *iter, double *tol, double *shift,
               double *droptol, double *mem, int *icntl, int *iprint, int *info, double *gap);
void zpjdrevcom_gep_(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib,
                     double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
                     int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                     double *shift, double *droptol, double *mem, int *icntl, int *ijob,
                     int *ndx1, int *ndx2, int *ndx3, int *iprint, int *info, double *gap);
void zjdrevcom_gep_(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                    double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
                    double *tol, int *ijob, int *ndx1, int *ndx2, int *ndx3, int *iprint,
                    int *info, double *gap);
/* The real symmetric form of order 2n that the complex drivers solve, for a
 * 1-based one-triangle Hermitian CSR (a: n or nnz complex values as
 * interleaved doubles; ja[0] < 0: diagonal).  Writes ia2 (2n+1), ja2 and a2
 * (1-based) when cap is at least the entry count; returns the entry count
 * (-1 on bad input).  No reference counterpart: exposed for checking. */
long jdhip_hermitian_real_form(int n, const double *a, const int *ja, const int *ia, int *ia2,
                               int *ja2, double *a2, long cap);
void zpjd(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x, int *lx,
          int *neig, double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
          double *tol, double *shift, double *droptol, double *mem, int *icntl, int *iprint,
          int *info, double *gap);
void zpjdrevcom(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
                int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
                int *iter, double *tol, double *shift, double *droptol, double *mem,
                int *icntl, int *ijob, int *ndx1, int *ndx2, int *iprint, int *info,
                double *gap);
void zjdrevcom(int *n, double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
               int *isearch, int *ninit, int *madspace, int *iter, double *tol, int *ijob,
               int *ndx1, int *ndx2, int *iprint, int *info, double *gap);
void zpjdcleanup(void);
void zpjd_gep(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib, double *eigs,
              double *res, double *x, int *lx, int *neig, double *sigma, int *isearch,
              int *ninit, int *madspace, int *iter, double *tol, double *shift,
              double *droptol, double *mem, int *icntl, int *iprint, int *info, double *gap);
void zpjdrevcom_gep(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib,
                    double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
                    int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                    double *shift, double *droptol, double *mem, int *icntl, int *ijob,
                    int *ndx1, int *ndx2, int *ndx3, int *iprint, int *info, double *gap);
void zjdrevcom_gep(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                   double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
                   double *tol, int *ijob, int *ndx1, int *ndx2, int *ndx3, int *iprint,
                   int *info, double *gap);

/* ---- context API: device-resident X ------------------------------------- */
typedef struct jdhip_ctx jdhip_ctx;

/* Create a solver context on HIP device `device`; NULL on failure.
 * No reference counterpart (the reference keeps one COMMON-block state,
 * JDaux.F:30-51); the legacy symbols use one process-global context. */
jdhip_ctx *jdhip_create(int device);
void jdhip_destroy(jdhip_ctx *ctx);

/* As DPJD (PJD.F:5) with x a DEVICE array of lx doubles. */
void jdhip_pjd(jdhip_ctx *ctx, int n, double *a, int *ja, int *ia, double *eigs, double *res,
               double *x_dev, long lx, int *neig, double sigma, int isearch, int ninit,
               int madspace, int *iter, double tol, double *shift, double *droptol,
               double *mem, int *icntl, int iprint, int *info, double *gap);
/* As DPJDREVCOM (PJD.F:268); NDX index the DEVICE array x_dev (1-based). */
void jdhip_pjdrevcom(jdhip_ctx *ctx, int n, double *a, int *ja, int *ia, double *eigs,
                     double *res, double *x_dev, long lx, int *neig, double sigma,
                     int isearch, int ninit, int madspace, int *iter, double tol,
                     double *shift, double *droptol, double *mem, int *icntl, int *ijob,
                     long *ndx1, long *ndx2, int iprint, int *info, double *gap);
/* As DJDREVCOM (PJD.F:369); NDX index the DEVICE array x_dev (1-based). */
void jdhip_jdrevcom(jdhip_ctx *ctx, int n, double *eigs, double *res, double *x_dev, long lx,
                    int *neig, double sigma, int isearch, int ninit, int madspace, int *iter,
                    double tol, int *ijob, long *ndx1, long *ndx2, int iprint, int *info,
                    double *gap);
/* As DPJD_GEP (PJD.F:492): A x = lambda B x, both matrices 1-based upper
 * CSR with every diagonal entry, x a DEVICE array.  The preconditioner
 * factors A - SHIFT B (symilupackfactorgep.c), the default shift is the
 * estimate of SYMILUPACKSAVEDIAGGEP (symamgsavediaggep.c). */
void jdhip_pjd_gep(jdhip_ctx *ctx, int n, double *a, int *ja, int *ia, double *b, int *jb,
                   int *ib, double *eigs, double *res, double *x_dev, long lx, int *neig,
                   double sigma, int isearch, int ninit, int madspace, int *iter, double tol,
                   double *shift, double *droptol, double *mem, int *icntl, int iprint,
                   int *info, double *gap);
/* As DPJDREVCOM_GEP (PJD.F:576); NDX (1-based) index the DEVICE array. */
void jdhip_pjdrevcom_gep(jdhip_ctx *ctx, int n, double *a, int *ja, int *ia, double *b,
                         int *jb, int *ib, double *eigs, double *res, double *x_dev, long lx,
                         int *neig, double sigma, int isearch, int ninit, int madspace,
                         int *iter, double tol, double *shift, double *droptol, double *mem,
                         int *icntl, int *ijob, long *ndx1, long *ndx2, long *ndx3,
                         int iprint, int *info, double *gap);
/* As DJDREVCOM_GEP (PJD.F:709); NDX (1-based) index the DEVICE array. */
void jdhip_jdrevcom_gep(jdhip_ctx *ctx, int n, double *eigs, double *res, double *x_dev,
                        long lx, int *neig, double sigma, int isearch, int ninit, int madspace,
                        int *iter, double tol, int *ijob, long *ndx1, long *ndx2, long *ndx3,
                        int iprint, int *info, double *gap);
/* As DPJDCLEANUP (PJD.F:435): release the preconditioner of ctx. */
void jdhip_cleanup(jdhip_ctx *ctx);

/* Eigenvalues wr + i wi and right eigenvectors v (column-major, ld ldv) of a
 * small real general n x n matrix a (overwritten): the DGEEV step of
 * SOLPROJEP for the harmonic extraction of the generalized problem
 * (JDaux.F:3293-3300), exposed for checking.  A complex pair stores the real
 * and imaginary parts of the first eigenvalue's vector in two columns.
 * 0 on success, k > 0 if the QR iteration did not converge, -1 bad input. */
int jdhip_dense_eig(int n, double *a, int lda, double *wr, double *wi, double *v, int ldv);

/* Approximate minimum degree ordering of the symmetric pattern given by a
 * 1-based upper (or lower) triangle CSR; perm[k] (1-based) is the original
 * index placed at position k.  Replaces the AMD call that orders each level
 * of the reference's factorization (src/ilupack/permamd.c,
 * src/ilupack/sympermMC64amd.c:503-509).  Returns 0, or -1 on bad input. */
int jdhip_order_amd(int n, const int *ia, const int *ja, int *perm);

/* Reset ctx to the state of a fresh process: the preconditioner is released
 * and the COMMON-block state of the reference (/PJDINITPJD/, /JDRVCOMJD/,
 * JDaux.F:30-47, which persists between calls) is zeroed. */
int jdhip_reset(jdhip_ctx *ctx);
/* The process-global context behind the legacy symbols (created on first
 * use on device $JDHIP_DEVICE, default 0). */
jdhip_ctx *jdhip_legacy_context(void);
/* The context behind the complex drivers (created on first use on device
 * $JDHIP_DEVICE, default 0), e.g. for jdhip_set_option.  No reference
 * counterpart. */
jdhip_ctx *jdhip_complex_context(void);

/* HIP stream (as void*) the context launches on; callers servicing IJOB
 * with their own kernels should use it (or synchronise with it). */
void *jdhip_stream(jdhip_ctx *ctx);

/* ---- batches of independent problems ------------------------------------ */
/* One problem of a batch: a 1-based upper CSR on the host and the DPJD
 * parameters (see jdhip_pjd); outputs are written into the item. */
typedef struct {
    int n;
    const int *ia, *ja;
    const double *a;
    int neig, isearch, madspace, maxiter;
    double sigma, tol, shift, droptol, mem;
    int icntl[5];
    double *eigs, *res; /* host, neig entries each */
    double *x;          /* host n*neig eigenvectors (column-major), or NULL */
    int info, iter, neig_out;
    double gap;
} jdhip_batch_item;
/* Solve every item on HIP device `device` with nworkers worker threads, each
 * owning a context and a HIP stream; items are taken largest first.  Each
 * item starts from the state of a fresh process and makes its set-up choices
 * from its inputs alone (options ildl_autotune = 0, ildl_tune_time = 0,
 * cost_clock = 1), so its results equal those of a single jdhip_pjd call with
 * these options.  No reference counterpart (the reference solves one problem
 * per call).  Returns the number of items with INFO != 0, or -1 on invalid
 * arguments. */
int jdhip_pjd_batch(int device, int nitems, jdhip_batch_item *items, int nworkers);
/* jdhip_pjd_batch with options (jdhip_set_option names and values) applied
 * to every worker context after the batch's own settings, e.g.
 * "ildl_sweeps" or "ildl_tune_tol".  "ildl_graph" 1 is refused (-1): while
 * one worker captures a graph, another worker's allocations may
 * synchronise the device and invalidate the capture.  An option that
 * jdhip_set_option rejects leaves every item with INFO = -54330. */
int jdhip_pjd_batch_opts(int device, int nitems, jdhip_batch_item *items, int nworkers,
                         int nopts, const char *const *names, const double *values);

/* ---- symmetric SpMV (MATVECA, JDaux.F:3397-3418) ------------------------- */
/* Upload a 1-based upper-triangle CSR (diagonal anywhere in each row) as the
 * device operator of ctx; 0 on success.  lanes_per_row: 0 = autotune (the
 * formats below timed on this matrix), -1 = sliced ELL (64-row slices, one
 * lane per row, 16-bit column offsets where a slice allows), k > 0 = CSR with
 * k lanes per row. */
int jdhip_set_matrix(jdhip_ctx *ctx, int n, const int *ia, const int *ja, const double *a,
                     int lanes_per_row);
/* y = A x on ctx's stream, x and y DEVICE arrays of n doubles. */
int jdhip_spmv(jdhip_ctx *ctx, const double *x_dev, double *y_dev);

/* Y = A X for k vectors (SpMM, MATVECA of JDaux.F:3397-3418 on a block): x_dev
 * and y_dev DEVICE arrays, column-major with leading dimensions ldx, ldy >= n
 * (the operator set by jdhip_set_matrix).  Every matrix entry is read once
 * per 8 vectors.  0 on success, -1 bad arguments, -2 no matrix / HIP error. */
int jdhip_spmm(jdhip_ctx *ctx, int k, const double *x_dev, long ldx, double *y_dev, long ldy);

/* ---- dense contractions on the FP64 matrix cores ------------------------- */
/* C = V^T W (k x m, k, m <= 64) for tall V (n x k) and W (n x m), DEVICE,
 * column-major: the Gram products of the Rayleigh-Ritz steps (JDaux.F:
 * 1651-1660, 2013-2031, 2165-2200), computed with v_mfma_f64_16x16x4_f64 and
 * summed deterministically; on ctx's stream, with ctx's 4 MiB of partial-sum
 * scratch.  0 on success, -1 bad arguments, -2 HIP error. */
int jdhip_gram(jdhip_ctx *ctx, long n, int k, int m, const double *v_dev, long ldv,
               const double *w_dev, long ldw, double *c_dev, int ldc);
/* jdhip_gram without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * bit-identical result as jdhip_gram, on `stream` (a hipStream_t of the
 * current device), with no context: the partial sums live in a static
 * device array of 4 MiB (524,288 doubles) that each GPU holds until the
 * process ends; k, m <= 64 (larger sizes return -1 with a message on
 * stderr).  Calls that overlap in time on one GPU share that array and race;
 * serialise them.  0 on success, -1 bad arguments, -2 HIP error. */
int jdhip_gram_noalloc(void *stream, long n, int k, int m, const double *v_dev, long ldv,
                       const double *w_dev, long ldw, double *c_dev, int ldc);
/* Y = V Q (n x j) for V (n x k), Q (k x j), k, j <= 64, DEVICE, column-major,
 * on the matrix cores: the basis update of SHRINK (JDaux.F:2418-2462). */
int jdhip_tall_gemm(jdhip_ctx *ctx, long n, int k, int j, const double *v_dev, long ldv,
                    const double *q_dev, int ldq, double *y_dev, long ldy);

/* ---- options ------------------------------------------------------------- */
/* Build options of ctx (no reference counterpart; defaults reproduce the
 * reference's behaviour, except "ildl_sweeps"): "ildl_sweeps" (0 = exact
 * level-scheduled triangular solves, k > 0 = k Jacobi sweeps per triangular
 * solve, -1 = the default: chosen at set-up by timing both on the device,
 * sweeps only where they reproduce the exact apply to "ildl_tune_tol", by
 * default 3e-2 for definite and 1e-2 for indefinite factors),
 * "ildl_autotune" (1 = kernel form of each factor product chosen by timing,
 * the default), "ildl_form" (-1 = automatic; 0, 1, 4, 8, 16, 32, 64 = that form for
 * every product), "ildl_ordering" (0 natural, 1 reverse Cuthill-McKee,
 * 2 approximate minimum degree, the default), "ildl_coupling" (0 = F blocks and
 * two solves with B per level, as the reference, the default; 1 = L_CB blocks,
 * see jdhip_ildl_coupling), "ildl_dense_max" (largest final
 * level factored densely), "spmv_lanes", "ildl_graph" (1: replay the
 * preconditioner apply from captured HIP graphs; default 0), "ildl_storage",
 * "ildl_xcd", "ildl_sigma" (device layout of the factor), "ildl_threads"
 * (host threads of the factorization), "sync_blocking" (1: the iteration's
 * waits for the device sleep on a blocking-sync event instead of spinning;
 * default 0), and for ICNTL(5) != 0: "ildl_vec32"
 * (work vectors of the apply: 1 FP32, 0 FP64, -1 FP32 for definite factors
 * only, the default), "inner_fp32" (1 = FP32 inner solver; default 0: FP64
 * inner solver as in the reference), "inner_spmv64" (FP32 inner solver: 1 =
 * its SpMV reads the FP64 values, the default), "ildl_tune_time" (1 = sweep
 * counts are kept only when timed faster than exact solves, the default; 0 =
 * whenever accurate), "fail_setup" (k > 0: the k-th preconditioner set-up of
 * a call reports failure, to exercise the INFO = -54321 path of
 * JDaux.F:431-441; default 0), "cost_clock" (clock of PJDRVCOM's
 * re-factorisation cost model, JDaux.F:282-449 and 505-535: 0 = wall clock,
 * the default; 1 = every interval reads 0 s, so the reference's floors
 * decide, as on a machine where the problem is small: the set-up sequence and
 * the suggested DROPTOL then follow from the residual history alone).
 * 0 on success, -2 for an
 * unknown name, -3 for an invalid value. */
int jdhip_set_option(jdhip_ctx *ctx, const char *name, double value);

/* z = P^-1 r with the preconditioner currently held by ctx (diagonal or
 * multilevel ILDL); DEVICE arrays.  The apply of SYMILUPACKSOL
 * (src/ilupack/symilupacksol.c:60-128) / DGLPRECSOL (src/jd/dglprecsol.c). */
int jdhip_precond_apply(jdhip_ctx *ctx, double *z_dev, const double *r_dev);

/* ---- multilevel ILDL factor (host; no device needed) --------------------- */
/* Factor A - shift*I (A: 1-based upper CSR) as the reference's
 * SYMILUPACKFACTOR / DSYMAMGfactor (src/ilupack/symilupackfactor.c:66-220,
 * symAMGsetup.c:121-1647) does: multilevel inverse-based ILDL^T with
 * droptol, condest, optional |D| conversion (spd, symspd.c:37) and FP32
 * arithmetic (single, singlesymilupackfactor.c:78-121).  *ierr: 0 ok, -3
 * elbow exceeded, >0 zero pivot in the final dense level. */
typedef struct jdhip_ildl jdhip_ildl;
jdhip_ildl *jdhip_ildl_factor(int n, const int *ia, const int *ja, const double *a, double shift,
                              double droptol, double condest, int spd, int single, int ordering,
                              int dense_max, int coupling, int *ierr);
void jdhip_ildl_free(jdhip_ildl *f);
/* Coupling block each level stores (the lc_* arrays of jdhip_ildl_level_data):
 * 0 = F^T = A_CB of the level's scaled, permuted input, applied with two
 * solves with B per level as DSYMAMGsol does (symAMGsol.c:110-631, F from
 * symAMGextract.c:112), the default; 1 = L_CB, the factor's rows of the
 * deferred pivots, applied as the block LDL^T substitution (one solve with
 * L_B and one with L_B^T per level). */
int jdhip_ildl_coupling(jdhip_ildl *f);
/* number of sparse levels; *dense_m = order of the dense final level */
int jdhip_ildl_nlevels(jdhip_ildl *f, int *dense_m);
int jdhip_ildl_level_sizes(jdhip_ildl *f, int lev, int *n, int *nB, long *nnz_lb, long *nnz_lc);
/* copy out level lev: perm[n] (position -> level input index), scal[n],
 * strict lower L_B (CSR, nB rows), coupling block (CSR, n-nB rows: F^T or
 * L_CB, see jdhip_ildl_coupling), pivots d[nB] */
int jdhip_ildl_level_data(jdhip_ildl *f, int lev, int *perm, double *scal, int *lb_ptr,
                          int *lb_idx, double *lb_val, int *lc_ptr, int *lc_idx, double *lc_val,
                          double *d);
/* block diagonal D of level lev: d[nB] its diagonal, and for 2x2 pivots
 * (the reference's 2x2 blocks, sympiluc.F:1042-1200) dpair[b] = the other
 * row of b's block (b-1 or b+1) with doff[b] = the block's off-diagonal
 * entry; dpair[b] = -1, doff[b] = 0 for a 1x1 pivot.  0, or -1 for a bad
 * level. */
int jdhip_ildl_level_pivots(jdhip_ildl *f, int lev, double *d, double *doff, int *dpair);
/* dense final level: explicit inverse (m x m, row-major) */
int jdhip_ildl_dense(jdhip_ildl *f, double *minv);
int jdhip_ildl_stats(jdhip_ildl *f, double *fill, int *nneg, double *condest, double *seconds);
/* Persisted factor: everything the apply reads (the reference keeps it in
 * memory as the DAMGlevelmat list, src/include/ilupack.h:197-211, for the
 * lifetime of one process; here it can also be written to a file and read
 * back by a later run, then attached with jdhip_ildl_attach and used with
 * ICNTL(2) = -1).  Binary, little-endian, versioned, with a checksum; the
 * layout is described in jadamilu_amd/csrc/persist.cpp.  save: 0, -1 bad
 * arguments, -2 I/O error.  load: the handle, or NULL with *ierr = -2 (I/O)
 * or -3 (not a factor file, version or checksum mismatch, inconsistent
 * structure). */
int jdhip_ildl_save(jdhip_ildl *f, const char *path);
jdhip_ildl *jdhip_ildl_load(const char *path, int *ierr);

/* MatrixMarket coordinate files (real; symmetric, or general holding a
 * symmetric matrix, of which the upper triangle is read; pattern files get
 * values 1) to the reference's input: 1-based upper-triangle CSR with every
 * diagonal entry stored (an absent diagonal entry becomes an explicit 0) and
 * duplicates summed.  info: order and stored entries; read: fills caller
 * arrays ia[n+1], ja[nnz], a[nnz] (nnz_cap >= the info count); write: a
 * symmetric file (lower triangle, 17 significant digits) from such a CSR.
 * 0 on success, -1 bad arguments, -2 I/O error, -3 malformed file, -4
 * unsupported kind (complex, Hermitian, skew), -5 size mismatch. */
int jdhip_mm_info(const char *path, int *n, long *nnz_upper);
int jdhip_mm_read(const char *path, int n, long nnz_cap, int *ia, int *ja, double *a);
int jdhip_mm_write(const char *path, int n, const int *ia, const int *ja, const double *a);
/* upload f as the preconditioner of ctx (sweeps as "ildl_sweeps") */
int jdhip_ildl_attach(jdhip_ctx *ctx, jdhip_ildl *f, int sweeps);
/* dependency levels of the uploaded factor and algorithmic bytes per apply */
int jdhip_ildl_device_info(jdhip_ctx *ctx, int *fwd_levels, int *bwd_levels, long *apply_bytes);
/* Triangular-solve mode of each sparse level of ctx's ILDL preconditioner
 * (0 exact level-scheduled, k Jacobi sweeps) into out[0..cap-1]; returns the
 * number of levels, or -1 without an ILDL preconditioner. */
int jdhip_ildl_level_sweeps(jdhip_ctx *ctx, int *out, int cap);
/* Kernel form of each product of each sparse level, four per level (forward
 * sweep over L_B, backward sweep over L_B^T, L_CB product, L_CB^T product):
 * 0 one lane per row on a SELL-C-sigma copy, 1 CSR-stream (row blocks of at
 * most 2048 entries staged through LDS), k >= 4 CSR with k lanes per row.
 * Chosen by timing at set-up (option ildl_autotune, default 1), else by the
 * average row length.  Fills out[0..cap-1]; returns the number of levels, or
 * -1 without an ILDL preconditioner.  No reference counterpart. */
int jdhip_ildl_level_forms(jdhip_ctx *ctx, int *out, int cap);
/* Diagnostic: time reps launches of one Jacobi sweep over the first level's
 * L_B (the largest single kernel of the preconditioner apply) on ctx's stream
 * with HIP events; *alg_bytes = its algorithmic bytes per launch.  No
 * reference counterpart.  0 on success, -1 without an ILDL preconditioner. */
int jdhip_ildl_time_sweep(jdhip_ctx *ctx, int reps, double *avg_ms, long *alg_bytes);

/* ---- eigenvalue adjoints (no reference counterpart) ---------------------- */
/* Gradient of L = sum_k w[k] lambda_k (simple, gap-separated eigenvalues,
 * unit eigenvectors X(:,k) = x_dev + k*ldx on the device) with respect to the
 * stored entries of the 1-based upper CSR (ia_dev, ja_dev on the device):
 * g[e(i,j)] = (i==j ? 1 : 2) * sum_k w[k] x_k(i) x_k(j).  w: host, k <= 64.
 * Launched on `stream` (a hipStream_t, e.g. jdhip_stream(ctx)); 0 on success. */
int jdhip_adjoint_csr(void *stream, int n, const int *ia_dev, const int *ja_dev,
                      const double *x_dev, long ldx, int k, const double *w, double *g_dev);
/* Gradient with respect to a diagonal potential V (A = H + diag(V)):
 * g[i] = sum_k w[k] x_k(i)^2. */
int jdhip_adjoint_diag(void *stream, int n, const double *x_dev, long ldx, int k,
                       const double *w, double *g_dev);

/* ---- diagnostics --------------------------------------------------------- */
typedef struct {
    long inner_iters;     /* inner PCG/sQMR steps of the last call */
    long host_syncs;      /* host<->device synchronisations of the last call */
    int n_precond_setups; /* preconditioner (re)factorisations */
    int spmv_lanes;       /* SpMV format chosen: k lanes per row, 0 sliced ELL */
    double setup_shifts[64];
    double precond_seconds; /* host factorisation + upload time of the last call */
    double fill;            /* stored factor entries / nnz(upper A) (ILDL) */
    int precond_levels;     /* sparse levels + dense level (ILDL) */
    int precond_negpiv;     /* negative pivots before |D| conversion */
    int precond_sweeps;     /* triangular solves: 0 exact, k Jacobi sweeps */
    /* per set-up of the last call (as printed by the reference, JDaux.F:1226-1232):
     * used DROPTOL and fill factor (nnzU + 2n) / nnz(A_upper) */
    double setup_droptols[64];
    double setup_fills[64];
} jdhip_stats;
void jdhip_get_stats(jdhip_ctx *ctx, jdhip_stats *out);
/* last error message of ctx ("" if none) */
const char *jdhip_last_error(jdhip_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
