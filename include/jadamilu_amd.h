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
/* As DPJDCLEANUP (PJD.F:435): release the preconditioner of ctx. */
void jdhip_cleanup(jdhip_ctx *ctx);

/* HIP stream (as void*) the context launches on; callers servicing IJOB
 * with their own kernels should use it (or synchronise with it). */
void *jdhip_stream(jdhip_ctx *ctx);

/* ---- symmetric SpMV (MATVECA, JDaux.F:3397-3418) ------------------------- */
/* Upload a 1-based upper-triangle CSR (diagonal anywhere in each row) as the
 * device operator of ctx; 0 on success.  lanes_per_row: 0 = autotune. */
int jdhip_set_matrix(jdhip_ctx *ctx, int n, const int *ia, const int *ja, const double *a,
                     int lanes_per_row);
/* y = A x on ctx's stream, x and y DEVICE arrays of n doubles. */
int jdhip_spmv(jdhip_ctx *ctx, const double *x_dev, double *y_dev);

/* ---- diagnostics --------------------------------------------------------- */
typedef struct {
    long inner_iters;     /* inner PCG/sQMR steps of the last call */
    long host_syncs;      /* host<->device synchronisations of the last call */
    int n_precond_setups; /* preconditioner (re)factorisations */
    int spmv_lanes;       /* lanes per row chosen for the SpMV */
    double setup_shifts[64];
} jdhip_stats;
void jdhip_get_stats(jdhip_ctx *ctx, jdhip_stats *out);
/* last error message of ctx ("" if none) */
const char *jdhip_last_error(jdhip_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
