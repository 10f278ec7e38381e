/*
 * jdo_engine.c -- CPU oracle: Jacobi-Davidson engine and revcom drivers.
 *
 * TEST INFRASTRUCTURE ONLY (see jd_oracle.h).  Plain-C restatement of
 *   JDRVCOM     src/jd/JDaux.F:1391-2415   outer JD state machine
 *   LSOLREVCOM  src/jd/JDaux.F:2989-3268   inner PCG / sQMR
 *   STOPTEST    src/jd/JDaux.F:2542-2672   inner stopping test
 *   REPGS       src/jd/JDaux.F:2674-2753   iterated classical Gram-Schmidt
 *   SHRINK      src/jd/JDaux.F:2418-2462   restart by QR of Ritz block
 *   SORTEV      src/jd/JDaux.F:2464-2540   sort + in-place column cycles
 *   SOLPROJEP   src/jd/JDaux.F:3271-3358   projected eigenproblem
 *   MATVECA     src/jd/JDaux.F:3397-3418
 *   PJDRVCOM    src/jd/JDaux.F:5-544       (diagonal + external precond)
 *   PJDINIT     src/jd/JDaux.F:549-717     (diagonal branch)
 *   DGLPRECSETUP/SOL  src/jd/dglprecsetup.c:11-82, dglprecsol.c:8-26
 * The Fortran SAVE / COMMON state (JDaux.F:30-51, 1457) lives in an explicit
 * jdo_state; the resume labels of the reference are kept as goto targets so
 * each step can be compared with the Fortran line by line.  Only the
 * standard eigenproblem (GEP=.FALSE.) is restated.  Offsets into X are
 * 0-based internally and 1-based (NDX) at the API.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "jd_oracle.h"
#include "jdo_internal.h"

#define EPSM (DBL_EPSILON * 0.5) /* dlamch('E') */
#define RINF DBL_MAX             /* dlamch('O') */
#define IMIN(a, b) ((a) < (b) ? (a) : (b))
#define IMAX(a, b) ((a) > (b) ? (a) : (b))
#define DMIN(a, b) ((a) <= (b) ? (a) : (b))
#define DMAX(a, b) ((a) >= (b) ? (a) : (b))

typedef struct {
    int rlbl;
    int ftflg, conv, uknown, initflg, neweig;
    int mvmax, midspace, modspace, nwant, itarg, iserch;
    long xx, xcg, xcg2, wcg, vv, ww, uu, reig, mm, nn, mp, np, yy, yy2;
    long wl, wl2, wl3, lwl, kxp, bx, bu, bt, bv, kbxp;
    int iterp, itcg, ifo, istep, istag;
    int i, ij, j, k, ii, jj, lijob;
    long l1, l2;
    double resid, sclr1, sclr2, theta, eta, thmt, delta1, deltao, eta0;
    double nreig, theta2, eta1, epsmsr, nreigp, trs, rndind, gr, gl, gg;
    double nrc, rinitm, alpha, gk, alprho, axpyalpha;
} jd_locals;

typedef struct {
    int rlbl, maxit, posdef;
    double beta, rho, rho1, gk2, ck2, ggk2, ttk2, ttk2m;
} lsol_locals;

typedef struct {
    int istag, istt, posdef, toupdate;
    double g0, gs, gtrs, gkmin, bk, beta, tn, un2, gkm1, alp;
} stop_locals;

struct jdo_state {
    /* COMMON /PJDINITPJD/ */
    double timefact, shift0, diagmin, shiftmax, slightlyless, toldiv;
    int factdgl, factspd, prvdr, reenterfirsttime, flagsingle;
    double *diag;
    int ndiag;
    /* COMMON /JDRVCOMJD/ */
    double timeilu, rinit, rcurr, rcum, shift1;
    int flagcg;
    /* PJDRVCOM SAVEd locals */
    int nw, ierr;
    double rprev;
    jd_locals jd;
    lsol_locals ls;
    stop_locals st;
    unsigned long long rng;
    /* diagnostics */
    double shifts[256];
    int nshifts, nprec;
    long inner;
    /* external preconditioner for jdo_pjd without a diagonal */
    jdo_psolve_fn psolve;
    void *pctx;
    double *work;
    long nwork;
};

jdo_state *jdo_create(void)
{
    jdo_state *s = (jdo_state *)calloc(1, sizeof(jdo_state));
    s->slightlyless = 0.8; /* src/blaslike/common.f:14-15 */
    s->toldiv = 2.0;
    s->rng = 0x2545F4914F6CDD1DULL;
    return s;
}

void jdo_destroy(jdo_state *s)
{
    if (!s) return;
    free(s->diag);
    free(s->work);
    free(s);
}

int jdo_nshift(const jdo_state *s) { return s->nshifts; }
double jdo_shift(const jdo_state *s, int i) { return s->shifts[i]; }
long jdo_inner_iters(const jdo_state *s) { return s->inner; }
int jdo_nprecond(const jdo_state *s) { return s->nprec; }

void jdo_set_psolve(jdo_state *s, jdo_psolve_fn fn, void *ctx)
{
    s->psolve = fn;
    s->pctx = ctx;
}

static double *scratch(jdo_state *s, long n)
{
    if (s->nwork < n) {
        free(s->work);
        s->work = (double *)malloc(sizeof(double) * n);
        s->nwork = n;
    }
    return s->work;
}

/* ---------------------------------------------------------------------- */
/* LARNV replacement for the REPGS breakdown restart (JDaux.F:2741-2745).  */
/* The reference draws from LAPACK dlaruv with seed (1,7,11,17); this path */
/* is only reached when a new direction is numerically dependent, which    */
/* the tested configurations never do.                                     */
static void larnv_u01(jdo_state *s, int n, double *w)
{
    for (int i = 0; i < n; i++) {
        s->rng ^= s->rng >> 12;
        s->rng ^= s->rng << 25;
        s->rng ^= s->rng >> 27;
        w[i] = (double)((s->rng * 2685821657736338717ULL) >> 11) * 0x1.0p-53;
    }
}

/* REPGS (JDaux.F:2674-2753).  y[0:jv) coefficients, y[jv] final norm. */
static void repgs(jdo_state *s, int n, const double *v, int jv, double *w,
                  double *y, int gam)
{
    if (jv <= 0) return;
    if (gam < 0) {
        jdo_gemv_t(n, jv, v, n, w, y);
        jdo_gemv_n(n, jv, -1.0, v, n, y, 1.0, w);
        return;
    }
    int nrnd = 1;
restart:;
    double nrn = jdo_dnrm2(n, w, 1);
    double nr = EPSM * nrn, nro;
    if (nrnd)
        for (int i = 0; i < jv; i++) y[i] = 0.0;
    for (;;) {
        nro = nrn;
        jdo_gemv_t(n, jv, v, n, w, y + jv);
        jdo_gemv_n(n, jv, -1.0, v, n, y + jv, 1.0, w);
        nrn = jdo_dnrm2(n, w, 1);
        if (nrnd)
            for (int i = 0; i < jv; i++) y[i] += y[jv + i];
        if (nrn <= nr) goto fail;
        if (!(nrn < 0.5 * nro)) break;
    }
    y[jv] = nrn;
    return;
fail:
    if (gam == 1) {
        if (jv < n) {
            larnv_u01(s, n, w);
            nrnd = 0;
            goto restart;
        }
    } else {
        for (int i = 0; i < n; i++) w[i] = 0.0;
        y[jv] = 0.0;
    }
}

/* SORTEV (JDaux.F:2464-2540), standard problem: keys by it, permute the
 * columns of v (n x j, ldv) and eigs (and r when |it|==4) in place. */
static void sortev(jdo_state *s, int n, int j, double *v, long ldv,
                   double *eigs, double *r, int it, double sigma)
{
    if (j <= 1) return;
    double *key = (double *)malloc(sizeof(double) * j);
    int *iw = (int *)malloc(sizeof(int) * j);
    double *w = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < j; i++) {
        if (it == 2 || it == -2)
            key[i] = -fabs(eigs[i] - sigma);
        else if (it == 4 || it == -4)
            key[i] = eigs[i];
        else
            key[i] = -eigs[i];
    }
    jdo_sortrx(j, key, iw);
    if (it == -2) {
        int is = iw[j - 1];
        for (int i = j - 2; i >= 0; i--) iw[i + 1] = iw[i];
        iw[0] = is;
    }
    int ic = 0;
    for (;;) {
        int is = 0;
        for (;;) {
            is++;
            if (iw[is - 1] < 0) continue;
            if (iw[is - 1] == is) {
                iw[is - 1] = -iw[is - 1];
                ic++;
                if (ic >= j) goto done;
                continue;
            }
            break;
        }
        int ir = iw[is - 1];
        memcpy(w, v + (long)(is - 1) * ldv, sizeof(double) * n);
        double r1 = 0.0, w1 = eigs[is - 1];
        if (it == 4 || it == -4) r1 = r[is - 1];
        for (;;) {
            memcpy(v + (long)(is - 1) * ldv, v + (long)(ir - 1) * ldv,
                   sizeof(double) * n);
            if (it == 4 || it == -4) r[is - 1] = r[ir - 1];
            eigs[is - 1] = eigs[ir - 1];
            iw[is - 1] = -ir;
            ic++;
            is = ir;
            ir = iw[is - 1];
            if (!(iw[ir - 1] > 0)) break;
        }
        memcpy(v + (long)(is - 1) * ldv, w, sizeof(double) * n);
        if (it == 4 || it == -4) r[is - 1] = r1;
        eigs[is - 1] = w1;
        iw[is - 1] = -ir;
        ic++;
        if (ic >= j) goto done;
    }
done:
    (void)s;
    free(key);
    free(iw);
    free(w);
}

/* SOLPROJEP (JDaux.F:3271-3358), standard problem.  w receives the
 * eigenvalues; for is==3 they are also stored in nn(:,1). */
static int solprojep(jdo_state *s, int j, double *mm, double *nn, int ld,
                     double *theta, double *theta2, double *w, int is)
{
    int info;
    if (is == 2 || is == -2) {
        info = jdo_sygv(j, mm, ld, nn, ld, w);
        if (info) return info;
        sortev(s, j, j, mm, ld, w, NULL, is, 0.0);
        *theta2 = w[1];
        return 0;
    }
    info = jdo_syev(j, mm, ld, w);
    if (info) return info;
    if (is == 3) {
        for (int i = 0; i < j; i++) nn[i] = w[i];
    } else {
        *theta = w[0];
        *theta2 = w[1];
    }
    return 0;
}

/* SHRINK (JDaux.F:2418-2462): QR of the k x j block u, V,W := [V,W] Q,
 * MM := Q^T MM Q (and NN for the harmonic case). */
static int shrink(int n, int j, int k, double *u, int ldu, double *tau,
                  double *v, double *w, double *mm, double *nn, int iserch)
{
    jdo_geqrf(k, j, u, ldu, tau);
    jdo_ormqr_rn(n, k, j, u, ldu, tau, v, n);
    jdo_ormqr_rn(n, k, j, u, ldu, tau, w, n);
    for (int ii = 1; ii < k; ii++)
        for (int jj = 0; jj < ii; jj++) mm[ii + jj * ldu] = mm[jj + ii * ldu];
    jdo_ormqr_rn(k, k, j, u, ldu, tau, mm, ldu);
    jdo_ormqr_lt(k, k, j, u, ldu, tau, mm, ldu);
    if (iserch == 2) {
        for (int ii = 1; ii < k; ii++)
            for (int jj = 0; jj < ii; jj++)
                nn[ii + jj * ldu] = nn[jj + ii * ldu];
        jdo_ormqr_rn(k, k, j, u, ldu, tau, nn, ldu);
        jdo_ormqr_lt(k, k, j, u, ldu, tau, nn, ldu);
    }
    return 0;
}

/* STOPTEST (JDaux.F:2542-2672), standard problem. */
static void stoptest(jdo_state *s, int n, const double *x, double tol,
                     double *rk, double alprho, double gk, double *theta,
                     double eta, int *info, double trs)
{
    stop_locals *t = &s->st;
    if (*info == 0) goto L10;
    t->posdef = (*info == -1);
    t->istag = 0;
    t->istt = 0;
    t->g0 = *rk;
    t->gkm1 = *rk;
    t->gtrs = trs * t->g0;
    if (t->gtrs < tol) t->gtrs = tol;
    t->gs = t->gtrs;
    t->gkmin = t->g0;
    t->bk = *theta - eta;
    t->un2 = -1.0;
L10:
    t->bk = t->bk - alprho;
    t->beta = fabs(t->bk);
    if (gk < t->gkmin) {
        t->gkmin = gk;
        t->istag = 0;
    }
    if (gk > t->gkm1) {
        t->istag++;
        t->istt++;
    }
    if (t->istag >= 25 || t->istt >= 100) goto L40;
    if (t->gkmin >= t->gtrs) goto L20;
    t->toupdate = 1;
    if (gk >= t->gs) goto L15;
L13:
    t->alp = 1.0;
    t->tn = jdo_dnrm2(n, x, 1);
    t->un2 = 1.0 + t->tn * t->tn;
    t->toupdate = 0;
    if (gk < 0.1 * t->g0)
        t->gs = tol;
    else
        t->gs = 0.1 * t->g0;
    if (t->gs < tol) t->gs = tol;
L15:
    *rk = sqrt(gk * gk * t->un2 + t->tn * t->tn * t->beta * t->beta) / t->un2;
    if (*rk < tol && t->toupdate) goto L13;
    if (*rk < tol) goto L40;
    if (gk * sqrt(t->un2) < 15 * t->tn * t->beta &&
        t->beta * t->tn > 0.5 * tol * t->un2 * t->alp)
        goto L40;
    if (t->istag >= 3) goto L40;
    if (t->istag >= 1 && t->beta * t->tn > 0.5 * tol * t->un2 * t->alp &&
        gk < t->gtrs)
        goto L40;
L20:
    *info = 0;
    t->gkm1 = gk;
    return;
L40:
    *info = IMAX(*info, 1);
    *theta = t->bk;
}

/* LSOLREVCOM (JDaux.F:2989-3268), standard problem.  work = [R Z P Q]
 * (n each).  On return *ijob: 1 matvec Q=A*P (offsets in *ndx1,*ndx2 relative
 * to work), 2 precond Z=P^-1 R, 3 stopping test, 4 finished. */
static void lsol(jdo_state *s, int n, double *x, double *work, double *x2,
                 int *iter, int *info, long *ndx1, long *ndx2, double *gk,
                 double *alpha, int *ijob)
{
    lsol_locals *l = &s->ls;
    double *R = work, *Z = work + n, *P = work + 2L * n, *Q = work + 3L * n;
    if (*ijob == 1) goto L1;
    if (*ijob == 2) {
        if (l->rlbl == 3) goto L3;
        if (l->rlbl == 4) goto L4;
        if (l->rlbl == 5) goto L5;
        *info = -6;
        goto L20;
    }
L1:
    l->posdef = (*info == 1);
    *info = 0;
    l->maxit = *iter;
    l->ggk2 = (*gk) * (*gk);
    if (l->posdef)
        for (int i = 0; i < n; i++) x[i] = 0.0;
    *iter = 0;
L10:
    (*iter)++;
    *ndx1 = n;
    *ndx2 = 0;
    l->rlbl = 3;
    *ijob = 2;
    return;
L3:
    l->rho = jdo_ddot(n, R, Z);
    if (l->posdef && l->rho <= 0.0) {
        fprintf(stderr,
                " Non fatal bug: the preconditioner is not positive definite\n");
        goto L40;
    }
    if (l->rho == 0.0) goto L40;
    if (*iter > 1) {
        l->beta = l->rho / l->rho1;
        jdo_daxpy(n, l->beta, P, Z);
        memcpy(P, Z, sizeof(double) * n);
    } else {
        memcpy(P, Z, sizeof(double) * n);
    }
    *ndx1 = 2L * n;
    *ndx2 = 3L * n;
    l->rlbl = 4;
    *ijob = 1;
    return;
L4:
    *alpha = jdo_ddot(n, P, Q);
    if ((l->posdef && *alpha <= 0.0) || *alpha == 0.0) goto L40;
    *alpha = l->rho / *alpha;
    jdo_daxpy(n, -*alpha, Q, R);
    *gk = jdo_dnrm2(n, R, 1);
    if (l->posdef) {
        jdo_daxpy(n, *alpha, P, x);
    } else {
        l->gk2 = (*gk) * (*gk);
        l->ttk2m = l->ttk2;
        l->ttk2 = l->gk2 / l->ggk2;
        l->ck2 = 1.0 / (1.0 + l->ttk2);
        l->ggk2 = l->ggk2 * l->ttk2 * l->ck2;
        if (*iter > 1) {
            jdo_daxpy(n, *alpha / l->ttk2m, P, x2);
            jdo_dscal(n, l->ttk2m * l->ck2, x2);
            jdo_daxpy(n, 1.0, x2, x);
        } else {
            memcpy(x2, P, sizeof(double) * n);
            jdo_dscal(n, *alpha * l->ck2, x2);
            memcpy(x, x2, sizeof(double) * n);
        }
    }
    l->rlbl = 5;
    *ijob = 3;
    return;
L5:
    if (*info == 1) goto L30;
    if (*iter == l->maxit) goto L40;
    l->rho1 = l->rho;
    goto L10;
L20:
    l->rlbl = -1;
    *ijob = 4;
    return;
L30:
    *info = 0;
    l->rlbl = -1;
    *ijob = 4;
    return;
L40:
    *info = 1;
    if (*alpha <= 0.0) *info = 2;
    l->rlbl = -1;
    *ijob = 4;
    if (*iter == 1) memcpy(x, P, sizeof(double) * n);
}

/* ---------------------------------------------------------------------- */
/* JDRVCOM (JDaux.F:1391-2415), standard problem.  *ijob on return:        */
/*   0 done, 1 matvec X(ndx2)=A X(ndx1), 2 precond X(ndx1)=P^-1 X(ndx2),   */
/*   3 matvec (final residual checks).  ndx are 0-based here.             */
static void jdrvcom(jdo_state *s, int n, double *eigs, double *res,
                    double *x, int lx, int *neig, double sigma, int isearch,
                    int ninit, int madspace, int *iter, double tol,
                    int *ijob, long *ndx1, long *ndx2, int *info, double *gap)
{
    jd_locals *v = &s->jd;
    if (*ijob == 0) goto L1;
    switch (v->rlbl) {
    case 1: goto L5;
    case 2: goto L225;
    case 3: goto L37;
    case 4: goto L38;
    case 5: goto L255;
    case 6: goto L55;
    default: break;
    }
    *info = -15;
    goto L190;

L1:
    s->flagcg = 1;
    v->neweig = 1;
    s->rcum = 0.0;
    v->rinitm = 1.0;
    *info = 0;
    if (n < 2)
        *info = -1;
    else if (*neig <= 0)
        *info = -5;
    else if (madspace <= 2)
        *info = -9;
    else if ((long)n * (2 * madspace + *neig + 4) + 3L * madspace * madspace +
                 IMAX(madspace * madspace, *neig) > lx)
        *info = -10;
    else if (*iter <= 0)
        *info = -11;
    else if (tol <= 0.0)
        *info = -12;
    if (*info != 0) goto L190;
    v->modspace = IMIN(madspace, n);
    v->nwant = IMIN(*neig, n);
    v->mvmax = *iter;
    /* JDaux.F:1530-1531: the second assignment overrides the first. */
    v->iserch = IMIN(isearch, 2);
    {
        long N = n, MS = v->modspace;
        v->xx = 0;
        v->vv = v->xx + (long)v->nwant * N;
        v->ww = v->vv + MS * N;
        v->bx = v->xx;
        v->bv = v->vv;
        v->mm = v->ww + MS * N;
        v->nn = v->mm + MS * MS;
        v->mp = v->nn + MS * MS;
        v->np = v->mp + MS * MS;
        v->wcg = v->np + (MS * MS > v->nwant ? MS * MS : v->nwant);
        v->xcg2 = v->ww + (MS - 1) * N;
        v->bt = v->bv + (MS - 1) * N;
        v->reig = v->wcg;
        v->yy = v->mp;
        v->yy2 = v->mp + MS;
        v->wl = v->reig + N;
        v->lwl = lx - v->wl;
        v->wl2 = v->wl + MS;
        v->wl3 = v->wl2 + MS;
        v->uu = v->xx + (long)(v->nwant - 1) * N;
        v->bu = v->bx + (long)(v->nwant - 1) * N;
        v->kxp = v->vv;
        v->kbxp = v->bv;
    }
    v->epsmsr = sqrt(EPSM);
    *neig = 0;
    *iter = 0;
    v->iterp = 0;
    v->istep = 0;
    v->eta1 = RINF;
    v->eta0 = sigma;
    v->delta1 = RINF;
    if (v->iserch == 0) v->eta0 = RINF;
    v->nreig = RINF;
    v->resid = RINF;
    v->rndind = 0.0;
    v->midspace = (v->modspace + 1) / 2;
    if (ninit > v->midspace) {
        v->initflg = 0;
        for (int q = 1; q < ninit; q++)
            if (eigs[q] != eigs[0]) v->initflg = 1;
        if (v->initflg) {
            sortev(s, n, ninit, x + v->xx, n, eigs, res, v->iserch, sigma);
        } else {
            for (int q = 1; q <= ninit / 2; q++) {
                int r = ninit - q + 1;
                memcpy(x + v->vv, x + (long)(r - 1) * n, sizeof(double) * n);
                memcpy(x + (long)(r - 1) * n, x + (long)(q - 1) * n,
                       sizeof(double) * n);
                memcpy(x + (long)(q - 1) * n, x + v->vv, sizeof(double) * n);
            }
        }
    }
    v->initflg = 0;

L2:
    if (ninit <= *neig) {
        v->j = 1;
        for (int q = 0; q < n; q++) x[v->vv + q] = 1.0;
    } else {
        v->j = IMIN(ninit - *neig, v->midspace);
        v->ij = ninit - *neig - v->j;
        if (v->ij > 0) v->initflg = 1;
        memmove(x + v->vv, x + (long)n * v->ij, sizeof(double) * n * v->j);
    }
    v->ij = 0;
L4:
    v->ij++;
    repgs(s, n, x + v->kxp, *neig + v->ij - 1, x + v->vv + (long)(v->ij - 1) * n,
          x + v->wl, 1);
    if (*neig + v->ij - 1 > 0)
        v->sclr1 = x[v->wl + *neig + v->ij - 1];
    else
        v->sclr1 = jdo_dnrm2(n, x + v->vv, 1);
    jdo_dscal(n, 1.0 / v->sclr1, x + v->vv + (long)(v->ij - 1) * n);
    *ndx1 = v->vv + (long)(v->ij - 1) * n;
    *ndx2 = v->ww + (long)(v->ij - 1) * n;
    *ijob = 1;
    v->rlbl = 1;
    return;
L5:
    (*iter)++;
    if (v->ij < v->j) goto L4;
    v->axpyalpha = -sigma;
    if (v->iserch == 2) {
        jdo_daxpy(n * v->j, v->axpyalpha, x + v->bv, x + v->ww);
        jdo_syrk_ut(v->j, n, x + v->ww, n, x + v->nn, v->modspace);
    }
    jdo_syr2k_ut(v->j, n, x + v->vv, n, x + v->ww, n, x + v->mm, v->modspace);
    jdo_dscal(v->modspace * v->modspace, 0.5, x + v->mm);
    v->uknown = (v->j <= 1);

L10: /* extraction */
    memcpy(x + v->mp, x + v->mm, sizeof(double) * v->modspace * v->modspace);
    if (v->iserch == 2)
        memcpy(x + v->np, x + v->nn,
               sizeof(double) * v->modspace * v->modspace);
    *info = solprojep(s, v->j, x + v->mp, x + v->np, v->modspace, &v->theta,
                      &v->theta2, x + v->reig, v->iserch);
    if (*info != 0) {
        if (v->iserch == 2) goto L191;
        goto L192;
    }
    if (!v->uknown) {
        if (v->iserch == 2) {
            v->sclr1 = jdo_dnrm2(v->j, x + v->yy, 1);
            jdo_dscal(v->j, 1.0 / v->sclr1, x + v->yy);
        }
        jdo_gemv_n(n, v->j, 1.0, x + v->vv, n, x + v->yy, 0.0, x + v->uu);
        jdo_gemv_n(n, v->j, 1.0, x + v->ww, n, x + v->yy, 0.0, x + v->reig);
    } else {
        x[v->yy] = 1.0;
        for (int q = 1; q < v->j; q++) x[v->yy + q] = 0.0;
        memcpy(x + v->uu, x + v->vv, sizeof(double) * n);
        memcpy(x + v->reig, x + v->ww, sizeof(double) * n);
    }
    if (v->iserch == 2) {
        jdo_symv_u(v->j, x + v->mm, v->modspace, x + v->yy, x + v->wl);
        v->thmt = jdo_ddot(v->j, x + v->wl, x + v->yy);
        v->theta = v->thmt + sigma;
    } else {
        v->thmt = v->theta;
        v->eta0 = DMIN(v->eta0, v->theta);
    }
    v->axpyalpha = -v->thmt;
    jdo_daxpy(n, v->axpyalpha, x + v->bu, x + v->reig);
    v->nreigp = v->nreig;
    v->nreig = jdo_dnrm2(n, x + v->reig, 1);
    if (v->neweig) {
        s->rinit = v->nreig;
        v->neweig = 0;
    }
    if (fabs(v->nreig / v->nreigp - 1.0) < 0.01)
        v->istag++;
    else
        v->istag = 0;
    if (v->nreig >= tol && v->istag <= 2) goto L20;
    if (v->nreig >= tol && v->rndind <= 1.0) goto L20;
    if (v->nreig >= tol && v->resid > 0.1 * v->nreig) goto L20;
    /* convergence detected */
    repgs(s, n, x + v->kxp, *neig, x + v->uu, x + v->wl, 0);
    if (*neig > 0 && x[v->wl + *neig] == 0.0) goto L195;
    if (*neig > 0)
        v->sclr1 = x[v->wl + *neig];
    else
        v->sclr1 = jdo_dnrm2(n, x + v->uu, 1);
    jdo_dscal(n, 1.0 / v->sclr1, x + v->uu);

L18: /* lock the converged pair */
    (*neig)++;
    v->kxp = v->uu;
    v->kbxp = v->bu;
    v->uu -= n;
    v->bu -= n;
    eigs[v->nwant - *neig] = v->theta;
    res[v->nwant - *neig] = v->nreig;
    v->rinitm *= s->rinit;
    v->iterp = *iter;
    *gap = v->theta2 - v->theta;
    if (*neig >= v->nwant) goto L200;
    v->delta1 = RINF;
    v->nreig = RINF;
    v->resid = RINF;
    v->rndind = 0.0;
    if (v->iserch != 2) {
        v->eta1 = DMIN(v->eta0, v->eta1);
        /* JDaux.F:1794-1796 reads EIGS(1..NEIG) although the converged
         * values sit at EIGS(NWANT-NEIG+1..NWANT). */
        for (int q = 0; q < *neig; q++) v->eta0 = DMAX(v->eta0, eigs[q]);
    }
    if (v->j == 1 || v->initflg) goto L2;
    v->k = v->j;
    v->j = IMIN(v->midspace + 1, v->j - 1);
    v->j = IMIN(v->j, v->modspace - 1);
    if (v->iserch == 2) {
        for (v->ij = 2; v->ij <= v->j + 1; v->ij++)
            repgs(s, v->k, x + v->yy, 1,
                  x + v->mp + (long)v->modspace * (v->ij - 1), x + v->wl, 1);
    }
    *info = shrink(n, v->j, v->k, x + v->yy2, v->modspace, x + v->np,
                   x + v->vv, x + v->ww, x + v->mm, x + v->nn, v->iserch);
    if (*info != 0) goto L193;
    v->uknown = 1;
    v->neweig = 1;
    goto L10;

L20: /* expansion */
    v->trs = sqrt(1.0 / 10);
    if (v->istag > 0) v->trs = 1.0 / 10;
    v->istep++;
    if (*iter >= v->mvmax - *neig - 1) goto L197;
    if (v->j >= v->modspace) {
        v->k = v->j;
        v->j = v->midspace;
        *info = shrink(n, v->j, v->k, x + v->mp, v->modspace, x + v->np,
                       x + v->vv, x + v->ww, x + v->mm, x + v->nn, v->iserch);
        if (*info != 0) goto L193;
        v->sclr1 = 0.0;
        for (int q = 0; q < n; q++)
            if (fabs(x[v->uu + q]) > v->sclr1) {
                v->sclr1 = fabs(x[v->uu + q]);
                v->jj = q;
            }
        v->axpyalpha = x[v->uu + v->jj] / x[v->vv + v->jj];
        x[v->yy] = v->axpyalpha / fabs(v->axpyalpha);
        for (int q = 1; q < v->j; q++) x[v->yy + q] = 0.0;
    }
    v->eta = v->eta0;
    if (v->j > 1) {
        v->deltao = v->delta1;
        if (v->iserch == 2) {
            v->sclr1 = jdo_dnrm2(v->j, x + v->yy2, 1);
            jdo_dscal(v->j, 1.0 / v->sclr1, x + v->yy2);
            jdo_symv_u(v->j, x + v->mm, v->modspace, x + v->yy2, x + v->wl);
            v->theta2 = jdo_ddot(v->j, x + v->yy2, x + v->wl);
            v->delta1 = fabs(v->theta2) - fabs(v->theta - sigma);
        } else {
            v->delta1 = v->theta2 - v->theta;
        }
        if (v->nreig < fabs(v->delta1) &&
            fabs(v->delta1 / v->deltao - 1.0) < 0.1)
            v->eta = v->theta;
    }
    repgs(s, n, x + v->bu, *neig + 1, x + v->reig, x + v->wl, -1);
    v->rndind = fabs(x[v->wl + *neig] / (v->nreig * v->epsmsr));
    v->resid = v->nreig;
    v->xcg = v->vv + (long)v->j * n;
    v->ftflg = 1;
    v->itcg = v->mvmax - *iter - *neig - 1;
    v->lijob = 1;
    v->ifo = 0;
    if (v->iserch != 2) v->ifo = 1;
    v->gk = v->resid;
    s->rcurr = v->nreig;
    if (*neig <= 0) {
        s->shift1 = v->theta;
        if (v->j > 1)
            *gap = v->theta2 - v->theta;
        else
            *gap = 0.0;
    }
L30:
    lsol(s, n, x + v->xcg, x + v->wcg, x + v->xcg2, &v->itcg, &v->ifo, &v->l1,
         &v->l2, &v->gk, &v->alpha, &v->lijob);
    switch (v->lijob) {
    case 1: goto L31;
    case 2: goto L32;
    case 3: goto L33;
    default: goto L34;
    }
L31:
    v->rlbl = 3;
    *ndx1 = v->l1 + v->wcg;
    *ndx2 = v->l2 + v->wcg;
    *ijob = 1;
    return;
L37:
    v->lijob = 2;
    v->axpyalpha = -v->eta;
    jdo_daxpy(n, v->axpyalpha, x + *ndx1, x + *ndx2);
    repgs(s, n, x + v->bu, 1, x + *ndx2, x + v->np, -1);
    v->alprho = x[v->np];
    (*iter)++;
    s->inner++;
    goto L30;
L32:
    v->rlbl = 4;
    *ndx1 = v->l1 + v->wcg;
    *ndx2 = v->l2 + v->wcg;
    *ijob = 2;
    return;
L38:
    v->lijob = 2;
    s->flagcg = 1;
    repgs(s, n, x + v->uu, *neig + 1, x + *ndx1, x + v->np, -1);
    goto L30;
L33:
    if (v->ftflg) v->ifo = -1;
    if (v->ftflg && v->iserch == 2) v->ifo = -2;
    v->axpyalpha = v->axpyalpha * v->alpha;
    v->alprho = v->alprho * v->alpha;
    stoptest(s, n, x + v->xcg, tol, &v->resid, v->alprho, v->gk, &v->theta,
             v->eta, &v->ifo, v->trs);
    v->ftflg = 0;
    v->lijob = 2;
    goto L30;
L34:
    if (v->iserch != 2 && v->eta <= v->eta0 && *neig > 0 && v->ifo == 2)
        v->eta0 = DMIN(v->eta1, v->eta0);
    v->conv = 0;
    s->flagcg = 0;
    s->rcum = s->rcum + log(v->nreig / v->resid);
    if (v->resid >= tol) goto L50;
    /* convergence reached during the inner solve */
    jdo_daxpy(n, -1.0, x + v->xcg, x + v->uu);
    repgs(s, n, x + v->kxp, *neig, x + v->uu, x + v->np, 0);
    if (*neig > 0 && x[v->np + *neig] == 0.0) goto L195;
    if (*neig > 0)
        v->sclr1 = x[v->np + *neig];
    else
        v->sclr1 = jdo_dnrm2(n, x + v->uu, 1);
    jdo_dscal(n, 1.0 / v->sclr1, x + v->uu);
    v->theta = v->theta / v->sclr1 + v->eta;
    v->conv = 1;
    v->nreig = v->resid;
    if (*neig >= v->nwant - 1) goto L18;
L50: /* expand the subspace with the correction */
    repgs(s, n, x + v->kxp, *neig + v->j, x + v->xcg, x + v->reig, 1);
    jdo_daxpy(v->j, -1.0, x + v->reig + *neig, x + v->yy);
    v->sclr1 = x[v->reig + *neig + v->j];
    jdo_dscal(n, 1.0 / v->sclr1, x + v->xcg);
    *ndx1 = v->xcg;
    *ndx2 = v->ww + (long)v->j * n;
    *ijob = 1;
    v->rlbl = 6;
    return;
L55:
    (*iter)++;
    x[v->yy + v->j] = -v->sclr1;
    v->axpyalpha = -sigma;
    if (v->iserch == 2)
        jdo_daxpy(n, v->axpyalpha, x + v->bv + (long)v->j * n, x + *ndx2);
    jdo_gemv_t(n, v->j + 1, x + v->vv, n, x + *ndx2,
               x + v->mm + (long)v->j * v->modspace);
    if (v->iserch == 2)
        jdo_gemv_t(n, v->j + 1, x + v->ww, n, x + *ndx2,
                   x + v->nn + (long)v->j * v->modspace);
    v->j++;
    v->uknown = 0;
    if (!v->conv) goto L10;
    /* convergence detected: prepare for restart */
    v->sclr1 = jdo_dnrm2(v->j, x + v->yy, 1);
    jdo_dscal(v->j, 1.0 / v->sclr1, x + v->yy);
    if (v->iserch == 2) {
        jdo_symv_u(v->j, x + v->nn, v->modspace, x + v->yy, x + v->wl2);
        v->sclr1 = jdo_ddot(v->j, x + v->yy, x + v->wl2);
        jdo_symv_u(v->j, x + v->mm, v->modspace, x + v->yy, x + v->wl);
        v->thmt = jdo_ddot(v->j, x + v->yy, x + v->wl);
        v->axpyalpha = -v->thmt / (2 * v->sclr1);
        jdo_daxpy(v->j, v->axpyalpha, x + v->wl2, x + v->wl);
        v->axpyalpha = -1.0 / v->sclr1;
        memcpy(x + v->mp, x + v->mm,
               sizeof(double) * v->modspace * v->modspace);
        jdo_syr2_u(v->j, v->axpyalpha, x + v->wl, x + v->wl2, x + v->mp,
                   v->modspace);
        v->theta = v->thmt + sigma;
        memcpy(x + v->np, x + v->nn,
               sizeof(double) * v->modspace * v->modspace);
    } else {
        jdo_symv_u(v->j, x + v->mm, v->modspace, x + v->yy, x + v->wl);
        v->theta = jdo_ddot(v->j, x + v->yy, x + v->wl);
        v->eta0 = DMIN(v->eta0, v->theta);
        v->sclr1 = jdo_dnrm2(v->j, x + v->mp, v->modspace + 1);
        if (v->theta < 0) v->sclr1 = v->sclr1 - 2 * v->theta;
        v->axpyalpha = v->sclr1;
        jdo_daxpy(v->j, v->axpyalpha, x + v->yy, x + v->wl);
        memcpy(x + v->wl2, x + v->yy, sizeof(double) * v->j);
        memcpy(x + v->mp, x + v->mm,
               sizeof(double) * v->modspace * v->modspace);
        jdo_syr2_u(v->j, -1.0, x + v->wl2, x + v->wl, x + v->mp, v->modspace);
    }
    *info = solprojep(s, v->j, x + v->mp, x + v->np, v->modspace, &v->sclr1,
                      &v->sclr2, x + v->reig, -v->iserch);
    if (*info != 0) {
        if (v->iserch == 2) goto L191;
        goto L192;
    }
    if (v->iserch == 2) {
        v->sclr1 = jdo_dnrm2(v->j, x + v->yy, 1);
        jdo_dscal(v->j, 1.0 / v->sclr1, x + v->yy);
    }
    goto L18;

    /* error exits (JDaux.F:2090-2160) */
L190:
    goto L301;
L191:
L192:
    if (*info > 0) *info = -*info;
    goto L198;
L193:
    if (*info > 0) *info = -*info;
    goto L198;
L195:
    *info = -54320;
    goto L198;
L197:
    *info = *neig + 1;
    v->rinitm *= s->rinit;
L198:
    if (*neig > 0) memcpy(x + v->vv, x + v->uu, sizeof(double) * n);
L200: /* normal termination */
    if (*neig < v->nwant) {
        s->rinit = pow(v->rinitm, 1.0 / (double)(*neig + 1));
        if (*neig > 0) {
            v->jj = IMIN(*neig, v->nwant - *neig);
            memmove(x + v->xx, x + (long)n * (v->nwant - v->jj) + v->xx,
                    sizeof(double) * n * v->jj);
            memmove(eigs, eigs + v->nwant - v->jj, sizeof(double) * v->jj);
            memmove(res, res + v->nwant - v->jj, sizeof(double) * v->jj);
            memcpy(x + (long)n * (*neig) + v->xx, x + v->vv, sizeof(double) * n);
            eigs[*neig] = v->theta;
            res[*neig] = v->nreig;
            v->theta2 = v->theta;
        } else {
            memcpy(x + v->xx, x + v->uu, sizeof(double) * n);
            eigs[0] = v->theta;
            res[0] = v->nreig;
            *gap = 0.0;
            goto L300;
        }
    } else {
        s->rinit = pow(v->rinitm, 1.0 / (double)(*neig));
    }
    sortev(s, n, *neig, x + v->xx, n, eigs, res, 4, 0.0);
    if (v->iserch >= 2) {
        for (v->itarg = 1; v->itarg <= *neig; v->itarg++)
            if (eigs[v->itarg - 1] > sigma) break;
    } else {
        v->itarg = 0;
    }
    if (v->iserch == 2)
        *gap = fabs(v->theta2 - sigma) -
               DMAX(fabs(eigs[*neig - 1] - sigma), fabs(eigs[0] - sigma));
    else
        *gap = v->theta2 - eigs[*neig - 1];
    if (v->iserch == 2)
        v->gg = eigs[0] + eigs[*neig - 1] - 2 * sigma;
    else
        v->gg = RINF;
    if (v->gg > 0.0) {
        v->ii = 1;
        v->jj = *neig - 1;
    } else {
        v->ii = 2;
        v->jj = *neig;
    }
    if (*neig == 1) {
        v->ii = 0;
        v->jj = 0;
    }
    v->i = 1;
    v->k = 0;
L210:
    if (v->i < v->ii || v->i >= v->jj) goto L250;
    if (v->i == 1)
        v->gl = v->gg - res[*neig - 1];
    else
        v->gl = eigs[v->i - 1] - eigs[v->i - 2] - res[v->i - 2];
    if (v->gl < 10 * (eigs[v->i] - eigs[v->i - 1])) goto L250;
    v->j = IMIN(v->i + v->modspace, v->jj + 1);
L215:
    v->j--;
    if (v->j == v->i) goto L250;
    if (v->j == *neig)
        v->gr = -v->gg - res[0];
    else
        v->gr = eigs[v->j] - eigs[v->j - 1] - res[v->j];
    if (10 * (eigs[v->j - 1] - eigs[v->i - 1]) > DMIN(v->gl, v->gr)) goto L215;
    /* cluster: additional Rayleigh-Ritz */
    v->k++;
    v->j = v->j - v->i + 1;
    memcpy(x + v->vv, x + (long)n * (v->i - 1) + v->xx,
           sizeof(double) * n * v->j);
    v->ij = 0;
L220:
    v->ij++;
    *ndx1 = v->vv + (long)(v->ij - 1) * n;
    *ndx2 = v->ww + (long)(v->ij - 1) * n;
    *ijob = 3;
    v->rlbl = 2;
    return;
L225:
    (*iter)++;
    if (v->ij < v->j) goto L220;
    jdo_syr2k_ut(v->j, n, x + v->vv, n, x + v->ww, n, x + v->mm, v->modspace);
    jdo_dscal(v->modspace * v->modspace, 0.5, x + v->mm);
    v->ij = solprojep(s, v->j, x + v->mm, x + v->np, v->modspace, &v->sclr1,
                      &v->sclr2, x + v->reig, 3);
    if (v->ij != 0) goto L250;
    {
        /* GEMM N x J x J twice (JDaux.F:2276-2279) */
        int jc = v->j;
        double *tmp = scratch(s, (long)n * jc);
        for (int c = 0; c < jc; c++)
            jdo_gemv_n(n, jc, 1.0, x + v->vv, n, x + v->mm + (long)c * v->modspace,
                       0.0, tmp + (long)c * n);
        memcpy(x + (long)n * (v->i - 1) + v->xx, tmp, sizeof(double) * n * jc);
        for (int c = 0; c < jc; c++)
            jdo_gemv_n(n, jc, 1.0, x + v->ww, n, x + v->mm + (long)c * v->modspace,
                       0.0, tmp + (long)c * n);
        memcpy(x + v->vv, tmp, sizeof(double) * n * jc);
    }
    v->nrc = 0.0;
    v->gl = v->gl - eigs[v->i - 1];
    v->gr = v->gr + eigs[v->i + v->j - 2];
    for (v->ij = 1; v->ij <= v->j; v->ij++) {
        v->theta = x[v->np + v->ij - 1];
        eigs[v->i + v->ij - 2] = v->theta;
        jdo_daxpy(n, -v->theta, x + (long)n * (v->i + v->ij - 2) + v->bx,
                  x + v->vv + (long)n * (v->ij - 1));
        v->nreig = jdo_dnrm2(n, x + v->vv + (long)n * (v->ij - 1), 1);
        res[v->i + v->ij - 2] = v->nreig;
        v->nrc += v->nreig * v->nreig;
    }
    for (v->ij = v->i; v->ij <= v->i + v->j - 1; v->ij++)
        if (v->ij == v->itarg) v->itarg--;
    v->i = v->i + v->j;
    if (v->i > *neig) goto L300;
    goto L210;
L250: /* isolated eigenvalue: true residual */
    *ndx1 = (long)n * (v->i - 1);
    *ndx2 = v->vv;
    *ijob = 3;
    v->rlbl = 5;
    return;
L255:
    (*iter)++;
    v->theta = jdo_ddot(n, x + (long)n * (v->i - 1), x + v->vv);
    jdo_daxpy(n, -v->theta, x + (long)n * (v->i - 1) + v->bx, x + v->vv);
    v->nreig = jdo_dnrm2(n, x + v->vv, 1);
    eigs[v->i - 1] = v->theta;
    res[v->i - 1] = v->nreig;
    if (v->i == v->itarg) v->itarg--;
    v->i++;
    if (v->i <= *neig) goto L210;
L300:
L301:
    *ijob = 0;
}

/* ---------------------------------------------------------------------- */
/* Preconditioner management (PJDRVCOM / PJDINIT, diagonal branch).        */

/* DGLPRECSETUP (dglprecsetup.c:11-82) */
static void dglprecsetup(jdo_state *s, const double *a, int n, int isearch,
                         double shift, double *shift0, double *diagmin)
{
    free(s->diag);
    s->diag = (double *)malloc(sizeof(double) * n);
    s->ndiag = n;
    double ta, del, ati, ata;
    *diagmin = a[0];
    ta = *diagmin;
    for (int i = 1; i < n; i++) {
        *diagmin = DMIN(*diagmin, a[i]);
        ta = DMAX(ta, a[i]);
    }
    ati = DMAX(*diagmin, -*diagmin);
    ata = DMAX(ta, -ta);
    ata = DMIN(ata, ati);
    del = DMAX(ta - *diagmin, ata);
    if (isearch <= 1) {
        *shift0 = *diagmin - del / n;
        s->factspd = -1;
        if (isearch == 1 && shift < *diagmin) {
            *shift0 = shift;
            s->factspd = 0;
        }
        for (int i = 0; i < n; i++) s->diag[i] = 1.0 / (a[i] - *shift0);
    } else {
        s->factspd = 0;
        del = 1. / del;
        for (int i = 0; i < n; i++) {
            double t = a[i] - *shift0;
            if (t > 0.0)
                t = DMAX(t, del);
            else
                t = DMIN(t, -del);
            s->diag[i] = 1.0 / t;
        }
    }
    s->factdgl = 1;
}

/* PJDINIT, diagonal branch (JDaux.F:693-717).  The COMMON diagmin is not
 * updated by this branch (the routine's result goes to a local copy). */
static void pjdinit_diag(jdo_state *s, int n, const double *a, double sigma,
                         int isearch, double shift)
{
    if (s->factdgl != 0) s->factdgl = 0; /* PJDCLEANUP1 */
    if (isearch >= 2) s->shift0 = sigma;
    double rdiagmin = s->diagmin, sshift0 = s->shift0;
    dglprecsetup(s, a, n, isearch, shift, &sshift0, &rdiagmin);
    s->shift0 = sshift0;
    if (s->nshifts < 256) s->shifts[s->nshifts++] = s->shift0;
}

static void pjdrvcom(jdo_state *s, int n, double *a, int *ja, int *ia,
                     double *eigs, double *res, double *x, int lx, int *neig,
                     double sigma, int isearch, int ninit, int madspace,
                     int *iter, double tol, double *shift, double *droptol,
                     double *mem, int *icntl, int *ijob, long *ndx1,
                     long *ndx2, int *info, double *gap, int internal_mv)
{
    (void)droptol;
    (void)mem;
    if (*ijob == 0) {
        s->flagsingle = icntl[4] != 0;
        s->ierr = 0;
        if ((icntl[1] != -1 && icntl[1] != -2) || s->factdgl == 0) {
            s->prvdr = 0;
            s->timefact = -1.0;
            s->factdgl = 0;
            if (ja[0] < 0) {
                pjdinit_diag(s, n, a, sigma, isearch, *shift);
            } else if (s->psolve) {
                s->factdgl = -1;
            } else {
                s->ierr = 1;
            }
            s->reenterfirsttime = 0;
        } else {
            s->prvdr = 0;
            s->reenterfirsttime = 1;
            s->rinit = s->rcurr = s->rcum = 0.0;
            if (ja[0] < 0 && isearch >= 2) s->shift0 = sigma;
        }
        if (s->ierr != 0) {
            *info = -54321;
            *ijob = 0;
            return;
        }
        s->rprev = 1.0;
        s->nw = *neig;
    }
    for (;;) {
        jdrvcom(s, n, eigs, res, x, lx, neig, sigma, isearch, ninit, madspace,
                iter, tol, ijob, ndx1, ndx2, info, gap);
        if (*ijob == 0) break;
        if (*ijob == 2) {
            int reinit = 0, factspd0 = 0;
            double tent;
            if (s->factdgl > 0 && !s->flagcg && isearch <= 1 &&
                (icntl[1] == 0 || icntl[1] == -2) && icntl[0] != 2) {
                if (*neig >= 1) {
                    tent = eigs[s->nw - 1];
                    factspd0 = -1;
                } else {
                    tent = s->shift1 - s->rcurr;
                    factspd0 = -1;
                    if (10 * s->rcurr >= s->diagmin - tent &&
                        s->shift1 > s->shift0)
                        factspd0 = s->factspd;
                }
                if (s->diagmin > tent && factspd0 < 0 &&
                    10 * fabs(s->shift0 - tent) >= (s->diagmin - s->shift0)) {
                    *shift = tent;
                    reinit = 1;
                }
            }
            if (reinit) {
                pjdinit_diag(s, n, a, sigma, IMAX(isearch, 1), *shift);
                if (factspd0 < 0 && s->factspd <= 0) s->factspd = factspd0;
            }
            s->nprec++;
            if (s->factdgl > 0) {
                for (int i = 0; i < n; i++)
                    x[*ndx1 + i] = s->diag[i] * x[*ndx2 + i];
            } else {
                s->psolve(s->pctx, n, x + *ndx1, x + *ndx2);
            }
            continue;
        }
        if (!internal_mv) return;
        jdo_matveca(n, ia, ja, a, x + *ndx1, x + *ndx2);
    }
    /* JDaux.F:476-482: suggested SHIFT */
    *shift = s->shift0;
    if (s->factdgl >= 0) *shift = eigs[0];
}

void jdo_matveca(int n, const int *ia, const int *ja, const double *a,
                 const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int i = 0; i < n; i++) {
        int k = ia[i] - 1;
        y[i] += a[k] * x[i];
        for (k = ia[i]; k < ia[i + 1] - 1; k++) {
            int c = ja[k] - 1;
            y[i] += a[k] * x[c];
            y[c] += a[k] * x[i];
        }
    }
}

void jdo_matvec_any(int n, const int *ia, const int *ja, const double *a,
                    const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int i = 0; i < n; i++)
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; k++) {
            int c = ja[k] - 1;
            if (c == i) {
                y[i] += a[k] * x[i];
            } else {
                y[i] += a[k] * x[c];
                y[c] += a[k] * x[i];
            }
        }
}

/* DPJDREVCOM (PJD.F:268-362) */
void jdo_pjdrevcom(jdo_state *s, int n, double *a, int *ja, int *ia,
                   double *eigs, double *res, double *x, int lx, int *neig,
                   double sigma, int isearch, int ninit, int madspace,
                   int *iter, double tol, double *shift, double *droptol,
                   double *mem, int *icntl, int *ijob, int *ndx1, int *ndx2,
                   int iprint, int *info, double *gap)
{
    (void)iprint;
    long l1 = *ndx1 - 1, l2 = *ndx2 - 1;
    if (*ijob == 0) s->inner = 0, s->nprec = 0;
    pjdrvcom(s, n, a, ja, ia, eigs, res, x, lx, neig, sigma, isearch, ninit,
             madspace, iter, tol, shift, droptol, mem, icntl, ijob, &l1, &l2,
             info, gap, icntl[0] == 1);
    if (*ijob == 3) *ijob = 1;
    *ndx1 = (int)(l1 + 1);
    *ndx2 = (int)(l2 + 1);
}

/* DJDREVCOM (PJD.F:369-427) */
void jdo_jdrevcom(jdo_state *s, int n, double *eigs, double *res, double *x,
                  int lx, int *neig, double sigma, int isearch, int ninit,
                  int madspace, int *iter, double tol, int *ijob, int *ndx1,
                  int *ndx2, int iprint, int *info, double *gap)
{
    (void)iprint;
    long l1 = *ndx1 - 1, l2 = *ndx2 - 1;
    if (*ijob == 0) s->inner = 0, s->nprec = 0;
    jdrvcom(s, n, eigs, res, x, lx, neig, sigma, isearch, ninit, madspace,
            iter, tol, ijob, &l1, &l2, info, gap);
    if (*ijob == 3) *ijob = 1;
    *ndx1 = (int)(l1 + 1);
    *ndx2 = (int)(l2 + 1);
}

/* DPJD (PJD.F:5-259): internal matvec; the preconditioner is the one
 * installed with jdo_set_psolve (ja[0] >= 0). */
void jdo_pjd(jdo_state *s, int n, double *a, int *ja, int *ia, double *eigs,
             double *res, double *x, int lx, int *neig, double sigma,
             int isearch, int ninit, int madspace, int *iter, double tol,
             double *shift, double *droptol, double *mem, int *icntl,
             int iprint, int *info, double *gap)
{
    (void)iprint;
    int keep = icntl[0];
    icntl[0] = 1;
    int ijob = 0;
    long l1 = 0, l2 = 0;
    s->inner = 0;
    s->nprec = 0;
    /* JDaux.F:215-229: move the diagonal to the first slot of every row */
    if (ja[0] >= 0) {
        for (int i = 0; i < n; i++)
            for (int k = ia[i] - 1; k < ia[i + 1] - 1; k++)
                if (ja[k] == i + 1) {
                    int l = ia[i] - 1, tj = ja[k];
                    double ta = a[k];
                    ja[k] = ja[l];
                    ja[l] = tj;
                    a[k] = a[l];
                    a[l] = ta;
                }
    }
    pjdrvcom(s, n, a, ja, ia, eigs, res, x, lx, neig, sigma, isearch, ninit,
             madspace, iter, tol, shift, droptol, mem, icntl, &ijob, &l1, &l2,
             info, gap, 1);
    icntl[0] = keep;
}

void jdo_cleanup(jdo_state *s)
{
    free(s->diag);
    s->diag = NULL;
    s->factdgl = 0;
}
