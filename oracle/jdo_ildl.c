/*
 * jdo_ildl.c -- CPU oracle of the multilevel ILDL apply.
 *
 * TEST INFRASTRUCTURE ONLY.  Sequential restatement of the multilevel
 * forward/backward substitution of DSYMAMGsol (src/ilupack/symAMGsol.c:
 * 110-631) for 1x1 pivots.  Per level: scale+permute; with the coupling
 * block F^T = A_CB (the reference's structure) y_B = B^-1 b_B (SYMPILUCSOL,
 * sympilucsol.F:104-201: forward sweep with L_B, D^-1, backward sweep with
 * L_B^T), coarse right-hand side b_C - F^T y_B (CSRmattvec.c), recursion on
 * the Schur complement, then x_B = y_B - B^-1 F x_C (CSRmatvec.c and a
 * second SYMPILUCSOL, symAMGsol.c:540-600); with L_CB instead the block
 * LDL^T substitution (forward sweep, coarse rhs with L_CB, D^-1, recursion,
 * L_CB^T product, backward sweep); unpermute + scale.  The final level is an
 * explicit dense inverse.  It consumes the factor exported through the
 * product's jdhip_ildl_level_data().
 */
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n, nB;
    const int *perm;
    const double *scal;
    const int *lb_ptr, *lb_idx;
    const double *lb_val;
    const int *lc_ptr, *lc_idx;
    const double *lc_val;
    const double *d;
    const double *doff; /* 2x2 pivots: off-diagonal entry of b's block */
    const int *dpair;   /* the other row of b's block, -1 for 1x1 (NULL: all 1x1) */
} jdo_ildl_level;

/* x_B = D^-1 x_B, D block diagonal with 1x1 and 2x2 blocks (symAMGsol.c:
 * 221-300) */
static void dsolve(const jdo_ildl_level *V, double *x)
{
    for (int b = 0; b < V->nB; b++) {
        const int p = V->dpair ? V->dpair[b] : -1;
        if (p < 0) {
            x[b] /= V->d[b];
        } else if (b < p) {
            const double a = V->d[b], o = V->doff[b], c = V->d[p];
            const double det = a * c - o * o;
            const double xb = x[b], xp = x[p];
            x[b] = (c * xb - o * xp) / det;
            x[p] = (a * xp - o * xb) / det;
        }
    }
}

/* 0: exact triangular solves; k > 0: each triangular solve replaced by k
 * Jacobi sweeps x <- b - N x from x = b (N the strictly triangular part),
 * the device's sweep mode. */
static int g_sweeps = 0;

void jdo_ildl_set_sweeps(int k) { g_sweeps = k > 0 ? k : 0; }

static void jacobi_lower(int nB, const int *ptr, const int *idx, const double *val, double *b,
                         int transposed)
{
    double *x = (double *)malloc(sizeof(double) * (nB > 0 ? nB : 1));
    double *xn = (double *)malloc(sizeof(double) * (nB > 0 ? nB : 1));
    memcpy(x, b, sizeof(double) * nB);
    for (int it = 0; it < g_sweeps; it++) {
        if (!transposed) {
            for (int i = 0; i < nB; i++) {
                double s = 0.0;
                for (int k = ptr[i]; k < ptr[i + 1]; k++) s += val[k] * x[idx[k]];
                xn[i] = b[i] - s;
            }
        } else {
            memcpy(xn, b, sizeof(double) * nB);
            for (int i = 0; i < nB; i++)
                for (int k = ptr[i]; k < ptr[i + 1]; k++) xn[idx[k]] -= val[k] * x[i];
        }
        double *t = x;
        x = xn;
        xn = t;
    }
    memcpy(b, x, sizeof(double) * nB);
    free(x);
    free(xn);
}

/* x_B = L_B^-T D^-1 L_B^-1 x_B in place (exact, or in sweep mode) */
static void solve_b(const jdo_ildl_level *V, double *x)
{
    const int nB = V->nB;
    if (g_sweeps > 0) jacobi_lower(nB, V->lb_ptr, V->lb_idx, V->lb_val, x, 0);
    else for (int i = 0; i < nB; i++) {
        double s = 0.0;
        for (int k = V->lb_ptr[i]; k < V->lb_ptr[i + 1]; k++) s += V->lb_val[k] * x[V->lb_idx[k]];
        x[i] -= s;
    }
    dsolve(V, x);
    if (g_sweeps > 0) jacobi_lower(nB, V->lb_ptr, V->lb_idx, V->lb_val, x, 1);
    else for (int i = nB - 1; i >= 0; i--) {
        double xi = x[i];
        for (int k = V->lb_ptr[i]; k < V->lb_ptr[i + 1]; k++) x[V->lb_idx[k]] -= V->lb_val[k] * xi;
    }
}

/* coupling block of the factor being applied: 0 F^T = A_CB, 1 L_CB */
static int g_coupling = 1;

static void level_apply(int l, int nlev, const jdo_ildl_level *L, int m, const double *minv,
                        const double *in, double *out);

static void level_apply_f(int l, int nlev, const jdo_ildl_level *L, int m, const double *minv,
                          const double *in, double *out)
{
    const jdo_ildl_level *V = &L[l];
    const int n = V->n, nB = V->nB, nC = n - nB;
    double *buf = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *zc = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1));
    double *t = (double *)malloc(sizeof(double) * (nB > 0 ? nB : 1));
    for (int i = 0; i < n; i++) {
        int p = V->perm[i];
        buf[i] = V->scal[p] * in[p];
    }
    solve_b(V, buf); /* y_B */
    for (int c = 0; c < nC; c++) {
        double s = 0.0;
        for (int k = V->lc_ptr[c]; k < V->lc_ptr[c + 1]; k++) s += V->lc_val[k] * buf[V->lc_idx[k]];
        buf[nB + c] -= s;
    }
    if (nC > 0) {
        level_apply(l + 1, nlev, L, m, minv, buf + nB, zc);
        for (int b = 0; b < nB; b++) t[b] = 0.0;
        for (int c = 0; c < nC; c++)
            for (int k = V->lc_ptr[c]; k < V->lc_ptr[c + 1]; k++)
                t[V->lc_idx[k]] += V->lc_val[k] * zc[c];
        solve_b(V, t);
        for (int b = 0; b < nB; b++) buf[b] -= t[b];
    }
    for (int i = 0; i < n; i++) {
        int p = V->perm[i];
        out[p] = V->scal[p] * (i < nB ? buf[i] : zc[i - nB]);
    }
    free(buf);
    free(zc);
    free(t);
}

static void level_apply(int l, int nlev, const jdo_ildl_level *L, int m, const double *minv,
                        const double *in, double *out)
{
    if (l < nlev && g_coupling == 0) {
        level_apply_f(l, nlev, L, m, minv, in, out);
        return;
    }
    if (l == nlev) {
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int j = 0; j < m; j++) s += minv[(long)i * m + j] * in[j];
            out[i] = s;
        }
        return;
    }
    const jdo_ildl_level *V = &L[l];
    const int n = V->n, nB = V->nB, nC = n - nB;
    double *buf = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *zc = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1));
    for (int i = 0; i < n; i++) {
        int p = V->perm[i];
        buf[i] = V->scal[p] * in[p];
    }
    /* forward: L_B y = x_B (rows ascending, entries j < i) */
    if (g_sweeps > 0) jacobi_lower(nB, V->lb_ptr, V->lb_idx, V->lb_val, buf, 0);
    else for (int i = 0; i < nB; i++) {
        double s = 0.0;
        for (int k = V->lb_ptr[i]; k < V->lb_ptr[i + 1]; k++) s += V->lb_val[k] * buf[V->lb_idx[k]];
        buf[i] -= s;
    }
    /* coarse right-hand side: x_C - L_CB y */
    for (int c = 0; c < nC; c++) {
        double s = 0.0;
        for (int k = V->lc_ptr[c]; k < V->lc_ptr[c + 1]; k++) s += V->lc_val[k] * buf[V->lc_idx[k]];
        buf[nB + c] -= s;
    }
    dsolve(V, buf);
    if (nC > 0) level_apply(l + 1, nlev, L, m, minv, buf + nB, zc);
    /* B rows minus L_CB^T z_C */
    for (int c = 0; c < nC; c++)
        for (int k = V->lc_ptr[c]; k < V->lc_ptr[c + 1]; k++)
            buf[V->lc_idx[k]] -= V->lc_val[k] * zc[c];
    /* backward: L_B^T z = t (column-oriented over the rows of L_B) */
    if (g_sweeps > 0) jacobi_lower(nB, V->lb_ptr, V->lb_idx, V->lb_val, buf, 1);
    else for (int i = nB - 1; i >= 0; i--) {
        double zi = buf[i];
        for (int k = V->lb_ptr[i]; k < V->lb_ptr[i + 1]; k++) buf[V->lb_idx[k]] -= V->lb_val[k] * zi;
    }
    for (int i = 0; i < n; i++) {
        int p = V->perm[i];
        out[p] = V->scal[p] * (i < nB ? buf[i] : zc[i - nB]);
    }
    free(buf);
    free(zc);
}

void jdo_ildl_set_coupling(int c) { g_coupling = c ? 1 : 0; }

void jdo_ildl_apply(int nlev, const jdo_ildl_level *levels, int m, const double *minv,
                    const double *r, double *z)
{
    level_apply(0, nlev, levels, m, minv, r, z);
}

/* Adapter with the jdo_psolve_fn signature (see jd_oracle.h). */
typedef struct {
    int nlev;
    const jdo_ildl_level *levels;
    int m;
    const double *minv;
} jdo_ildl_ctx;

void jdo_ildl_psolve(void *ctx, int n, double *out, const double *in)
{
    const jdo_ildl_ctx *c = (const jdo_ildl_ctx *)ctx;
    (void)n;
    level_apply(0, c->nlev, c->levels, c->m, c->minv, in, out);
}
