// complex.cpp -- the complex Hermitian drivers ZPJD, ZPJDREVCOM, ZJDREVCOM,
// ZPJD_GEP, ZPJDREVCOM_GEP, ZJDREVCOM_GEP and ZPJDCLEANUP (argument meaning
// Doc/ZPJD_prototypes.f, C prototypes src/include/jadamilu1.h:62-108).
//
// A Hermitian matrix H = H_r + i H_i of order n acts on z = x + i y like the
// real symmetric matrix of order 2n
//
//     R(H) = [ H_r  -H_i ]   on   [ x ]
//            [ H_i   H_r ]        [ y ],
//
// and in the interleaved order (x_1, y_1, x_2, y_2, ...), which is how a
// DOUBLE COMPLEX array lies in memory, the entry H_jk = a + i b becomes the
// 2 x 2 block [[a, -b], [b, a]] at rows 2j-1, 2j and columns 2k-1, 2k.  The
// caller's complex vectors therefore ARE vectors of R(H), with no copy or
// reordering.  R(H) has the spectrum of H with every eigenvalue twice (the
// eigenvectors z and i z), so each driver here runs the real solver of this
// library (device SpMV, multilevel ILDL factor of R(H), device Jacobi-Davidson
// engine) on R(H) for twice the wanted number of eigenpairs, then keeps one
// vector per complex direction.  The real kernels, the FP64 matrix-core Gram
// products and the factor's level schedules serve the complex problem
// unchanged; the price is that each complex eigenvalue is converged twice.
//
// Selection: the converged real vectors v_1, v_2, ... (in the order the
// solver returns them) are taken in turn; v_k is kept when its component
// outside the complex span of the vectors already kept is larger than 1e-4
// (the second copy i z of a kept z lies in that span up to the convergence
// error).  Of the kept directions (one more than the pairs when a pair is
// broken at the end of the run) the NEIG smallest eigenvalues, or for
// ISEARCH >= 2 the NEIG nearest SIGMA, are returned in that order.  For the
// generalized problem the copy of a direction with the smaller residual is
// the one returned.  For the standard problem a kept vector is made orthogonal, in the
// complex inner product, to the kept vectors, so that a multiple eigenvalue
// returns complex-orthonormal vectors as the reference's complex iteration
// does; RES is then recomputed from H (ZPJD) or bounded by the triangle
// inequality from the real residuals (reverse communication, where H is the
// caller's).  For the generalized problem the kept vectors are returned as
// the solver produced them (B-normalised) with the solver's residuals
// ||A z - lambda B z|| / ||B z||.
//
// Reverse communication: the real solver's requests refer to its own device
// workspace (twice the eigenvector columns of the caller's LX); each request
// is forwarded to the caller with fixed slots of the caller's X: the input
// vector in X(1:N), A's product in X(N+1:2N), B's product in X(2N+1:3N)
// (IJOB = 2: the right-hand side in X(N+1:2N), the solution in X(1:N)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "../../include/jadamilu_amd.h"

namespace {

typedef std::complex<double> cplx;

// Real form R(H) of a Hermitian CSR (1-based, complex values interleaved),
// 1-based CSR of order 2n in the triangle of the input.  Exact zeros of the
// real or imaginary part of an off-diagonal entry are not stored; every
// diagonal entry is.  ja[0] < 0 (a diagonal matrix, a[0..n) its entries):
// R(H) diagonal in the same convention.
void real_form(int n, const double *a, const int *ja, const int *ia, std::vector<int> &ia2,
               std::vector<int> &ja2, std::vector<double> &a2) {
    ia2.assign(2 * (size_t)n + 1, 0);
    ja2.clear();
    a2.clear();
    if (ja[0] < 0) {
        ja2.assign(2 * (size_t)n, 0);
        ja2[0] = -1;
        a2.resize(2 * (size_t)n);
        for (int i = 0; i < n; ++i) a2[2 * i] = a2[2 * i + 1] = a[2 * (size_t)i];
        for (int r = 0; r <= 2 * n; ++r) ia2[r] = r + 1;
        return;
    }
    const long nnz = ia[n] - 1;
    ja2.reserve(4 * nnz);
    a2.reserve(4 * nnz);
    for (int i = 0; i < n; ++i) {
        for (int half = 0; half < 2; ++half) {
            const int row = 2 * i + half;
            ia2[row] = (int)ja2.size() + 1;
            for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
                const int j = ja[k] - 1;
                const double re = a[2 * (size_t)k], im = a[2 * (size_t)k + 1];
                if (j == i) {  // [[h, 0], [0, h]] with h = Re H_ii
                    ja2.push_back(row + 1);
                    a2.push_back(re);
                    continue;
                }
                // row 2i: (2j, a), (2j+1, -b); row 2i+1: (2j, b), (2j+1, a)
                const double v0 = half == 0 ? re : im, v1 = half == 0 ? -im : re;
                if (v0 != 0.0) {
                    ja2.push_back(2 * j + 1);
                    a2.push_back(v0);
                }
                if (v1 != 0.0) {
                    ja2.push_back(2 * j + 2);
                    a2.push_back(v1);
                }
            }
        }
    }
    ia2[2 * n] = (int)ja2.size() + 1;
}

// y = H z for the caller's Hermitian CSR (each off-diagonal entry stands for
// itself and its conjugate transpose).
void herm_matvec(int n, const double *a, const int *ja, const int *ia, const cplx *z, cplx *y) {
    std::fill(y, y + n, cplx(0.0, 0.0));
    if (ja[0] < 0) {
        for (int i = 0; i < n; ++i) y[i] = a[2 * (size_t)i] * z[i];
        return;
    }
    for (int i = 0; i < n; ++i)
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
            const int j = ja[k] - 1;
            const cplx h(a[2 * (size_t)k], a[2 * (size_t)k + 1]);
            if (j == i) {
                y[i] += h.real() * z[i];
            } else {
                y[i] += h * z[j];
                y[j] += std::conj(h) * z[i];
            }
        }
}

cplx cdot(long n, const cplx *u, const cplx *v) {  // u^H v
    cplx s(0.0, 0.0);
    for (long i = 0; i < n; ++i) s += std::conj(u[i]) * v[i];
    return s;
}

double cnorm(long n, const cplx *u) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += std::norm(u[i]);
    return std::sqrt(s);
}

// One complex direction per pair of real eigenvectors (see the header
// comment).  v: m real vectors of R(H) (2n doubles each, as complex n-vectors),
// eigs/res: the solver's values.  Writes at most `want` vectors to out
// (complex, n each) with their eigenvalues and residual norms (for orth, the
// triangle-inequality bound); returns the count.  orth: orthogonalise the
// kept vectors (standard problem).
int select_complex(int n, int m, const double *v, const double *eigs, const double *res,
                   int want, bool orth, cplx *out, double *eig_out, double *res_out) {
    const cplx *V = reinterpret_cast<const cplx *>(v);
    std::vector<cplx> Q;  // orthonormal basis of the kept complex span
    std::vector<cplx> r(n);
    std::vector<cplx> coef;
    int cnt = 0;
    for (int k = 0; k < m && cnt < want; ++k) {
        const cplx *z = V + (long)k * n;
        std::copy(z, z + n, r.begin());
        const double nz = cnorm(n, z);
        if (!(nz > 0.0)) continue;
        coef.assign(cnt, cplx(0.0, 0.0));
        for (int pass = 0; pass < 2; ++pass)  // classical Gram-Schmidt, twice
            for (int q = 0; q < cnt; ++q) {
                const cplx *Qq = Q.data() + (long)q * n;
                const cplx c = cdot(n, Qq, r.data());
                coef[q] += c;
                for (int i = 0; i < n; ++i) r[i] -= c * Qq[i];
            }
        const double nr = cnorm(n, r.data());
        if (nr <= 1e-4 * nz) {  // the partner of a kept direction
            if (!orth && cnt > 0) {
                // keep whichever copy has the smaller residual (the kept
                // direction is the one z projects on most)
                int q = 0;
                for (int t = 1; t < cnt; ++t)
                    if (std::abs(coef[t]) > std::abs(coef[q])) q = t;
                if (res[k] < res_out[q]) {
                    std::copy(z, z + n, out + (long)q * n);
                    res_out[q] = res[k];
                    eig_out[q] = eigs[k];
                }
            }
            continue;
        }
        for (int i = 0; i < n; ++i) r[i] /= nr;
        Q.insert(Q.end(), r.begin(), r.end());
        cplx *o = out + (long)cnt * n;
        if (orth) {
            std::copy(r.begin(), r.end(), o);
            // z = sum_q coef_q Q_q + nr r, so H r - eig r = (H z - eig z
            // - sum_q coef_q (H Q_q - eig Q_q)) / nr; |H Q_q - eig Q_q| is
            // bounded by res_q + |eig_q - eig|
            double b = res[k] / nz;
            for (int q = 0; q < cnt; ++q)
                b += std::abs(coef[q]) / nz * (res_out[q] + std::abs(eig_out[q] - eigs[k]));
            res_out[cnt] = b * nz / nr;
        } else {  // as the solver normalised it, with its residual
            std::copy(z, z + n, o);
            res_out[cnt] = res[k];
        }
        eig_out[cnt] = eigs[k];
        ++cnt;
    }
    return cnt;
}

// ---- state of the complex drivers ----------------------------------------

struct ZState {
    jdhip_ctx *ctx = nullptr;
    double *xd = nullptr;  // device workspace of the real solver
    long lxd = 0;
    // real form of the caller's matrices, kept across reverse-communication
    // calls (the real driver reads them again when it refactors)
    std::vector<int> ia2, ja2, ib2, jb2;
    std::vector<double> a2, b2;
    std::vector<double> eigs2, res2;
    int neig2 = 0, nwant = 0, n = 0, isearch = 0;
    double sigma = 0.0;
    bool gep = false;
    int pend = 0;                    // IJOB handed to the caller (0: none)
    long ndx1 = 0, ndx2 = 0, ndx3 = 0;  // the real solver's request (1-based)
};

std::mutex g_zmu;
ZState *g_z = nullptr;

ZState *zstate() {
    std::lock_guard<std::mutex> lk(g_zmu);
    if (!g_z) {
        g_z = new ZState();
        int dev = 0;
        if (const char *e = std::getenv("JDHIP_DEVICE")) dev = std::atoi(e);
        g_z->ctx = jdhip_create(dev);
    }
    return g_z->ctx ? g_z : nullptr;
}

void sync_copy(ZState &Z, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    hipStream_t s = (hipStream_t)jdhip_stream(Z.ctx);
    if (hipMemcpyAsync(dst, src, bytes, kind, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        throw std::runtime_error("complex driver: HIP copy failed");
}

void ensure_ws(ZState &Z, long lx) {
    if (lx <= Z.lxd) return;
    if (Z.xd) (void)hipFree(Z.xd);
    Z.xd = nullptr;
    Z.lxd = 0;
    if (hipMalloc(&Z.xd, sizeof(double) * lx) != hipSuccess) {
        Z.xd = nullptr;
        throw std::runtime_error("complex driver: device workspace allocation failed");
    }
    Z.lxd = lx;
}

// Argument checks of the complex drivers in the caller's units (JDRVCOM's,
// JDaux.F:1502-1524, with the complex LX), plus this library's limit on the
// columns of the real solver (twice the wanted count + MADSPACE <= 64).
int check_args(int n, int lx, int neig, int isearch, int madspace, int iter, double tol,
               bool gep) {
    const long N = n, M = madspace;
    if (n < 2) return -1;
    if (neig <= 0) return -5;
    if (madspace <= 2) return -9;
    if (N * (2 * M + neig + 4) + 3 * M * M + std::max<long>(M * M, neig) > lx) return -10;
    if (gep && N * (3 * M + 2L * neig + 4) + 3 * M * M + std::max<long>(M * M, neig) > lx)
        return -10;
    if (gep && isearch >= 2 &&
        N * (3 * M + 2L * neig) + 6 * M + 3 * M * M + std::max<long>(M * M, neig) > lx)
        return -10;
    if (iter <= 0) return -11;
    if (tol <= 0.0) return -12;
    if (2 * std::min(neig, n) + std::min(madspace, 2 * n) > 64) return -9;
    return 0;
}

// Start of a run: real forms, device workspace, initial guesses.
void start(ZState &Z, int n, const double *a, const int *ja, const int *ia, const double *b,
           const int *jb, const int *ib, const double *eigs, const double *x, int neig,
           int ninit, int madspace, bool gep, int isearch, double sigma) {
    Z.n = n;
    Z.isearch = isearch;
    Z.sigma = sigma;
    Z.gep = gep;
    Z.nwant = neig;
    Z.neig2 = 2 * neig;
    Z.pend = 0;
    Z.ndx1 = Z.ndx2 = Z.ndx3 = 0;
    if (a) real_form(n, a, ja, ia, Z.ia2, Z.ja2, Z.a2);
    if (b) real_form(n, b, jb, ib, Z.ib2, Z.jb2, Z.b2);
    const long n2 = 2L * n, M = madspace;
    // optimal workspace of the real drivers (PJD.F:99-103, 544-545) for
    // order 2n and twice the wanted count
    const long lx2 = gep ? n2 * (4 * M + 2L * Z.neig2 + 1) + 3 * M * M +
                               std::max<long>(M * M, Z.neig2)
                         : n2 * (3 * M + Z.neig2 + 1) + 3 * M * M +
                               std::max<long>(M * M, Z.neig2);
    ensure_ws(Z, lx2);
    Z.eigs2.assign(Z.neig2, 0.0);
    Z.res2.assign(Z.neig2, 0.0);
    const int ni = std::max(0, std::min(ninit, neig));
    for (int k = 0; k < ni; ++k) Z.eigs2[k] = eigs[k];
    if (ni > 0) sync_copy(Z, Z.xd, x, sizeof(double) * n2 * ni, hipMemcpyHostToDevice);
}

// End of a run: one vector per complex direction into the caller's X.
void finish(ZState &Z, const double *a, const int *ja, const int *ia, double *eigs, double *res,
            double *x, int *neig) {
    const int n = Z.n;
    const long n2 = 2L * n;
    const int m = std::max(0, std::min(*neig, 2 * Z.nwant));
    if (m == 0) {
        *neig = 0;
        return;
    }
    std::vector<double> v(n2 * m);
    sync_copy(Z, v.data(), Z.xd, sizeof(double) * v.size(), hipMemcpyDeviceToHost);
    // every complex direction among the converged vectors (a pair broken at
    // the end of the run leaves one more direction than pairs) ...
    std::vector<cplx> cand((size_t)m * n);
    std::vector<double> e(m), r(m);
    const int cnt = select_complex(n, m, v.data(), Z.eigs2.data(), Z.res2.data(), m, !Z.gep,
                                   cand.data(), e.data(), r.data());
    // ... of which the wanted count is kept: the smallest eigenvalues
    // (ISEARCH <= 1) or those nearest SIGMA (ISEARCH >= 2)
    std::vector<int> ord(cnt);
    for (int k = 0; k < cnt; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) {
        return Z.isearch >= 2 ? std::abs(e[p] - Z.sigma) < std::abs(e[q] - Z.sigma) : e[p] < e[q];
    });
    const int keep = std::min(cnt, Z.nwant);
    cplx *out = reinterpret_cast<cplx *>(x);
    for (int k = 0; k < keep; ++k) {
        std::copy(cand.begin() + (size_t)ord[k] * n, cand.begin() + (size_t)(ord[k] + 1) * n,
                  out + (size_t)k * n);
        eigs[k] = e[ord[k]];
        res[k] = r[ord[k]];
    }
    if (a && !Z.gep) {  // the matrix is known: exact residual norms
        std::vector<cplx> y(n);
        for (int k = 0; k < keep; ++k) {
            const cplx *z = out + (long)k * n;
            herm_matvec(n, a, ja, ia, z, y.data());
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += std::norm(y[i] - eigs[k] * z[i]);
            res[k] = std::sqrt(s);
        }
    }
    *neig = keep;
}

// Forward the real solver's request to the caller's X slots.
void forward(ZState &Z, double *x, int ijob2, int *ijob, int *ndx1, int *ndx2, int *ndx3) {
    const long n2 = 2L * Z.n;
    const int N = Z.n;
    Z.pend = ijob2;
    if (ijob2 == 2) {  // X(NDX1) = P^-1 X(NDX2): rhs to X(N+1:2N)
        sync_copy(Z, x + n2, Z.xd + (Z.ndx2 - 1), sizeof(double) * n2, hipMemcpyDeviceToHost);
        *ndx1 = 1;
        *ndx2 = N + 1;
        if (ndx3) *ndx3 = 2 * N + 1;
    } else {  // products of X(NDX1)
        sync_copy(Z, x, Z.xd + (Z.ndx1 - 1), sizeof(double) * n2, hipMemcpyDeviceToHost);
        *ndx1 = 1;
        *ndx2 = N + 1;
        if (ndx3) *ndx3 = 2 * N + 1;
    }
    *ijob = ijob2;
}

// Take back the caller's answer to the pending request.
void take_back(ZState &Z, const double *x) {
    const long n2 = 2L * Z.n;
    switch (Z.pend) {
    case 1:
        sync_copy(Z, Z.xd + (Z.ndx2 - 1), x + n2, sizeof(double) * n2, hipMemcpyHostToDevice);
        if (Z.gep)
            sync_copy(Z, Z.xd + (Z.ndx3 - 1), x + 2 * n2, sizeof(double) * n2,
                      hipMemcpyHostToDevice);
        break;
    case 2:
        sync_copy(Z, Z.xd + (Z.ndx1 - 1), x, sizeof(double) * n2, hipMemcpyHostToDevice);
        break;
    case 3:
        sync_copy(Z, Z.xd + (Z.ndx2 - 1), x + n2, sizeof(double) * n2, hipMemcpyHostToDevice);
        break;
    case 4:
        sync_copy(Z, Z.xd + (Z.ndx3 - 1), x + 2 * n2, sizeof(double) * n2,
                  hipMemcpyHostToDevice);
        break;
    default:
        break;
    }
}

enum Kind { kPjd, kPjdRevcom, kJdRevcom };

// One call of a complex reverse-communication driver (or a whole ZPJD run
// when kind = kPjd).  a/b: the caller's complex matrices (null where the
// driver has none).
void zdrive(Kind kind, bool gep, int *n, double *a, int *ja, int *ia, double *b, int *jb,
            int *ib, double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
            int *isearch, int *ninit, int *madspace, int *iter, double *tol, double *shift,
            double *droptol, double *mem, int *icntl, int *ijob, int *ndx1, int *ndx2,
            int *ndx3, int *iprint, int *info, double *gap) {
    ZState *Zp = zstate();
    if (!Zp) {
        *info = JDHIP_INFO_DEVICE_ERROR;
        if (ijob) *ijob = 0;
        return;
    }
    ZState &Z = *Zp;
    try {
        const bool first = kind == kPjd || *ijob == 0;
        int ijob2 = 0;  // the real driver's IJOB, handed back unchanged
        if (first) {
            const int chk = check_args(*n, *lx, *neig, *isearch, *madspace, *iter, *tol, gep);
            if (chk != 0) {
                *info = chk;
                if (ijob) *ijob = 0;
                return;
            }
            start(Z, *n, kind == kJdRevcom ? nullptr : a, ja, ia,
                  (gep && kind != kJdRevcom) ? b : nullptr, jb, ib, eigs, x, *neig, *ninit,
                  *madspace, gep, *isearch, *sigma);
        } else {
            if (Z.pend == 0) {  // no request outstanding: a bad resume (JDaux.F:1500)
                *info = -15;
                *ijob = 0;
                return;
            }
            take_back(Z, x);
            ijob2 = Z.pend;
            Z.pend = 0;
        }
        const int n2 = 2 * Z.n;
        long d1 = Z.ndx1, d2 = Z.ndx2, d3 = Z.ndx3;
        const int ni = std::max(0, std::min(*ninit, *neig));
        if (kind == kPjd) {
            if (gep)
                jdhip_pjd_gep(Z.ctx, n2, Z.a2.data(), Z.ja2.data(), Z.ia2.data(), Z.b2.data(),
                              Z.jb2.data(), Z.ib2.data(), Z.eigs2.data(), Z.res2.data(), Z.xd,
                              Z.lxd, &Z.neig2, *sigma, *isearch, ni, *madspace, iter, *tol,
                              shift, droptol, mem, icntl, *iprint, info, gap);
            else
                jdhip_pjd(Z.ctx, n2, Z.a2.data(), Z.ja2.data(), Z.ia2.data(), Z.eigs2.data(),
                          Z.res2.data(), Z.xd, Z.lxd, &Z.neig2, *sigma, *isearch, ni,
                          *madspace, iter, *tol, shift, droptol, mem, icntl, *iprint, info,
                          gap);
            ijob2 = 0;
        } else if (kind == kPjdRevcom) {
            if (gep)
                jdhip_pjdrevcom_gep(Z.ctx, n2, Z.a2.data(), Z.ja2.data(), Z.ia2.data(),
                                    Z.b2.data(), Z.jb2.data(), Z.ib2.data(), Z.eigs2.data(),
                                    Z.res2.data(), Z.xd, Z.lxd, &Z.neig2, *sigma, *isearch, ni,
                                    *madspace, iter, *tol, shift, droptol, mem, icntl,
                                    &ijob2, &d1, &d2, &d3, *iprint, info, gap);
            else
                jdhip_pjdrevcom(Z.ctx, n2, Z.a2.data(), Z.ja2.data(), Z.ia2.data(),
                                Z.eigs2.data(), Z.res2.data(), Z.xd, Z.lxd, &Z.neig2, *sigma,
                                *isearch, ni, *madspace, iter, *tol, shift, droptol, mem,
                                icntl, &ijob2, &d1, &d2, *iprint, info, gap);
        } else {
            if (gep)
                jdhip_jdrevcom_gep(Z.ctx, n2, Z.eigs2.data(), Z.res2.data(), Z.xd, Z.lxd,
                                   &Z.neig2, *sigma, *isearch, ni, *madspace, iter, *tol,
                                   &ijob2, &d1, &d2, &d3, *iprint, info, gap);
            else
                jdhip_jdrevcom(Z.ctx, n2, Z.eigs2.data(), Z.res2.data(), Z.xd, Z.lxd,
                               &Z.neig2, *sigma, *isearch, ni, *madspace, iter, *tol, &ijob2,
                               &d1, &d2, *iprint, info, gap);
        }
        Z.ndx1 = d1;
        Z.ndx2 = d2;
        Z.ndx3 = d3;
        if (ijob2 != 0) {  // a request for the caller
            forward(Z, x, ijob2, ijob, ndx1, ndx2, gep ? ndx3 : nullptr);
            return;
        }
        finish(Z, kind == kPjd ? a : nullptr, ja, ia, eigs, res, x, &Z.neig2);
        *neig = Z.neig2;
        // the products ran out with one copy of the last direction missing:
        // every wanted complex eigenpair was found
        if (*info > 0 && Z.neig2 == Z.nwant) *info = 0;
        if (ijob) *ijob = 0;
    } catch (const std::exception &) {
        *info = JDHIP_INFO_DEVICE_ERROR;
        if (ijob) *ijob = 0;
        Z.pend = 0;
    }
}

}  // namespace

extern "C" {

void zpjd_(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x, int *lx,
           int *neig, double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
           double *tol, double *shift, double *droptol, double *mem, int *icntl, int *iprint,
           int *info, double *gap) {
    zdrive(kPjd, false, n, a, ja, ia, nullptr, nullptr, nullptr, eigs, res, x, lx, neig, sigma,
           isearch, ninit, madspace, iter, tol, shift, droptol, mem, icntl, nullptr, nullptr,
           nullptr, nullptr, iprint, info, gap);
}

void zpjdrevcom_(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
                 int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
                 int *iter, double *tol, double *shift, double *droptol, double *mem, int *icntl,
                 int *ijob, int *ndx1, int *ndx2, int *iprint, int *info, double *gap) {
    zdrive(kPjdRevcom, false, n, a, ja, ia, nullptr, nullptr, nullptr, eigs, res, x, lx, neig,
           sigma, isearch, ninit, madspace, iter, tol, shift, droptol, mem, icntl, ijob, ndx1,
           ndx2, nullptr, iprint, info, gap);
}

void zjdrevcom_(int *n, double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
                int *isearch, int *ninit, int *madspace, int *iter, double *tol, int *ijob,
                int *ndx1, int *ndx2, int *iprint, int *info, double *gap) {
    zdrive(kJdRevcom, false, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, eigs, res,
           x, lx, neig, sigma, isearch, ninit, madspace, iter, tol, nullptr, nullptr, nullptr,
           nullptr, ijob, ndx1, ndx2, nullptr, iprint, info, gap);
}

void zpjd_gep_(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib, double *eigs,
               double *res, double *x, int *lx, int *neig, double *sigma, int *isearch,
               int *ninit, int *madspace, int *iter, double *tol, double *shift,
               double *droptol, double *mem, int *icntl, int *iprint, int *info, double *gap) {
    zdrive(kPjd, true, n, a, ja, ia, b, jb, ib, eigs, res, x, lx, neig, sigma, isearch, ninit,
           madspace, iter, tol, shift, droptol, mem, icntl, nullptr, nullptr, nullptr, nullptr,
           iprint, info, gap);
}

void zpjdrevcom_gep_(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib,
                     double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
                     int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                     double *shift, double *droptol, double *mem, int *icntl, int *ijob,
                     int *ndx1, int *ndx2, int *ndx3, int *iprint, int *info, double *gap) {
    zdrive(kPjdRevcom, true, n, a, ja, ia, b, jb, ib, eigs, res, x, lx, neig, sigma, isearch,
           ninit, madspace, iter, tol, shift, droptol, mem, icntl, ijob, ndx1, ndx2, ndx3,
           iprint, info, gap);
}

void zjdrevcom_gep_(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                    double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
                    double *tol, int *ijob, int *ndx1, int *ndx2, int *ndx3, int *iprint,
                    int *info, double *gap) {
    zdrive(kJdRevcom, true, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, eigs, res, x,
           lx, neig, sigma, isearch, ninit, madspace, iter, tol, nullptr, nullptr, nullptr,
           nullptr, ijob, ndx1, ndx2, ndx3, iprint, info, gap);
}

jdhip_ctx *jdhip_complex_context(void) {
    ZState *Z = zstate();
    return Z ? Z->ctx : nullptr;
}

// The real symmetric form used by the drivers, exposed for checking.
long jdhip_hermitian_real_form(int n, const double *a, const int *ja, const int *ia, int *ia2,
                               int *ja2, double *a2, long cap) {
    if (n <= 0 || !a || !ja || (ja[0] >= 0 && !ia)) return -1;
    std::vector<int> i2, j2;
    std::vector<double> v2;
    real_form(n, a, ja, ia, i2, j2, v2);
    const long cnt = (long)j2.size();
    if (!ia2 || !ja2 || !a2 || cap < cnt) return cnt;
    std::copy(i2.begin(), i2.end(), ia2);
    std::copy(j2.begin(), j2.end(), ja2);
    std::copy(v2.begin(), v2.end(), a2);
    return cnt;
}

void zpjdcleanup_(void) {
    std::lock_guard<std::mutex> lk(g_zmu);
    if (!g_z) return;
    if (g_z->ctx) jdhip_cleanup(g_z->ctx);
    if (g_z->xd) (void)hipFree(g_z->xd);
    g_z->xd = nullptr;
    g_z->lxd = 0;
    g_z->pend = 0;
}

// names of a reference build without -D__UNDERSCORE__ (namesjadamilu.h:53-88)
void zpjd(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x, int *lx,
          int *neig, double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
          double *tol, double *shift, double *droptol, double *mem, int *icntl, int *iprint,
          int *info, double *gap) {
    zpjd_(n, a, ja, ia, eigs, res, x, lx, neig, sigma, isearch, ninit, madspace, iter, tol, shift,
          droptol, mem, icntl, iprint, info, gap);
}
void zpjdrevcom(int *n, double *a, int *ja, int *ia, double *eigs, double *res, double *x,
                int *lx, int *neig, double *sigma, int *isearch, int *ninit, int *madspace,
                int *iter, double *tol, double *shift, double *droptol, double *mem, int *icntl,
                int *ijob, int *ndx1, int *ndx2, int *iprint, int *info, double *gap) {
    zpjdrevcom_(n, a, ja, ia, eigs, res, x, lx, neig, sigma, isearch, ninit, madspace, iter, tol,
                shift, droptol, mem, icntl, ijob, ndx1, ndx2, iprint, info, gap);
}
void zjdrevcom(int *n, double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
               int *isearch, int *ninit, int *madspace, int *iter, double *tol, int *ijob,
               int *ndx1, int *ndx2, int *iprint, int *info, double *gap) {
    zjdrevcom_(n, eigs, res, x, lx, neig, sigma, isearch, ninit, madspace, iter, tol, ijob, ndx1,
               ndx2, iprint, info, gap);
}
void zpjdcleanup(void) { zpjdcleanup_(); }
void zpjd_gep(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib, double *eigs,
              double *res, double *x, int *lx, int *neig, double *sigma, int *isearch,
              int *ninit, int *madspace, int *iter, double *tol, double *shift, double *droptol,
              double *mem, int *icntl, int *iprint, int *info, double *gap) {
    zpjd_gep_(n, a, ja, ia, b, jb, ib, eigs, res, x, lx, neig, sigma, isearch, ninit, madspace,
              iter, tol, shift, droptol, mem, icntl, iprint, info, gap);
}
void zpjdrevcom_gep(int *n, double *a, int *ja, int *ia, double *b, int *jb, int *ib,
                    double *eigs, double *res, double *x, int *lx, int *neig, double *sigma,
                    int *isearch, int *ninit, int *madspace, int *iter, double *tol,
                    double *shift, double *droptol, double *mem, int *icntl, int *ijob,
                    int *ndx1, int *ndx2, int *ndx3, int *iprint, int *info, double *gap) {
    zpjdrevcom_gep_(n, a, ja, ia, b, jb, ib, eigs, res, x, lx, neig, sigma, isearch, ninit,
                    madspace, iter, tol, shift, droptol, mem, icntl, ijob, ndx1, ndx2, ndx3,
                    iprint, info, gap);
}
void zjdrevcom_gep(int *n, double *eigs, double *res, double *x, int *lx, int *neig,
                   double *sigma, int *isearch, int *ninit, int *madspace, int *iter,
                   double *tol, int *ijob, int *ndx1, int *ndx2, int *ndx3, int *iprint,
                   int *info, double *gap) {
    zjdrevcom_gep_(n, eigs, res, x, lx, neig, sigma, isearch, ninit, madspace, iter, tol, ijob,
                   ndx1, ndx2, ndx3, iprint, info, gap);
}

}  // extern "C"
