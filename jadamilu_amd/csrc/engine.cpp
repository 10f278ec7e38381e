// engine.cpp -- device-resident Jacobi-Davidson engine (see engine.hpp).
//
// Step-by-step restatement of JDRVCOM (src/jd/JDaux.F:1391-2415).  Labels
// L<n> carry the reference's statement labels so each block can be checked
// against the Fortran.  n-length operations are HIP kernels on ctx->stream;
// the projected J x J problems run on the host.
#include "engine.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "dense.hpp"

namespace jdhip {

static constexpr double kEps = DBL_EPSILON * 0.5;  // dlamch('E')
static constexpr double kInf = DBL_MAX;            // dlamch('O')

// ---------------------------------------------------------------------------
void DevCtx::init(int dev) {
    device = dev;
    JDHIP_CHECK(hipSetDevice(dev));
    JDHIP_CHECK(hipStreamCreate(&stream));
    JDHIP_CHECK(hipMalloc(&rb.partials, sizeof(double) * kMaxPartials));
    JDHIP_CHECK(hipMalloc(&rb.counter, sizeof(unsigned) * 16));
    JDHIP_CHECK(hipMemsetAsync(rb.counter, 0, sizeof(unsigned) * 16, stream));
    JDHIP_CHECK(hipMalloc(&d_small, sizeof(double) * 8192));
    JDHIP_CHECK(hipMemsetAsync(d_small, 0, sizeof(double) * 8192, stream));
    rb.out = d_small;
    JDHIP_CHECK(hipMalloc(&d_sc, sizeof(InnerScalars)));
    JDHIP_CHECK(hipMemsetAsync(d_sc, 0, sizeof(InnerScalars), stream));
    JDHIP_CHECK(hipStreamSynchronize(stream));
    JDHIP_CHECK(hipHostMalloc(&h_pin, sizeof(double) * 8192, hipHostMallocDefault));
    JDHIP_CHECK(hipHostMalloc(&h_sc, sizeof(InnerScalars), hipHostMallocDefault));
    std::memset(h_sc, 0, sizeof(InnerScalars));
    JDHIP_CHECK(hipMalloc(&d_gram, sizeof(double) * kGramScratch));
    JDHIP_CHECK(hipMalloc(&d_gout, sizeof(double) * kMaxCols * kMaxCols));
    JDHIP_CHECK(hipEventCreateWithFlags(&ev_sync, hipEventBlockingSync | hipEventDisableTiming));
}

void DevCtx::release() {
    if (!stream) return;
    hipSetDevice(device);
    hipStreamSynchronize(stream);
    hipFree(rb.partials);
    hipFree(rb.counter);
    hipFree(d_small);
    hipFree(d_sc);
    hipFree(d_tmp);
    hipFree(d_gram);
    hipFree(d_gout);
    hipHostFree(h_pin);
    hipHostFree(h_sc);
    if (ev_sync) hipEventDestroy(ev_sync);
    ev_sync = nullptr;
    hipStreamDestroy(stream);
    stream = nullptr;
    d_tmp = nullptr;
    tmp_len = 0;
}

void DevCtx::ensure_tmp(long len) {
    if (len <= tmp_len) return;
    if (d_tmp) {
        JDHIP_CHECK(hipStreamSynchronize(stream));
        JDHIP_CHECK(hipFree(d_tmp));
    }
    JDHIP_CHECK(hipMalloc(&d_tmp, sizeof(double) * len));
    tmp_len = len;
}

// ---------------------------------------------------------------------------
void JdEngine::fetch(double *host, const double *dev, long cnt) {
    if (cnt <= 0) return;
    JDHIP_CHECK(hipMemcpyAsync(ctx_->h_pin, dev, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                               ctx_->stream));
    ctx_->sync();
    ++nsync;
    std::memcpy(host, ctx_->h_pin, sizeof(double) * cnt);
}

void JdEngine::put(double *dev, const double *host, long cnt) {
    if (cnt <= 0) return;
    std::memcpy(ctx_->h_pin + 4096, host, sizeof(double) * cnt);
    JDHIP_CHECK(hipMemcpyAsync(dev, ctx_->h_pin + 4096, sizeof(double) * cnt,
                               hipMemcpyHostToDevice, ctx_->stream));
    ctx_->sync();
}

void JdEngine::ensure_f32(long n, int nq) {
    long need = n * (6L + nq);
    if (need <= fbuf_len_) return;
    if (fbuf_) {
        ctx_->sync();
        JDHIP_CHECK(hipFree(fbuf_));
    }
    JDHIP_CHECK(hipMalloc(&fbuf_, sizeof(float) * need));
    fbuf_len_ = need;
}

double JdEngine::dnorm(const double *w, long n) {
    double *d = ctx_->slot(2);
    dot(ctx_->stream, ctx_->rb, w, w, n, d);
    double v;
    fetch(&v, d, 1);
    return std::sqrt(v);
}

double JdEngine::ddot(const double *a, const double *b, long n) {
    double *d = ctx_->slot(2);
    dot(ctx_->stream, ctx_->rb, a, b, n, d);
    double v;
    fetch(&v, d, 1);
    return v;
}

// REPGS (JDaux.F:2674-2753): iterated classical Gram-Schmidt of w, the
// coefficients y = BV^T w from the jv columns at bvoff and the update
// w -= V y with those at voff (V = BV for the standard problem).  y[0:jv)
// coefficients, y[jv] final norm (gam >= 0).
void JdEngine::repgs2(int n, long voff, long bvoff, int jv, long woff, std::vector<double> &y,
                      int gam) {
    if (jv <= 0) return;
    if ((int)y.size() < 2 * jv + 2) y.resize(2 * jv + 2);
    hipStream_t s = ctx_->stream;
    const double *V = X_ + voff;
    const double *BV = X_ + bvoff;
    double *w = X_ + woff;
    double *dy = ctx_->slot(0);
    if (gam < 0) {
        multidot(s, ctx_->rb, BV, n, jv, w, n, dy);
        gemv_n(s, ctx_->rb, w, V, n, jv, dy, -1.0, 1.0, n, nullptr);
        fetch(y.data(), dy, jv);
        return;
    }
    bool nrnd = true;
restart : {
    double nrn = dnorm(w, n), nro;
    double nr = kEps * nrn;
    if (nrnd) std::fill(y.begin(), y.begin() + jv, 0.0);
    for (;;) {
        nro = nrn;
        multidot(s, ctx_->rb, BV, n, jv, w, n, dy);
        gemv_n(s, ctx_->rb, w, V, n, jv, dy, -1.0, 1.0, n, ctx_->slot(1));
        double buf[kMaxCols + 1];
        fetch(buf, dy, kMaxCols + 1);
        nrn = std::sqrt(buf[kMaxCols]);
        if (nrnd)
            for (int c = 0; c < jv; ++c) y[c] += buf[c];
        if (nrn <= nr) goto fail;
        if (!(nrn < 0.5 * nro)) break;
    }
    y[jv] = nrn;
    return;
}
fail:
    if (gam == 1) {
        if (jv < n) {
            // breakdown: expand with a pseudo-random vector (JDaux.F:2741-2745)
            std::vector<double> r(n);
            for (int i = 0; i < n; ++i) {
                rng_ ^= rng_ >> 12;
                rng_ ^= rng_ << 25;
                rng_ ^= rng_ >> 27;
                r[i] = (double)((rng_ * 2685821657736338717ULL) >> 11) * 0x1.0p-53;
            }
            JDHIP_CHECK(hipMemcpyAsync(w, r.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                       ctx_->stream));
            JDHIP_CHECK(hipStreamSynchronize(ctx_->stream));
            nrnd = false;
            goto restart;
        }
    } else {
        fill(s, w, 0.0, n);
        y[jv] = 0.0;
    }
}

// SORTEV (JDaux.F:2464-2540): the permutation is computed on the host, the
// column cycles are applied on the device.
void JdEngine::sortev_dev(int n, int j, long voff, double *eigs, double *r, int it,
                          double sigma, long v2off) {
    if (j <= 1) return;
    std::vector<double> key(j);
    std::vector<int> iw(j);
    for (int i = 0; i < j; ++i) {
        if (it == 2 || it == -2) key[i] = -std::fabs(eigs[i] - sigma);
        else if (it == 4 || it == -4) key[i] = eigs[i];
        else key[i] = -eigs[i];
    }
    sortrx(j, key.data(), iw.data());
    if (it == -2) {
        int is = iw[j - 1];
        for (int i = j - 2; i >= 0; --i) iw[i + 1] = iw[i];
        iw[0] = is;
    }
    // IT = -4 moves the columns of a second matrix (B X) along (JDaux.F:2511-2532)
    const bool two = (it == -4) && v2off >= 0;
    ctx_->ensure_tmp((long)n * (two ? 2 : 1));
    hipStream_t s = ctx_->stream;
    const size_t bytes = sizeof(double) * n;
    auto col = [&](int k) { return X_ + voff + (long)(k - 1) * n; };
    auto col2 = [&](int k) { return X_ + v2off + (long)(k - 1) * n; };
    double *tmp2 = ctx_->d_tmp + n;
    int ic = 0;
    for (;;) {
        int is = 0;
        for (;;) {
            ++is;
            if (iw[is - 1] < 0) continue;
            if (iw[is - 1] == is) {
                iw[is - 1] = -iw[is - 1];
                if (++ic >= j) return;
                continue;
            }
            break;
        }
        int ir = iw[is - 1];
        JDHIP_CHECK(hipMemcpyAsync(ctx_->d_tmp, col(is), bytes, hipMemcpyDeviceToDevice, s));
        if (two) JDHIP_CHECK(hipMemcpyAsync(tmp2, col2(is), bytes, hipMemcpyDeviceToDevice, s));
        double r1 = 0.0, w1 = eigs[is - 1];
        if (it == 4 || it == -4) r1 = r[is - 1];
        for (;;) {
            JDHIP_CHECK(hipMemcpyAsync(col(is), col(ir), bytes, hipMemcpyDeviceToDevice, s));
            if (two)
                JDHIP_CHECK(hipMemcpyAsync(col2(is), col2(ir), bytes, hipMemcpyDeviceToDevice, s));
            if (it == 4 || it == -4) r[is - 1] = r[ir - 1];
            eigs[is - 1] = eigs[ir - 1];
            iw[is - 1] = -ir;
            ++ic;
            is = ir;
            ir = iw[is - 1];
            if (!(iw[ir - 1] > 0)) break;
        }
        JDHIP_CHECK(hipMemcpyAsync(col(is), ctx_->d_tmp, bytes, hipMemcpyDeviceToDevice, s));
        if (two) JDHIP_CHECK(hipMemcpyAsync(col2(is), tmp2, bytes, hipMemcpyDeviceToDevice, s));
        if (it == 4 || it == -4) r[is - 1] = r1;
        eigs[is - 1] = w1;
        iw[is - 1] = -ir;
        if (++ic >= j) return;
    }
}

// SOLPROJEP (JDaux.F:3271-3358) on host matrices: SYEV (|is| != 2), SYGV
// (harmonic, standard problem) or POSV + GEEV (harmonic, generalized).
int JdEngine::solprojep(int j, double *mm, double *nn, double *th, double *th2,
                        std::vector<double> &w, int is) {
    if ((int)w.size() < std::max(j, 2)) w.resize(std::max(j, 2), 0.0);
    int info;
    if (is == 2 || is == -2) {
        if (gep_) {
            // JDaux.F:3293-3300: MM := NN^-1 MM (POSV), its eigenvectors
            // (GEEV, the real parts of the eigenvalues kept) over MM
            info = chol_solve(j, nn, modspace, j, mm, modspace);
            for (int c = 0; c < j; ++c)
                std::memcpy(nn + (long)c * modspace, mm + (long)c * modspace,
                            sizeof(double) * j);
            if (info) return info;
            std::vector<double> wi(j);
            info = real_eig(j, nn, modspace, w.data(), wi.data(), mm, modspace);
        } else {
            info = sym_gen_eig(j, mm, modspace, nn, modspace, w.data());
        }
        if (info) return info;
        // SORTEV on the small eigenvector matrix (JDaux.F:3350-3352)
        if (j > 1) {
            std::vector<double> key(j);
            std::vector<int> iw(j);
            for (int i = 0; i < j; ++i) key[i] = -std::fabs(w[i]);
            sortrx(j, key.data(), iw.data());
            if (is == -2) {
                int t = iw[j - 1];
                for (int i = j - 2; i >= 0; --i) iw[i + 1] = iw[i];
                iw[0] = t;
            }
            std::vector<double> m2((size_t)modspace * j), w2(j);
            for (int c = 0; c < j; ++c) {
                w2[c] = w[iw[c] - 1];
                std::memcpy(&m2[(size_t)c * modspace], mm + (long)(iw[c] - 1) * modspace,
                            sizeof(double) * j);
            }
            for (int c = 0; c < j; ++c) {
                w[c] = w2[c];
                std::memcpy(mm + (long)c * modspace, &m2[(size_t)c * modspace],
                            sizeof(double) * j);
            }
        }
        *th2 = w[1];
        return 0;
    }
    info = sym_eig(j, mm, modspace, w.data());
    if (info) return info;
    if (is == 3) {
        for (int i = 0; i < j; ++i) nn[i] = w[i];
    } else {
        *th = w[0];
        *th2 = w[1];
    }
    return 0;
}

// SHRINK (JDaux.F:2418-2462): QR of the k x j Ritz block at MP(:, ucol_off..),
// [V W] := [V W] Q(:,1:j) on the device, MM (and NN) := Q^T . Q on the host.
int JdEngine::shrink(int n, int j, int k, int ucol_off, int iserch) {
    double *U = MP.data() + ucol_off;
    std::vector<double> tau(j), q((size_t)k * j);
    house_qr(k, j, U, modspace, tau.data());
    form_q(k, j, j, U, modspace, tau.data(), q.data());
    double *dq = ctx_->d_small + 4096;
    put(dq, q.data(), (long)k * j);
    ctx_->ensure_tmp((long)n * j);
    hipStream_t s = ctx_->stream;
    tall_gemm_mfma(s, X_ + vv, n, k, dq, k, j, ctx_->d_tmp, n, n);
    JDHIP_CHECK(hipMemcpyAsync(X_ + vv, ctx_->d_tmp, sizeof(double) * n * j,
                               hipMemcpyDeviceToDevice, s));
    tall_gemm_mfma(s, X_ + ww, n, k, dq, k, j, ctx_->d_tmp, n, n);
    JDHIP_CHECK(hipMemcpyAsync(X_ + ww, ctx_->d_tmp, sizeof(double) * n * j,
                               hipMemcpyDeviceToDevice, s));
    if (gep_) {  // BV := BV Q as well (JDaux.F:2439)
        tall_gemm_mfma(s, X_ + bv, n, k, dq, k, j, ctx_->d_tmp, n, n);
        JDHIP_CHECK(hipMemcpyAsync(X_ + bv, ctx_->d_tmp, sizeof(double) * n * j,
                                   hipMemcpyDeviceToDevice, s));
    }
    const int ld = modspace;
    if (iserch != 2 || !gep_)  // MM is full (not symmetric) for the harmonic GEP
        for (int a = 1; a < k; ++a)
            for (int b = 0; b < a; ++b) MM[a + b * ld] = MM[b + a * ld];
    apply_q_right(k, k, j, U, ld, tau.data(), MM.data(), ld);
    apply_qt_left(k, k, j, U, ld, tau.data(), MM.data(), ld);
    if (iserch == 2) {
        for (int a = 1; a < k; ++a)
            for (int b = 0; b < a; ++b) NN[a + b * ld] = NN[b + a * ld];
        apply_q_right(k, k, j, U, ld, tau.data(), NN.data(), ld);
        apply_qt_left(k, k, j, U, ld, tau.data(), NN.data(), ld);
    }
    return 0;
}

// Initial projected matrices (JDaux.F:1646-1660).
void JdEngine::gram_initial(int n) {
    hipStream_t s = ctx_->stream;
    const int ld = modspace;
    std::vector<double> C((size_t)J * J);
    if (iserch == 2) {
        // W := W - sigma V over the J initial columns (contiguous)
        // (sigma folded into axpyalpha by the caller)
        axpy(s, X_ + ww, axpyalpha, X_ + bv, (long)n * J);
        // NN = W^T W (upper triangle), one MFMA Gram product
        gram(s, n, J, J, X_ + ww, n, X_ + ww, n, ctx_->d_gout, J, ctx_->d_gram);
        fetch(C.data(), ctx_->d_gout, (long)J * J);
        for (int c = 0; c < J; ++c)
            for (int r = 0; r <= c; ++r) NN[r + c * ld] = C[r + (size_t)c * J];
    }
    if (iserch == 2 && gep_) {
        // MM = W^T (B V), not symmetric (JDaux.F:1653-1655)
        gram(s, n, J, J, X_ + ww, n, X_ + bv, n, ctx_->d_gout, J, ctx_->d_gram);
        fetch(C.data(), ctx_->d_gout, (long)J * J);
        for (int c = 0; c < J; ++c)
            for (int r = 0; r < J; ++r) MM[r + c * ld] = C[r + (size_t)c * J];
        return;
    }
    // MM = (V^T W + W^T V) / 2 from one Gram product C = V^T W
    gram_sym(n, X_ + vv, X_ + ww, J, C.data());
    for (auto &v : MM) v *= 0.5;
}

// MM(r, c) = V(:,r).W(:,c) + W(:,r).V(:,c) for r <= c < J (upper triangle),
// from one MFMA Gram product C = V^T W (the column-by-column GEMV 'C'
// updates of JDaux.F:1651-1660 and 2165-2200 in one launch).
void JdEngine::gram_sym(int n, const double *V, const double *W, int J, double *C) {
    hipStream_t s = ctx_->stream;
    const int ld = modspace;
    gram(s, n, J, J, V, n, W, n, ctx_->d_gout, J, ctx_->d_gram);
    fetch(C, ctx_->d_gout, (long)J * J);
    for (int c = 0; c < J; ++c)
        for (int r = 0; r <= c; ++r)
            MM[r + c * ld] = C[r + (size_t)c * J] + C[c + (size_t)r * J];
}

// STOPTEST (JDaux.F:2542-2672), standard problem; xn2 = ||x||^2 (device).
void JdEngine::stoptest(int n, double tol, double *rk, double alprho_, double gk_,
                        double *theta_, double eta_, int *info, double trs_, double xn2) {
    (void)n;
    if (*info == 0) goto L10;
    st_istag = 0;
    st_istt = 0;
    st_g0 = *rk;
    st_gkm1 = *rk;
    st_gtrs = trs_ * st_g0;
    if (st_gtrs < tol) st_gtrs = tol;
    st_gs = st_gtrs;
    st_gkmin = st_g0;
    st_bk = *theta_ - eta_;
    st_un2 = -1.0;
L10:
    st_bk = st_bk - alprho_;
    st_beta = std::fabs(st_bk);
    if (gk_ < st_gkmin) {
        st_gkmin = gk_;
        st_istag = 0;
    }
    if (gk_ > st_gkm1) {
        ++st_istag;
        ++st_istt;
    }
    if (st_istag >= 25 || st_istt >= 100) goto L40;
    if (st_gkmin >= st_gtrs) goto L20;
    st_toupdate = true;
    if (gk_ >= st_gs) goto L15;
L13:
    st_alp = 1.0;
    st_tn = std::sqrt(xn2);
    st_un2 = 1.0 + st_tn * st_tn;
    st_toupdate = false;
    st_gs = (gk_ < 0.1 * st_g0) ? tol : 0.1 * st_g0;
    if (st_gs < tol) st_gs = tol;
L15:
    *rk = std::sqrt(gk_ * gk_ * st_un2 + st_tn * st_tn * st_beta * st_beta) / st_un2;
    if (*rk < tol && st_toupdate) goto L13;
    if (*rk < tol) goto L40;
    if (gk_ * std::sqrt(st_un2) < 15 * st_tn * st_beta &&
        st_beta * st_tn > 0.5 * tol * st_un2 * st_alp)
        goto L40;
    if (st_istag >= 3) goto L40;
    if (st_istag >= 1 && st_beta * st_tn > 0.5 * tol * st_un2 * st_alp && gk_ < st_gtrs)
        goto L40;
L20:
    *info = 0;
    st_gkm1 = gk_;
    return;
L40:
    *info = std::max(*info, 1);
    *theta_ = st_bk;
}

// STOPTEST with the terms of the generalized problem (JDaux.F:2590-2668):
// the residual norm is scaled by ||B u||, the Rayleigh-quotient estimate
// carries the B-corrections pcad (accumulated BU.Q alpha) and pcr (BU.r at
// the start), and the norm of the correction is measured through BT = B x.
void JdEngine::stoptest_gep(int n, double tol, double *rk, double *alprho_, double gk_,
                            double *theta_, double eta_, int *info, double trs_, double pcad,
                            double pcr) {
    if (*info == 0) goto L10;
    st_istag = 0;
    st_istt = 0;
    st_g0 = *rk;
    st_gkm1 = *rk;
    st_gtrs = trs_ * st_g0;
    if (st_gtrs < tol) st_gtrs = tol;
    st_gs = st_gtrs;
    st_gkmin = st_g0;
    st_bk = *theta_ - eta_;
    st_bkgep = *theta_ - eta_ + pcr;
    st_un2 = -1.0;
L10:
    st_bk = st_bk - *alprho_;
    st_bkgep = st_bkgep - pcad;
    st_beta = std::fabs(st_bkgep);
    gk_ = gk_ / nbu;
    if (gk_ < st_gkmin) {
        st_gkmin = gk_;
        st_istag = 0;
    }
    if (gk_ > st_gkm1) {
        ++st_istag;
        ++st_istt;
    }
    if (st_istag >= 25 || st_istt >= 100) goto L40;
    if (st_gkmin >= st_gtrs) goto L20;
    st_toupdate = true;
    if (gk_ >= st_gs) goto L15;
L13 : {
    const double tnb = dnorm(X_ + bt, n) / nbu;
    const double bubt = -ddot(X_ + bu, X_ + bt, n) / (nbu * nbu);
    st_alp = std::fabs(1.0 + bubt);
    st_tn = std::sqrt((tnb + bubt) * (tnb - bubt)) / st_alp;
}
    st_un2 = 1.0 + st_tn * st_tn;
    st_toupdate = false;
    st_gs = (gk_ < 0.1 * st_g0) ? tol : 0.1 * st_g0;
    if (st_gs < tol) st_gs = tol;
L15:
    if (st_beta < gk_ * st_tn)
        *rk = std::sqrt((gk_ * gk_ + st_beta * st_beta) / st_un2) / st_alp;
    else
        *rk = (gk_ + st_beta * st_tn) / (st_alp * st_un2);
    if (*rk < tol && st_toupdate) goto L13;
    if (*rk < tol) goto L40;
    if (gk_ * std::sqrt(st_un2) < 15 * st_tn * st_beta &&
        st_beta * st_tn > 0.5 * tol * st_un2 * st_alp)
        goto L40;
    if (st_istag >= 3) goto L40;
    if (st_istag >= 1 && st_beta * st_tn > 0.5 * tol * st_un2 * st_alp && gk_ < st_gtrs)
        goto L40;
L20:
    *info = 0;
    st_gkm1 = gk_;
    return;
L40:
    *info = std::max(*info, 1);
    *alprho_ = st_un2 * st_alp * st_alp;
    *theta_ = st_bk;
}

// ---------------------------------------------------------------------------
// Progress line on stderr at most every 30 s of a solve when JDHIP_VERBOSE
// is set (a solve of a large problem can run for minutes without returning).
void JdEngine::progress(int matvecs, int converged) {
    static const bool on = [] {
        const char *v = std::getenv("JDHIP_VERBOSE");
        return v && *v && *v != '0';
    }();
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    if (!progress_started_) {
        progress_started_ = true;
        progress_t0_ = progress_last_ = now;
        return;
    }
    if (std::chrono::duration<double>(now - progress_last_).count() < 30.0) return;
    progress_last_ = now;
    std::fprintf(stderr, "jdhip: %d MATVECs, %d eigenpairs converged, %.0f s\n", matvecs,
                 converged, std::chrono::duration<double>(now - progress_t0_).count());
}

void JdEngine::step(int n, double *eigs, double *res, double *x, long lx, int *neig,
                    double sigma, int isearch, int ninit, int madspace, int *iter, double tol,
                    int *ijob, long *ndx1, long *ndx2, long *ndx3, int iprint, int *info,
                    double *gap, bool gep) {
    (void)iprint;
    X_ = x;
    hipStream_t s = ctx_->stream;
    const RedBuf &rb = ctx_->rb;
    const size_t nb = sizeof(double) * (size_t)n;
    auto D = [&](long off) { return X_ + off; };

    if (*ijob == 0) {
        progress_started_ = false;
        gep_ = gep;
        goto L1;
    }
    switch (rlbl) {
    case 1: goto L5;
    case 2: goto L225;
    case 3: goto L37;
    case 4: goto L38;
    case 5: goto L255;
    case 6: goto L55;
    case 11: goto L15;
    case 12: goto L45;
    case 13: goto LG38;
    case 14: goto LG37;
    default: break;
    }
    *info = -15;
    goto L190;

L1:
    flagcg = true;
    neweig = true;
    rcum = 0.0;
    rinitm = 1.0;
    inner_iters = 0;
    nsync = 0;
    *info = 0;
    if (n < 2) *info = -1;
    else if (*neig <= 0) *info = -5;
    else if (madspace <= 2) *info = -9;
    else if ((long)n * (2 * madspace + *neig + 4) + 3L * madspace * madspace +
                 std::max(madspace * madspace, *neig) > lx)
        *info = -10;
    else if (gep_ && (long)n * (3 * madspace + 2 * *neig + 4) + 3L * madspace * madspace +
                         std::max(madspace * madspace, *neig) > lx)
        *info = -10;  // JDaux.F:1512-1515
    else if (gep_ && isearch >= 2 &&
             (long)n * (3 * madspace + 2 * *neig) + 6L * madspace + 3L * madspace * madspace +
                     std::max(madspace * madspace, *neig) > lx)
        *info = -10;  // JDaux.F:1516-1519
    else if (*iter <= 0) *info = -11;
    else if (tol <= 0.0) *info = -12;
    else if (std::min(*neig, n) + std::min(madspace, n) > kMaxCols)
        *info = -9;  // tall-skinny kernels hold at most kMaxCols columns
    if (*info != 0) goto L190;
    modspace = std::min(madspace, n);
    nwant = std::min(*neig, n);
    mvmax = *iter;
    iserch = std::min(isearch, 2);  // JDaux.F:1530-1531 (second assignment wins)
    {
        const long N = n, MS = modspace;
        xx = 0;
        vv = xx + (long)nwant * N;
        ww = vv + MS * N;
        long mm_off;
        if (gep_) {  // JDaux.F:1537-1544
            bx = ww + MS * N;
            bv = bx + (long)nwant * N;
            mm_off = bv + MS * N;
        } else {
            bx = xx;
            bv = vv;
            mm_off = ww + MS * N;
        }
        // MM, NN, MP, NP live on the host; WCG keeps the reference offset
        wcg = mm_off + 3 * MS * MS + std::max(MS * MS, (long)nwant);
        xcg2 = ww + (MS - 1) * N;
        bt = bv + (MS - 1) * N;
        reig = wcg;
        uu = xx + (long)(nwant - 1) * N;
        bu = bx + (long)(nwant - 1) * N;
        kxp = vv;
        kbxp = bv;
    }
    MM.assign((size_t)modspace * modspace, 0.0);
    NN.assign((size_t)modspace * modspace, 0.0);
    MP.assign((size_t)modspace * modspace, 0.0);
    NP.assign((size_t)modspace * modspace, 0.0);
    WL.assign(3 * modspace + 4, 0.0);
    WL2.assign(modspace + 2, 0.0);
    WL3.assign(modspace + 2, 0.0);
    REIGW.assign(std::max(modspace, 2) + 2, 0.0);
    ycoef.assign(2 * (nwant + modspace) + 4, 0.0);
    epsmsr = std::sqrt(kEps);
    *neig = 0;
    *iter = 0;
    iterp = 0;
    istep = 0;
    eta1 = kInf;
    eta0 = sigma;
    delta1 = kInf;
    if (iserch == 0) eta0 = kInf;
    nreig = kInf;
    resid = kInf;
    rndind = 0.0;
    midspace = (modspace + 1) / 2;
    if (ninit > midspace) {
        initflg = false;
        for (int q = 1; q < ninit; ++q)
            if (eigs[q] != eigs[0]) initflg = true;
        if (initflg) {
            sortev_dev(n, ninit, xx, eigs, res, iserch, sigma);
        } else {
            ctx_->ensure_tmp(n);
            for (int q = 1; q <= ninit / 2; ++q) {
                int r = ninit - q + 1;
                JDHIP_CHECK(hipMemcpyAsync(ctx_->d_tmp, D((long)(r - 1) * n), nb,
                                           hipMemcpyDeviceToDevice, s));
                JDHIP_CHECK(hipMemcpyAsync(D((long)(r - 1) * n), D((long)(q - 1) * n), nb,
                                           hipMemcpyDeviceToDevice, s));
                JDHIP_CHECK(hipMemcpyAsync(D((long)(q - 1) * n), ctx_->d_tmp, nb,
                                           hipMemcpyDeviceToDevice, s));
            }
        }
    }
    initflg = false;

L2:
    if (ninit <= *neig) {
        J = 1;
        fill(s, D(vv), 1.0, n);
    } else {
        J = std::min(ninit - *neig, midspace);
        IJ = ninit - *neig - J;
        if (IJ > 0) initflg = true;
        JDHIP_CHECK(hipMemcpyAsync(D(vv), D((long)n * IJ), nb * J, hipMemcpyDeviceToDevice, s));
    }
    IJ = 0;
L4:
    ++IJ;
    repgs2(n, kxp, kbxp, *neig + IJ - 1, vv + (long)(IJ - 1) * n, ycoef, 1);
    if (!gep_) {
        if (*neig + IJ - 1 > 0) sclr1 = ycoef[*neig + IJ - 1];
        else sclr1 = dnorm(D(vv), n);
        scal(s, D(vv + (long)(IJ - 1) * n), 1.0 / sclr1, n);
    }
    *ndx1 = vv + (long)(IJ - 1) * n;
    *ndx2 = ww + (long)(IJ - 1) * n;
    *ndx3 = bv + (long)(IJ - 1) * n;
    rlbl = 1;
    if (internal_mv()) {
        do_matvec(D(*ndx1), D(*ndx2));
        if (gep_) do_bmatvec(D(*ndx1), D(*ndx3));
        goto L5;
    }
    *ijob = 1;
    return;
L5:
    ++*iter;
    if (gep_) {  // B-normalise v, B v and A v (JDaux.F:1637-1644)
        const long c = (long)(IJ - 1) * n;
        sclr1 = ddot(D(vv + c), D(bv + c), n);
        if (sclr1 <= 0.0) goto L196;
        sclr1 = std::sqrt(sclr1);
        scal(s, D(bv + c), 1.0 / sclr1, n);
        scal(s, D(vv + c), 1.0 / sclr1, n);
        scal(s, D(ww + c), 1.0 / sclr1, n);
    }
    if (IJ < J) goto L4;
    axpyalpha = -sigma;
    gram_initial(n);
    uknown = (J <= 1);

L10:  // extraction
    MP = MM;
    if (iserch == 2) NP = NN;
    *info = solprojep(J, MP.data(), NP.data(), &theta, &theta2, REIGW, iserch);
    if (*info != 0) goto L192;
    if (!uknown) {
        if (iserch == 2) {
            sclr1 = small_nrm2(J, YY());
            for (int c = 0; c < J; ++c) YY()[c] *= 1.0 / sclr1;
        }
        put(ctx_->slot(3), YY(), J);
        gemv_n(s, rb, D(uu), D(vv), n, J, ctx_->slot(3), 1.0, 0.0, n, nullptr);
        gemv_n(s, rb, D(reig), D(ww), n, J, ctx_->slot(3), 1.0, 0.0, n, nullptr);
        if (gep_) gemv_n(s, rb, D(bu), D(bv), n, J, ctx_->slot(3), 1.0, 0.0, n, nullptr);
    } else {
        YY()[0] = 1.0;
        for (int c = 1; c < J; ++c) YY()[c] = 0.0;
        JDHIP_CHECK(hipMemcpyAsync(D(uu), D(vv), nb, hipMemcpyDeviceToDevice, s));
        JDHIP_CHECK(hipMemcpyAsync(D(reig), D(ww), nb, hipMemcpyDeviceToDevice, s));
        if (gep_) JDHIP_CHECK(hipMemcpyAsync(D(bu), D(bv), nb, hipMemcpyDeviceToDevice, s));
    }
    if (gep_) nbu = dnorm(D(bu), n);
    if (iserch == 2) {
        if (gep_) {  // MM is not symmetric here (JDaux.F:1699-1702)
            gemv_host(J, J, MM.data(), modspace, YY(), WL.data(), false);
            thmt = small_dot(J, WL.data(), YY()) / (nbu * nbu);
        } else {
            symv_upper(J, MM.data(), modspace, YY(), WL.data());
            thmt = small_dot(J, WL.data(), YY());
        }
        theta = thmt + sigma;
    } else {
        thmt = theta;
        eta0 = std::min(eta0, theta);
    }
    axpyalpha = -thmt;
    {
        double *d = ctx_->slot(2);
        axpy_nrm2(s, rb, D(reig), axpyalpha, D(bu), n, d);
        double v;
        fetch(&v, d, 1);
        nreigp = nreig;
        nreig = std::sqrt(v);
        if (gep_) nreig = nreig / nbu;
    }
    if (neweig) {
        rinit = nreig;
        neweig = false;
    }
    if (std::fabs(nreig / nreigp - 1.0) < 0.01) ++istag;
    else istag = 0;
    if (nreig >= tol && istag <= 2) goto L20;
    if (nreig >= tol && rndind <= 1.0) goto L20;
    if (nreig >= tol && resid > 0.1 * nreig) goto L20;
    // convergence detected
    repgs2(n, kxp, kbxp, *neig, uu, ycoef, 0);
    if (*neig > 0 && ycoef[*neig] == 0.0) goto L195;
    if (gep_) {  // B u of the orthogonalised u, then B-normalise (labels 14-15)
        *ndx1 = uu;
        *ndx3 = bu;
        rlbl = 11;
        if (internal_mv()) {
            do_bmatvec(D(*ndx1), D(*ndx3));
            goto L15;
        }
        *ijob = 4;
        return;
    }
    sclr1 = (*neig > 0) ? ycoef[*neig] : dnorm(D(uu), n);
    goto L16;
L15:
    sclr1 = ddot(D(uu), D(bu), n);
    if (sclr1 <= 0.0) goto L196;
    sclr1 = std::sqrt(sclr1);
    scal(s, D(bu), 1.0 / sclr1, n);
L16:
    scal(s, D(uu), 1.0 / sclr1, n);

L18:  // lock the converged pair
    ++*neig;
    kxp = uu;
    kbxp = bu;
    uu -= n;
    bu -= n;
    eigs[nwant - *neig] = theta;
    res[nwant - *neig] = nreig;
    rinitm *= rinit;
    iterp = *iter;
    *gap = theta2 - theta;
    if (*neig >= nwant) goto L200;
    delta1 = kInf;
    nreig = kInf;
    resid = kInf;
    rndind = 0.0;
    if (iserch != 2) {
        eta1 = std::min(eta0, eta1);
        // JDaux.F:1794-1796 reads EIGS(1..NEIG) (caller's input values for
        // not-yet-filled slots) although converged values sit at the top.
        for (int q = 0; q < *neig; ++q) eta0 = std::max(eta0, eigs[q]);
    }
    if (J == 1 || initflg) goto L2;
    K = J;
    J = std::min(midspace + 1, J - 1);
    J = std::min(J, modspace - 1);
    if (iserch == 2) {
        // orthogonalise MP columns 2..J+1 against YY (small, host)
        for (IJ = 2; IJ <= J + 1; ++IJ) {
            double *w = MP.data() + (long)modspace * (IJ - 1);
            const double *v = YY();
            bool nrnd = true;
            double nrn = small_nrm2(K, w), nro, nr = kEps * nrn;
            for (;;) {
                nro = nrn;
                double c = small_dot(K, v, w);
                for (int t = 0; t < K; ++t) w[t] -= c * v[t];
                nrn = small_nrm2(K, w);
                if (nrn <= nr) break;
                if (!(nrn < 0.5 * nro)) break;
            }
            (void)nrnd;
        }
    }
    *info = shrink(n, J, K, modspace /* YY2 */, iserch);
    if (*info != 0) goto L193;
    uknown = true;
    neweig = true;
    goto L10;

L20:  // expansion
    trs = std::sqrt(1.0 / 10);
    if (istag > 0) trs = 1.0 / 10;
    ++istep;
    progress(*iter, I - 1);
    if (*iter >= mvmax - *neig - 1) goto L197;
    if (J >= modspace) {
        K = J;
        J = midspace;
        *info = shrink(n, J, K, 0 /* MP */, iserch);
        if (*info != 0) goto L193;
        double am[2], vjj;
        argmax_abs(s, rb, D(uu), n, ctx_->slot(2));
        fetch(am, ctx_->slot(2), 2);
        JJ = (int)am[0];
        fetch(&vjj, D(vv + JJ), 1);
        axpyalpha = am[1] / vjj;
        YY()[0] = axpyalpha / std::fabs(axpyalpha);
        for (int c = 1; c < J; ++c) YY()[c] = 0.0;
    }
    eta = eta0;
    if (J > 1) {
        deltao = delta1;
        if (iserch == 2) {
            if (gep_) {
                theta2 = 1.0 / theta2;
            } else {
                sclr1 = small_nrm2(J, YY2());
                for (int c = 0; c < J; ++c) YY2()[c] *= 1.0 / sclr1;
                symv_upper(J, MM.data(), modspace, YY2(), WL.data());
                theta2 = small_dot(J, YY2(), WL.data());
            }
            delta1 = std::fabs(theta2) - std::fabs(theta - sigma);
        } else {
            delta1 = theta2 - theta;
        }
        if (nreig < std::fabs(delta1) && std::fabs(delta1 / deltao - 1.0) < 0.1) eta = theta;
    }
    repgs2(n, bu, uu, *neig + 1, reig, ycoef, -1);
    rndind = std::fabs(ycoef[*neig] / (nreig * epsmsr));
    resid = nreig;
    xcg = vv + (long)J * n;
    ftflg = true;
    itcg = mvmax - *iter - *neig - 1;
    ifo = (iserch != 2) ? 1 : 0;
    gk = resid;
    rcurr = nreig;
    if (*neig <= 0) {
        shift1 = theta;
        *gap = (J > 1) ? theta2 - theta : 0.0;
    }
    if (gep_) {
        axpybeta = ddot(D(bu), D(reig), n) / (nbu * nbu);
        goto LG0;
    }

    // ---- LSOLREVCOM init (JDaux.F:3086-3116) ----
    {
        ls_posdef = (ifo == 1);
        ifo = 0;
        ls_maxit = itcg;
        ls_iter = 0;
        f32_ = use_f32();
        if (f32_) {
            // mixed path: the correction equation is solved on FP32 copies
            // of r and of the projector basis [u, X]
            ensure_f32(n, nwant + 1);
            convert(s, fbuf_, D(wcg), n);
            convert(s, fbuf_ + 6L * n, D(uu), (long)(*neig + 1) * n);
            if (ls_posdef)
                JDHIP_CHECK(hipMemsetAsync(fbuf_ + 4L * n, 0, sizeof(float) * n, s));
        } else if (ls_posdef) {
            fill(s, D(xcg), 0.0, n);
        }
        InnerScalars &h = *ctx_->h_sc;
        h.posdef = ls_posdef ? 1 : 0;
        h.flag = 0;
        h.ggk2 = gk * gk;
        h.eta = eta;
        JDHIP_CHECK(hipMemcpyAsync(ctx_->d_sc, &h, sizeof(InnerScalars), hipMemcpyHostToDevice,
                                   s));
        ctx_->sync();
    }
L10i:  // new inner iteration: preconditioner solve Z = P^-1 R
    ++ls_iter;
    ++inner_iters;
    if (f32_) {
        float *rf = fbuf_, *zf = fbuf_ + n, *pf = fbuf_ + 2L * n, *qf = fbuf_ + 3L * n;
        float *xf = fbuf_ + 4L * n, *x2f = fbuf_ + 5L * n, *Qf = fbuf_ + 6L * n;
        ops_->precond_f(*this, zf, rf);
        flagcg = true;
        cg_project_rho<float>(s, rb, zf, Qf, *neig + 1, rf, n, ctx_->d_sc, ctx_->slot(0));
        cg_direction<float>(s, zf, pf, n, ctx_->d_sc, ls_iter);
        ops_->matvec_f(pf, qf, eta, Qf, &ctx_->d_sc->alprho);
        cg_project_alpha<float>(s, rb, qf, Qf, pf, n, ctx_->d_sc);
        if (ls_posdef) {
            cg_update_posdef<float>(s, rb, rf, qf, xf, pf, n, ctx_->d_sc);
        } else {
            qmr_update_r<float>(s, rb, rf, qf, n, ctx_->d_sc);
            qmr_update_x<float>(s, rb, xf, x2f, pf, n, ctx_->d_sc, ls_iter);
        }
        goto Lsync;
    }
    *ndx1 = wcg + n;  // Z
    *ndx2 = wcg;      // R
    rlbl = 4;
    if (internal_pc()) {
        ops_->precond(*this, D(*ndx1), D(*ndx2));
        goto L38;
    }
    *ijob = 2;
    return;
L38:  // project Z against [u, X] ; rho = r.z ; direction
    flagcg = true;
    cg_project_rho(s, rb, D(wcg + n), D(uu), *neig + 1, D(wcg), n, ctx_->d_sc, ctx_->slot(0));
    cg_direction(s, D(wcg + n), D(wcg + 2L * n), n, ctx_->d_sc, ls_iter);
    *ndx1 = wcg + 2L * n;  // P
    *ndx2 = wcg + 3L * n;  // Q
    rlbl = 3;
    if (internal_mv()) {
        ops_->matvec(D(*ndx1), D(*ndx2), nullptr, eta, D(bu), &ctx_->d_sc->alprho);
        goto L37b;
    }
    // external matvec: the breakdown test must be known before returning
    JDHIP_CHECK(hipMemcpyAsync(ctx_->h_sc, ctx_->d_sc, sizeof(InnerScalars),
                               hipMemcpyDeviceToHost, s));
    ctx_->sync();
    ++nsync;
    if (ctx_->h_sc->flag == 1) goto L40i;
    *ijob = 1;
    return;
L37:  // after an external matvec: Q -= eta P ; alprho = u.Q
    axpy(s, D(wcg + 3L * n), -eta, D(wcg + 2L * n), n);
    multidot(s, rb, D(bu), n, 1, D(wcg + 3L * n), n, &ctx_->d_sc->alprho);
    ++*iter;
    goto L4i;
L37b:
L4i:
    cg_project_alpha(s, rb, D(wcg + 3L * n), D(bu), D(wcg + 2L * n), n, ctx_->d_sc);
    if (ls_posdef) {
        cg_update_posdef(s, rb, D(wcg), D(wcg + 3L * n), D(xcg), D(wcg + 2L * n), n,
                         ctx_->d_sc);
    } else {
        qmr_update_r(s, rb, D(wcg), D(wcg + 3L * n), n, ctx_->d_sc);
        qmr_update_x(s, rb, D(xcg), D(xcg2), D(wcg + 2L * n), n, ctx_->d_sc, ls_iter);
    }
Lsync:
    JDHIP_CHECK(hipMemcpyAsync(ctx_->h_sc, ctx_->d_sc, sizeof(InnerScalars),
                               hipMemcpyDeviceToHost, s));
    ctx_->sync();
    ++nsync;
    {
        const InnerScalars &h = *ctx_->h_sc;
        if (internal_mv() && h.flag != 1) ++*iter;
        if (h.flag == 1) goto L40i;
        if (h.flag == 2) {
            alpha = h.pq;
            goto L40i;
        }
        alpha = h.alpha;
        gk = std::sqrt(h.gk2);
        alprho = h.alprho;
        // L33: stopping test
        if (ftflg) ifo = -1;
        if (ftflg && iserch == 2) ifo = -2;
        axpyalpha = axpyalpha * alpha;
        alprho = alprho * alpha;
        stoptest(n, tol, &resid, alprho, gk, &theta, eta, &ifo, trs, h.xn2);
        ftflg = false;
        // L5 of LSOLREVCOM
        if (ifo == 1) {
            ifo = 0;
            goto L34;
        }
        if (ls_iter == ls_maxit) goto L40i;
        goto L10i;
    }
L40i:  // inner iteration fails (JDaux.F:3255-3266)
    ifo = 1;
    if (alpha <= 0.0) ifo = 2;
    if (ls_iter == 1) {
        if (f32_)
            JDHIP_CHECK(hipMemcpyAsync(fbuf_ + 4L * n, fbuf_ + 2L * n, sizeof(float) * n,
                                       hipMemcpyDeviceToDevice, s));
        else
            JDHIP_CHECK(
                hipMemcpyAsync(D(xcg), D(wcg + 2L * n), nb, hipMemcpyDeviceToDevice, s));
    }
    goto L34;
    // ---- inner solver of the generalized problem: LSOLREVCOM with the GEP
    // terms of JDRVCOM labels 30-38 (JDaux.F:1883-1934, 2989-3268), scalars on
    // the host.  Z holds B p after each matvec (NDX3 = Z) and BT accumulates
    // alpha B p.
LG0:
    ls_posdef = (ifo == 1);
    ifo = 0;
    ls_maxit = itcg;
    ls_iter = 0;
    f32_ = false;
    if (ls_posdef) fill(s, D(xcg), 0.0, n);
    fill(s, D(bt), 0.0, n);
    g_ggk2 = gk * gk;
LG10:
    ++ls_iter;
    ++inner_iters;
    *ndx1 = wcg + n;  // Z
    *ndx2 = wcg;      // R
    rlbl = 13;
    if (internal_pc()) {
        ops_->precond(*this, D(*ndx1), D(*ndx2));
        goto LG38;
    }
    *ijob = 2;
    return;
LG38:
    flagcg = true;
    repgs2(n, uu, bu, *neig + 1, wcg + n, ycoef, -1);  // z -= U (BU^T z)
    g_rho = ddot(D(wcg), D(wcg + n), n);
    if ((ls_posdef && g_rho <= 0.0) || g_rho == 0.0) goto LG40;
    if (ls_iter > 1) {
        axpy(s, D(wcg + n), g_rho / g_rho1, D(wcg + 2L * n), n);
        JDHIP_CHECK(hipMemcpyAsync(D(wcg + 2L * n), D(wcg + n), nb, hipMemcpyDeviceToDevice, s));
    } else {
        JDHIP_CHECK(hipMemcpyAsync(D(wcg + 2L * n), D(wcg + n), nb, hipMemcpyDeviceToDevice, s));
    }
    *ndx1 = wcg + 2L * n;  // P
    *ndx2 = wcg + 3L * n;  // Q = A P
    *ndx3 = wcg + n;       // Z = B P
    rlbl = 14;
    if (internal_mv()) {
        do_matvec(D(*ndx1), D(*ndx2));
        do_bmatvec(D(*ndx1), D(*ndx3));
        goto LG37;
    }
    *ijob = 1;
    return;
LG37:
    axpy(s, D(wcg + 3L * n), -eta, D(wcg + n), n);  // Q = (A - eta B) P
    axpyalpha = ddot(D(bu), D(wcg + 3L * n), n) / (nbu * nbu);
    repgs2(n, bu, uu, 1, wcg + 3L * n, ycoef, -1);  // Q -= BU (U^T Q)
    alprho = ycoef[0];
    ++*iter;
    {  // LSOLREVCOM label 4
        const double pq = ddot(D(wcg + 2L * n), D(wcg + 3L * n), n);
        alpha = pq;
        if ((ls_posdef && pq <= 0.0) || pq == 0.0) goto LG40;
        alpha = g_rho / pq;
    }
    {
        double *d = ctx_->slot(2), v;
        axpy_nrm2(s, rb, D(wcg), -alpha, D(wcg + 3L * n), n, d);  // R -= alpha Q
        fetch(&v, d, 1);
        gk = std::sqrt(v);
    }
    axpy(s, D(bt), alpha, D(wcg + n), n);  // BT += alpha B P
    if (ls_posdef) {
        axpy(s, D(xcg), alpha, D(wcg + 2L * n), n);
    } else {
        const double gk2 = gk * gk;
        g_ttk2m = g_ttk2;
        g_ttk2 = gk2 / g_ggk2;
        const double ck2 = 1.0 / (1.0 + g_ttk2);
        g_ggk2 = g_ggk2 * g_ttk2 * ck2;
        if (ls_iter > 1) {
            axpy(s, D(xcg2), alpha / g_ttk2m, D(wcg + 2L * n), n);
            scal(s, D(xcg2), g_ttk2m * ck2, n);
            axpy(s, D(xcg), 1.0, D(xcg2), n);
        } else {
            JDHIP_CHECK(
                hipMemcpyAsync(D(xcg2), D(wcg + 2L * n), nb, hipMemcpyDeviceToDevice, s));
            scal(s, D(xcg2), alpha * ck2, n);
            JDHIP_CHECK(hipMemcpyAsync(D(xcg), D(xcg2), nb, hipMemcpyDeviceToDevice, s));
        }
    }
    // L33: stopping test
    if (ftflg) ifo = -1;
    if (ftflg && iserch == 2) ifo = -2;
    axpyalpha = axpyalpha * alpha;
    alprho = alprho * alpha;
    stoptest_gep(n, tol, &resid, &alprho, gk, &theta, eta, &ifo, trs, axpyalpha, axpybeta);
    ftflg = false;
    if (ifo == 1) {
        ifo = 0;
        goto L34;
    }
    if (ls_iter == ls_maxit) goto LG40;
    g_rho1 = g_rho;
    goto LG10;
LG40:  // inner iteration fails
    ifo = 1;
    if (alpha <= 0.0) ifo = 2;
    if (ls_iter == 1)
        JDHIP_CHECK(hipMemcpyAsync(D(xcg), D(wcg + 2L * n), nb, hipMemcpyDeviceToDevice, s));
L34:
    f32_last_ = f32_;
    if (f32_) {
        convert(s, D(xcg), fbuf_ + 4L * n, n);  // correction back to FP64
        f32_ = false;
    }
    itcg = ls_iter;
    if (iserch != 2 && eta <= eta0 && *neig > 0 && ifo == 2) eta0 = std::min(eta1, eta0);
    conv = false;
    flagcg = false;
    rcum = rcum + std::log(nreig / resid);
    // The FP32 inner solver's recurrences cannot certify a residual below
    // TOL (their rounding floor lies above it): its correction always goes
    // through the outer step, whose FP64 residual decides convergence.
    if (resid >= tol || f32_last_) goto L50;
    // convergence reached during the inner solve
    axpy(s, D(uu), -1.0, D(xcg), n);
    repgs2(n, kxp, kbxp, *neig, uu, ycoef, 0);
    if (*neig > 0 && ycoef[*neig] == 0.0) goto L195;
    if (gep_) {  // B u, then B-normalise (labels 44-45)
        *ndx1 = uu;
        *ndx3 = bu;
        rlbl = 12;
        if (internal_mv()) {
            do_bmatvec(D(*ndx1), D(*ndx3));
            goto L45;
        }
        *ijob = 4;
        return;
    }
    sclr1 = (*neig > 0) ? ycoef[*neig] : dnorm(D(uu), n);
    goto L46;
L45:
    sclr1 = ddot(D(uu), D(bu), n);
    if (sclr1 <= 0.0) goto L196;
    sclr1 = std::sqrt(sclr1);
    scal(s, D(bu), 1.0 / sclr1, n);
L46:
    scal(s, D(uu), 1.0 / sclr1, n);
    theta = theta / sclr1 + eta;
    conv = true;
    nreig = resid;
    if (*neig >= nwant - 1) goto L18;
L50:  // expand the subspace with the correction
    repgs2(n, kxp, kbxp, *neig + J, xcg, ycoef, 1);
    for (int c = 0; c < J; ++c) YY()[c] += -1.0 * ycoef[*neig + c];
    if (!gep_) {
        sclr1 = ycoef[*neig + J];
        scal(s, D(xcg), 1.0 / sclr1, n);
    }
    *ndx1 = xcg;
    *ndx2 = ww + (long)J * n;
    *ndx3 = bv + (long)J * n;
    rlbl = 6;
    if (internal_mv()) {
        do_matvec(D(*ndx1), D(*ndx2));
        if (gep_) do_bmatvec(D(*ndx1), D(*ndx3));
        goto L55;
    }
    *ijob = 1;
    return;
L55:
    ++*iter;
    if (gep_) {  // B-normalise the new column (JDaux.F:2000-2008)
        const long c = (long)J * n;
        sclr1 = ddot(D(xcg), D(bv + c), n);
        if (sclr1 <= 0.0) goto L196;
        sclr1 = std::sqrt(sclr1);
        scal(s, D(xcg), 1.0 / sclr1, n);
        scal(s, D(ww + c), 1.0 / sclr1, n);
        scal(s, D(bv + c), 1.0 / sclr1, n);
    }
    YY()[J] = -sclr1;
    axpyalpha = -sigma;
    if (iserch == 2) axpy(s, D(*ndx2), axpyalpha, D(bv + (long)J * n), n);
    if (iserch == 2 && gep_) {
        // MM(0:J, J) = W^T (B v), MM(J, 0:J-1) = (B V)^T w (JDaux.F:2012-2016)
        double buf[3 * kMaxCols];
        multidot(s, rb, D(ww), n, J + 1, D(bv + (long)J * n), n, ctx_->slot(0));
        multidot(s, rb, D(ww), n, J + 1, D(*ndx2), n, ctx_->slot(1));
        if (J > 0) multidot(s, rb, D(bv), n, J, D(*ndx2), n, ctx_->slot(2));
        fetch(buf, ctx_->slot(0), 3 * kMaxCols);
        for (int r = 0; r <= J; ++r) MM[r + (long)J * modspace] = buf[r];
        for (int c = 0; c < J; ++c) MM[J + (long)c * modspace] = buf[2 * kMaxCols + c];
        for (int r = 0; r <= J; ++r) NN[r + (long)J * modspace] = buf[kMaxCols + r];
    } else {
        double buf[2 * kMaxCols];
        multidot(s, rb, D(vv), n, J + 1, D(*ndx2), n, ctx_->slot(0));
        if (iserch == 2) multidot(s, rb, D(ww), n, J + 1, D(*ndx2), n, ctx_->slot(1));
        fetch(buf, ctx_->slot(0), iserch == 2 ? 2 * kMaxCols : J + 1);
        for (int r = 0; r <= J; ++r) MM[r + (long)J * modspace] = buf[r];
        if (iserch == 2)
            for (int r = 0; r <= J; ++r) NN[r + (long)J * modspace] = buf[kMaxCols + r];
    }
    ++J;
    uknown = false;
    if (!conv) goto L10;
    // convergence detected: prepare for restart (JDaux.F:2040-2081)
    if (iserch != 2 || !gep_) {
        sclr1 = small_nrm2(J, YY());
        for (int c = 0; c < J; ++c) YY()[c] *= 1.0 / sclr1;
    }
    if (iserch == 2 && gep_) {
        symv_upper(J, NN.data(), modspace, YY(), WL2.data());
        sclr1 = small_dot(J, YY(), WL2.data());
        gemv_host(J, J, MM.data(), modspace, YY(), WL.data(), false);
        axpyalpha = small_dot(J, YY(), WL.data());
        thmt = axpyalpha / (alprho * nbu * nbu);
        axpyalpha = -axpyalpha / sclr1;
        for (int c = 0; c < J; ++c) WL[c] += axpyalpha * WL2[c];
        gemv_host(J, J, MM.data(), modspace, YY(), WL3.data(), true);
        MP = MM;
        ger_host(J, J, -1.0 / sclr1, WL.data(), WL2.data(), MP.data(), modspace);
        ger_host(J, J, -1.0 / sclr1, WL2.data(), WL3.data(), MP.data(), modspace);
        theta = thmt + sigma;
        NP = NN;
    } else if (iserch == 2) {
        symv_upper(J, NN.data(), modspace, YY(), WL2.data());
        sclr1 = small_dot(J, YY(), WL2.data());
        symv_upper(J, MM.data(), modspace, YY(), WL.data());
        thmt = small_dot(J, YY(), WL.data());
        axpyalpha = -thmt / (2 * sclr1);
        for (int c = 0; c < J; ++c) WL[c] += axpyalpha * WL2[c];
        axpyalpha = -1.0 / sclr1;
        MP = MM;
        syr2_upper(J, axpyalpha, WL.data(), WL2.data(), MP.data(), modspace);
        theta = thmt + sigma;
        NP = NN;
    } else {
        symv_upper(J, MM.data(), modspace, YY(), WL.data());
        theta = small_dot(J, YY(), WL.data());
        eta0 = std::min(eta0, theta);
        sclr1 = small_nrm2(J, MP.data(), modspace + 1);
        if (theta < 0) sclr1 = sclr1 - 2 * theta;
        axpyalpha = sclr1;
        for (int c = 0; c < J; ++c) WL[c] += axpyalpha * YY()[c];
        for (int c = 0; c < J; ++c) WL2[c] = YY()[c];
        MP = MM;
        syr2_upper(J, -1.0, WL2.data(), WL.data(), MP.data(), modspace);
    }
    *info = solprojep(J, MP.data(), NP.data(), &sclr1, &sclr2, REIGW, -iserch);
    if (*info != 0) goto L192;
    if (iserch == 2) {
        sclr1 = small_nrm2(J, YY());
        for (int c = 0; c < J; ++c) YY()[c] *= 1.0 / sclr1;
    }
    goto L18;

    // ---- error exits (JDaux.F:2090-2160) ----
L190:
    goto L301;
L192:
    if (*info > 0) *info = -*info;
    goto L198;
L193:
    if (*info > 0) *info = -*info;
    goto L198;
L195:
    *info = -54320;
    goto L198;
L196:  // the mass matrix B is not positive definite
    *info = -54322;
    goto L198;
L197:
    *info = *neig + 1;
    rinitm *= rinit;
L198:
    if (*neig > 0) JDHIP_CHECK(hipMemcpyAsync(D(vv), D(uu), nb, hipMemcpyDeviceToDevice, s));
    if (*neig > 0 && gep_)
        JDHIP_CHECK(hipMemcpyAsync(D(bv), D(bu), nb, hipMemcpyDeviceToDevice, s));
L200:  // normal termination
    if (*neig < nwant) {
        rinit = std::pow(rinitm, 1.0 / (double)(*neig + 1));
        if (*neig > 0) {
            JJ = std::min(*neig, nwant - *neig);
            JDHIP_CHECK(hipMemcpyAsync(D(xx), D((long)n * (nwant - JJ) + xx), nb * JJ,
                                       hipMemcpyDeviceToDevice, s));
            std::memmove(eigs, eigs + nwant - JJ, sizeof(double) * JJ);
            std::memmove(res, res + nwant - JJ, sizeof(double) * JJ);
            JDHIP_CHECK(hipMemcpyAsync(D((long)n * (*neig) + xx), D(vv), nb,
                                       hipMemcpyDeviceToDevice, s));
            if (gep_) {
                JDHIP_CHECK(hipMemcpyAsync(D(bx), D((long)n * (nwant - JJ) + bx), nb * JJ,
                                           hipMemcpyDeviceToDevice, s));
                JDHIP_CHECK(hipMemcpyAsync(D((long)n * (*neig) + bx), D(bv), nb,
                                           hipMemcpyDeviceToDevice, s));
            }
            eigs[*neig] = theta;
            res[*neig] = nreig;
            theta2 = theta;
        } else {
            JDHIP_CHECK(hipMemcpyAsync(D(xx), D(uu), nb, hipMemcpyDeviceToDevice, s));
            if (gep_) JDHIP_CHECK(hipMemcpyAsync(D(bx), D(bu), nb, hipMemcpyDeviceToDevice, s));
            eigs[0] = theta;
            res[0] = nreig;
            *gap = 0.0;
            goto L300;
        }
    } else {
        rinit = std::pow(rinitm, 1.0 / (double)(*neig));
    }
    sortev_dev(n, *neig, xx, eigs, res, gep_ ? -4 : 4, 0.0, bx);
    if (iserch >= 2) {
        for (itarg = 1; itarg <= *neig; ++itarg)
            if (eigs[itarg - 1] > sigma) break;
    } else {
        itarg = 0;
    }
    if (iserch == 2)
        *gap = std::fabs(theta2 - sigma) -
               std::max(std::fabs(eigs[*neig - 1] - sigma), std::fabs(eigs[0] - sigma));
    else
        *gap = theta2 - eigs[*neig - 1];
    gg = (iserch == 2) ? eigs[0] + eigs[*neig - 1] - 2 * sigma : kInf;
    if (gg > 0.0) {
        II = 1;
        JJ = *neig - 1;
    } else {
        II = 2;
        JJ = *neig;
    }
    if (*neig == 1) {
        II = 0;
        JJ = 0;
    }
    I = 1;
    K = 0;
L210:
    if (I < II || I >= JJ) goto L250;
    gl = (I == 1) ? gg - res[*neig - 1] : eigs[I - 1] - eigs[I - 2] - res[I - 2];
    if (gl < 10 * (eigs[I] - eigs[I - 1])) goto L250;
    J = std::min(I + modspace, JJ + 1);
L215:
    --J;
    if (J == I) goto L250;
    gr = (J == *neig) ? -gg - res[0] : eigs[J] - eigs[J - 1] - res[J];
    if (10 * (eigs[J - 1] - eigs[I - 1]) > std::min(gl, gr)) goto L215;
    // cluster detected: additional Rayleigh-Ritz
    ++K;
    J = J - I + 1;
    JDHIP_CHECK(hipMemcpyAsync(D(vv), D((long)n * (I - 1) + xx), nb * J,
                               hipMemcpyDeviceToDevice, s));
    IJ = 0;
L220:
    ++IJ;
    *ndx1 = vv + (long)(IJ - 1) * n;
    *ndx2 = ww + (long)(IJ - 1) * n;
    rlbl = 2;
    if (internal_mv()) {
        do_matvec(D(*ndx1), D(*ndx2));
        goto L225;
    }
    *ijob = 3;
    return;
L225:
    ++*iter;
    if (IJ < J) goto L220;
    {
        std::vector<double> C((size_t)J * J);
        gram_sym(n, D(vv), D(ww), J, C.data());
        for (auto &v : MM) v *= 0.5;
    }
    IJ = solprojep(J, MM.data(), NP.data(), &sclr1, &sclr2, REIGW, 3);
    if (IJ != 0) goto L250;
    {
        std::vector<double> q((size_t)J * J);
        for (int c = 0; c < J; ++c)
            for (int r = 0; r < J; ++r) q[r + (size_t)c * J] = MM[r + (size_t)c * modspace];
        double *dq = ctx_->d_small + 4096;
        put(dq, q.data(), (long)J * J);
        ctx_->ensure_tmp((long)n * J);
        tall_gemm_mfma(s, D(vv), n, J, dq, J, J, ctx_->d_tmp, n, n);
        JDHIP_CHECK(hipMemcpyAsync(D((long)n * (I - 1) + xx), ctx_->d_tmp, nb * J,
                                   hipMemcpyDeviceToDevice, s));
        tall_gemm_mfma(s, D(ww), n, J, dq, J, J, ctx_->d_tmp, n, n);
        JDHIP_CHECK(hipMemcpyAsync(D(vv), ctx_->d_tmp, nb * J, hipMemcpyDeviceToDevice, s));
        if (gep_) {  // BX of the cluster := BX Q (JDaux.F:2296-2300)
            tall_gemm_mfma(s, D((long)n * (I - 1) + bx), n, J, dq, J, J, ctx_->d_tmp, n, n);
            JDHIP_CHECK(hipMemcpyAsync(D((long)n * (I - 1) + bx), ctx_->d_tmp, nb * J,
                                       hipMemcpyDeviceToDevice, s));
        }
    }
    nrc = 0.0;
    gl = gl - eigs[I - 1];
    gr = gr + eigs[I + J - 2];
    for (IJ = 1; IJ <= J; ++IJ) {
        theta = NP[IJ - 1];
        eigs[I + IJ - 2] = theta;
        double v;
        axpy_nrm2(s, rb, D(vv + (long)n * (IJ - 1)), -theta, D((long)n * (I + IJ - 2) + bx), n,
                  ctx_->slot(2));
        fetch(&v, ctx_->slot(2), 1);
        nreig = std::sqrt(v);
        if (gep_) nreig = nreig / dnorm(D((long)n * (I + IJ - 2) + bx), n);
        res[I + IJ - 2] = nreig;
        nrc += nreig * nreig;
    }
    for (IJ = I; IJ <= I + J - 1; ++IJ)
        if (IJ == itarg) --itarg;
    I = I + J;
    if (I > *neig) goto L300;
    goto L210;
L250:  // isolated eigenvalue: true residual
    *ndx1 = (long)n * (I - 1);
    *ndx2 = vv;
    rlbl = 5;
    if (internal_mv()) {
        do_matvec(D(*ndx1), D(*ndx2));
        goto L255;
    }
    *ijob = 3;
    return;
L255:
    ++*iter;
    {
        double d[1], v;
        dot(s, rb, D((long)n * (I - 1)), D(vv), n, ctx_->slot(2));
        fetch(d, ctx_->slot(2), 1);
        theta = d[0];
        axpy_nrm2(s, rb, D(vv), -theta, D((long)n * (I - 1) + bx), n, ctx_->slot(2));
        fetch(&v, ctx_->slot(2), 1);
        nreig = std::sqrt(v);
        if (gep_) nreig = nreig / dnorm(D((long)n * (I - 1) + bx), n);
    }
    eigs[I - 1] = theta;
    res[I - 1] = nreig;
    if (I == itarg) --itarg;
    ++I;
    if (I <= *neig) goto L210;
L300:
L301:
    *ijob = 0;
    ctx_->sync();
}

}  // namespace jdhip
