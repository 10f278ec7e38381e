// engine.hpp -- device-resident Jacobi-Davidson engine (host control).
//
// JdEngine restates JDRVCOM (src/jd/JDaux.F:1391-2415) with its inner solver
// LSOLREVCOM (2989-3268) and STOPTEST (2542-2672) as an explicit-state
// coroutine.  All n-length data lives in a device workspace X laid out like
// the reference's X array (so reverse-communication offsets NDX1/NDX2 mean
// the same thing); the J x J projected matrices, the Ritz coefficients and
// the scalar control stay on the host.  Matrix-vector products and
// preconditioner solves are either returned to the caller (reverse
// communication) or executed inline through an Ops object, in which case the
// inner iteration runs as a short chain of fused kernels with one host
// synchronisation per step.
#pragma once
#include <chrono>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace jdhip {

#define JDHIP_CHECK(expr)                                                              \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) + \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// Per-device resources of a solver context.
struct DevCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    RedBuf rb{};
    double *d_small = nullptr;   // 4 x kMaxCols coefficient slots + scalars
    InnerScalars *d_sc = nullptr;
    double *h_pin = nullptr;     // pinned host staging (4 x kMaxCols + 8)
    InnerScalars *h_sc = nullptr;
    double *d_tmp = nullptr;     // n x ncols scratch (tall products)
    long tmp_len = 0;
    double *d_gram = nullptr;    // partial sums of the MFMA Gram products (kGramScratch)
    double *d_gout = nullptr;    // Gram result, up to kMaxCols x kMaxCols

    void init(int dev);
    void release();
    void ensure_tmp(long len);
    // the host's waits for the stream: spinning (hipStreamSynchronize, the
    // default) or, with blocking_sync, sleeping on an event created with
    // hipEventBlockingSync, which leaves the core to other host threads
    hipEvent_t ev_sync = nullptr;
    bool blocking_sync = false;
    void sync() {
        if (blocking_sync) {
            JDHIP_CHECK(hipEventRecord(ev_sync, stream));
            JDHIP_CHECK(hipEventSynchronize(ev_sync));
        } else {
            JDHIP_CHECK(hipStreamSynchronize(stream));
        }
    }
    // device slots
    double *slot(int k) { return d_small + (long)k * kMaxCols; }
};

class JdEngine;

// Inline operators.  When present, the engine executes them without
// returning to the caller.
struct Ops {
    virtual ~Ops() = default;
    virtual bool has_matvec() const { return false; }
    // y = A x ; if eta_dev/eta: y -= eta x ; if u: *udot_dev = u.y
    virtual void matvec(const double *x, double *y, const double *eta_dev, double eta,
                        const double *u, double *udot_dev) {}
    virtual bool has_precond() const { return false; }
    // called when the reference would return IJOB=2 to PJDRVCOM
    virtual void precond(JdEngine &e, double *z, const double *r) {}
    // FP32 inner solver (mixed path): operator and preconditioner on float
    // vectors; only used when has_f32() and both operators are inline.
    virtual bool has_f32() const { return false; }
    virtual void matvec_f(const float *x, float *y, double eta, const float *u, double *udot) {}
    virtual void precond_f(JdEngine &e, float *z, const float *r) {}
    // generalized problem A x = lambda B x: y = B x (inline when has_matvec)
    virtual void bmatvec(const double *x, double *y) {}
};

class JdEngine {
  public:
    explicit JdEngine(DevCtx *ctx) : ctx_(ctx) {}
    ~JdEngine() {
        if (fbuf_) (void)hipFree(fbuf_);
    }

    // One JDRVCOM call.  x is the DEVICE workspace (lx doubles), eigs/res
    // host arrays.  On return *ijob: 0 done, 1 matvec X(ndx2)=A X(ndx1) and,
    // for the generalized problem, X(ndx3)=B X(ndx1), 2 precond
    // X(ndx1)=P^-1 X(ndx2), 3 matvec X(ndx2)=A X(ndx1) only (final checks),
    // 4 X(ndx3)=B X(ndx1) only.  ndx are 0-based offsets into x.  gep (read
    // when *ijob = 0): A x = lambda B x with B symmetric positive definite
    // (JDRVCOM's GEP branches).
    void step(int n, double *eigs, double *res, double *x, long lx, int *neig, double sigma,
              int isearch, int ninit, int madspace, int *iter, double tol, int *ijob,
              long *ndx1, long *ndx2, long *ndx3, int iprint, int *info, double *gap,
              bool gep = false);
    bool gep() const { return gep_; }

    void set_ops(Ops *ops) { ops_ = ops; }

    // COMMON /JDRVCOMJD/ (JDaux.F:46-47): read by the preconditioner manager
    double timeilu = 0, rinit = 0, rcurr = 0, rcum = 0, shift1 = 0;
    bool flagcg = false;
    long inner_iters = 0;  // total inner steps (diagnostic)
    long nsync = 0;        // host synchronisations (diagnostic)

  private:
    // ---- device helpers ----
    double dnorm(const double *w, long n);
    void repgs(int n, long voff, int jv, long woff, std::vector<double> &y, int gam) {
        repgs2(n, voff, voff, jv, woff, y, gam);
    }
    // REPGS with separate bases: coefficients from the columns at bvoff,
    // subtraction with those at voff (the same for the standard problem)
    void repgs2(int n, long voff, long bvoff, int jv, long woff, std::vector<double> &y,
                int gam);
    void sortev_dev(int n, int j, long voff, double *eigs, double *r, int it, double sigma,
                    long v2off = -1);
    int solprojep(int j, double *mm, double *nn, double *theta, double *theta2,
                  std::vector<double> &w, int is);
    int shrink(int n, int j, int k, int ucol_off, int iserch);
    void gram_initial(int n);
    void gram_sym(int n, const double *V, const double *W, int J, double *C);
    void fetch(double *host, const double *dev, long cnt);
    void put(double *dev, const double *host, long cnt);
    void stoptest(int n, double tol, double *rk, double alprho, double gk, double *theta,
                  double eta, int *info, double trs, double xn2);
    // generalized problem: the GEP terms of STOPTEST (pcad, pcr, ||BT||,
    // BU.BT); may change alprho
    void stoptest_gep(int n, double tol, double *rk, double *alprho, double gk, double *theta,
                      double eta, int *info, double trs, double pcad, double pcr);
    void do_bmatvec(const double *xin, double *yout) { ops_->bmatvec(xin, yout); }
    double ddot(const double *a, const double *b, long n);
    bool internal_mv() const { return ops_ && ops_->has_matvec(); }
    bool internal_pc() const { return ops_ && ops_->has_precond(); }
    bool use_f32() const { return internal_mv() && internal_pc() && ops_->has_f32(); }
    void ensure_f32(long n, int nq);
    // FP32 inner-solver buffers: r z p q x x2 (n each) and Q ((NEIG+1) n)
    float *fbuf_ = nullptr;
    long fbuf_len_ = 0;
    bool f32_ = false;
    bool f32_last_ = false;  // the last inner solve ran in FP32
    void do_matvec(const double *xin, double *yout) {
        ops_->matvec(xin, yout, nullptr, 0.0, nullptr, nullptr);
    }

    DevCtx *ctx_;
    Ops *ops_ = nullptr;
    double *X_ = nullptr;  // device workspace of the current call

    // host small matrices: MM, NN, MP, NP (modspace^2 each) and work
    std::vector<double> MM, NN, MP, NP, WL, WL2, WL3, REIGW;
    double *YY() { return MP.data(); }
    double *YY2() { return MP.data() + modspace; }

    // ---- JDRVCOM locals (SAVE) ----
    int rlbl = 0;
    bool ftflg = false, conv = false, uknown = false, initflg = false, neweig = false;
    int mvmax = 0, midspace = 0, modspace = 0, nwant = 0, itarg = 0, iserch = 0;
    long xx = 0, xcg = 0, xcg2 = 0, wcg = 0, vv = 0, ww = 0, uu = 0, reig = 0;
    long bx = 0, bu = 0, bv = 0, kxp = 0, kbxp = 0, bt = 0;
    bool gep_ = false;
    double nbu = 1.0, axpybeta = 0.0;
    // inner solver of the generalized problem (host scalars)
    double g_rho = 0, g_rho1 = 0, g_ttk2 = 0, g_ttk2m = 0, g_ggk2 = 0;
    int iterp = 0, itcg = 0, ifo = 0, istep = 0, istag = 0;
    int I = 0, IJ = 0, J = 0, K = 0, II = 0, JJ = 0;
    double resid = 0, sclr1 = 0, sclr2 = 0, theta = 0, eta = 0, thmt = 0, delta1 = 0,
           deltao = 0, eta0 = 0, nreig = 0, theta2 = 0, eta1 = 0, epsmsr = 0, nreigp = 0,
           trs = 0, rndind = 0, gr = 0, gl = 0, gg = 0, nrc = 0, rinitm = 0, alpha = 0,
           gk = 0, alprho = 0, axpyalpha = 0;
    std::vector<double> ycoef;

    // ---- LSOLREVCOM locals ----
    int ls_iter = 0, ls_maxit = 0;
    bool ls_posdef = false;

    // ---- STOPTEST locals (SAVE) ----
    int st_istag = 0, st_istt = 0;
    bool st_toupdate = false;
    double st_g0 = 0, st_gs = 0, st_gtrs = 0, st_gkmin = 0, st_bk = 0, st_beta = 0,
           st_tn = 0, st_un2 = 0, st_gkm1 = 0, st_alp = 0, st_bkgep = 0;

    unsigned long long rng_ = 0x2545F4914F6CDD1DULL;

    // progress reporting (JDHIP_VERBOSE)
    void progress(int matvecs, int converged);
    bool progress_started_ = false;
    std::chrono::steady_clock::time_point progress_t0_, progress_last_;
};

}  // namespace jdhip
