// ildl.hpp -- multilevel incomplete LDL^T preconditioner: host factor
// (ildl_factor.cpp), device apply (ildl_kernels.hip).
#pragma once
#include <memory>
#include <vector>

#include "engine.hpp"
#include "ildl_factor.hpp"

namespace jdhip {

struct IldlOptions {
    double droptol = 1e-3;
    double condest = 5.0;
    double mem = 0.0;    // elbow space (stored entries / nnz(A)), <= 0: unlimited
    bool spd = false;    // |D| conversion (ISEARCH <= 1)
    bool single = false; // FP32 factor, storage and apply (ICNTL(5) != 0)
    int sweeps = -1;     // 0: exact level-scheduled; k > 0: k Jacobi sweeps; -1: tuned
    int ordering = 2;  // 0 natural, 1 RCM, 2 AMD
    int dense_max = 1500;
    int threads = 0;     // host threads of the factorisation (0: all cores, at most 16)
    int coupling = 0;    // 0: F blocks, two B solves per level (the reference's); 1: L_CB
};

struct IldlResult {
    int ierr = 0;
    int nlev = 0;
    long nnz = 0;
    int nneg = 0;
    double condest = 0.0;
    double fill = 0.0;  // stored entries / nnz(upper A)
    double seconds = 0.0;
    std::vector<int> level_n, level_nB, fwd_levels, bwd_levels;
    int sweeps = 0;  // triangular-solve mode in use (0 exact, k Jacobi sweeps)
};

class IldlImplBase {
  public:
    virtual ~IldlImplBase() = default;
    virtual void apply(double *z, const double *r) = 0;
    virtual void apply_f(float *z, const float *r) = 0;  // FP32 in/out
    virtual int nlevels() const = 0;
    virtual int dense_m() const = 0;
    // level data in double precision (host copies of the factor)
    virtual void level_sizes(int lev, int *n, int *nB, long *nnz_lb, long *nnz_lc) const = 0;
    virtual void level_data(int lev, int *perm, double *scal, int *lb_ptr, int *lb_idx,
                            double *lb_val, int *lc_ptr, int *lc_idx, double *lc_val,
                            double *d) const = 0;
    virtual void dense_data(double *minv) const = 0;
    virtual long apply_bytes() const = 0;  // algorithmic bytes of one apply
    virtual bool single() const = 0;
    // Pick the triangular-solve mode by measurement: the fastest of the exact
    // level-scheduled solve and k Jacobi sweeps whose apply stays within
    // relative distance tol of the exact apply on a fixed random vector.
    // by_time = false skips the timing comparison: sweeps are kept whenever
    // they are accurate (a choice that depends on the inputs only).
    virtual int tune_sweeps(double tol, bool by_time) = 0;
    // Time reps launches of one Jacobi sweep over the first level's L_B (the
    // largest single kernel of the apply) on the context stream.
    virtual int time_sweep(int reps, double *avg_ms, long *alg_bytes) = 0;
    virtual int sweeps() const = 0;
    virtual std::vector<int> level_sweeps() const = 0;
    // Choose every product's kernel form (SELL or CSR lanes) by timing.
    virtual void autotune_forms() = 0;
    // Set every product's form (0 SELL, 1 CSR-stream, 4/8/16/32/64 CSR lanes
    // per row).
    virtual void set_forms(int form) = 0;
    // per level: forms of the forward sweep, backward sweep, L_CB, L_CB^T
    // (0 SELL, 1 CSR-stream, k >= 4 lanes per row CSR)
    virtual std::vector<int> level_forms() const = 0;
    // replay the apply from captured HIP graphs (default on)
    virtual void set_graphs(bool on) = 0;
};

// Host factor (no device needed).
struct IldlHost {
    bool single = false;
    bool spd = false;  // factor converted to |D| (ISEARCH <= 1)
    IldlMultilevel<double> fd;
    IldlMultilevel<float> ff;
    IldlResult res;
    long nnz_upper = 0;
};

// Factor (A - shift I), A given as a 1-based upper-triangle CSR.
void ildl_factor_host(int n, const int *ia, const int *ja, const double *a, double shift,
                      const IldlOptions &opt, IldlHost &out);

class IldlPrecond {
  public:
    explicit IldlPrecond(DevCtx *ctx) : ctx_(ctx) {}
    // factor on the host, then upload
    IldlResult setup(int n, const int *ia, const int *ja, const double *a, double shift,
                     const IldlOptions &opt);
    // upload an existing host factor (copied)
    void attach(const IldlHost &h, int sweeps);
    void apply(double *z, const double *r) { impl_->apply(z, r); }
    IldlImplBase *impl() { return impl_.get(); }
    bool ready() const { return (bool)impl_; }
    IldlResult result;
    // device layout: storage order of each level (0 factor order, 1 original
    // matrix index order) and XCD-contiguous block order in the sweeps
    int storage_order = 1;
    int xcd_remap = 1;
    int sell_sigma = 256;  // SELL-C-sigma sorting window of the sweep copies
    bool graphs = false;   // replay the apply from captured HIP graphs
    // work vectors of an FP32 factor's apply: -1 FP32 for a definite
    // problem and FP64 otherwise, 0 always FP64, 1 always FP32
    int vec32 = -1;
    bool autotune = true;  // kernel forms by measurement (else by row length)
    int force_form = -1;   // >= 0: this form for every product (tests, experiments)
    double tune_tol = 0.0;  // sweep tuning tolerance (0: 3e-2 definite, 1e-2 indefinite)
    bool tune_time = true;  // sweeps kept only if faster than exact solves (timed)

  private:
    DevCtx *ctx_;
    std::unique_ptr<IldlImplBase> impl_;
};

}  // namespace jdhip
