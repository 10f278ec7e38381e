// ildl.hpp -- multilevel incomplete LDL^T preconditioner (host factor, device apply).
#pragma once
#include <memory>
#include <vector>

#include "engine.hpp"

namespace jdhip {

struct IldlOptions {
    int isearch = 0;
    double sigma = 0.0;
    double shift = 0.0;
    double droptol = 1e-3;
    double mem = 0.0;
    int icntl3 = 0, icntl4 = 0;
    bool single = false;  // FP32 factor and apply (ICNTL(5) != 0)
};

struct IldlResult {
    int ierr = 0;
    double shift = 0.0, diagmin = 0.0, shiftmax = 0.0, droptol = 0.0, time = 0.0;
    double fill = 0.0;
    int factspd = 0, nlev = 0;
};

struct IldlLevelsDev;

class IldlPrecond {
  public:
    explicit IldlPrecond(DevCtx *ctx);
    ~IldlPrecond();
    IldlResult setup(int n, const int *ia, int *ja, double *a, const IldlOptions &opt);
    void apply(double *z, const double *r);  // z = M^-1 r on ctx->stream

  private:
    DevCtx *ctx_;
    std::unique_ptr<IldlLevelsDev> dev_;
};

}  // namespace jdhip
