// ildl.cpp -- multilevel incomplete LDL^T preconditioner.
#include "ildl.hpp"

#include <stdexcept>

namespace jdhip {

struct IldlLevelsDev {};

IldlPrecond::IldlPrecond(DevCtx *ctx) : ctx_(ctx) {}
IldlPrecond::~IldlPrecond() = default;

IldlResult IldlPrecond::setup(int, const int *, int *, double *, const IldlOptions &) {
    IldlResult r;
    r.ierr = -54321;
    return r;
}

void IldlPrecond::apply(double *, const double *) {
    throw std::runtime_error("ILDL preconditioner not available");
}

}  // namespace jdhip
