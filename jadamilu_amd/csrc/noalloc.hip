// noalloc.hip -- jdhip_gram_noalloc: the Gram product C = V^T W of jdhip_gram
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The partial sums that jdhip_gram keeps in its context's buffer live here in
// a file-scope device array of kGramScratch doubles (4 MiB), which every GPU
// the library runs on holds until the process ends.  Host code never passes
// the array's name to a kernel (in host code the name is only a stand-in for
// the symbol): the device address of each GPU's copy is looked up once per
// device with hipGetSymbolAddress, with that device current, and that address
// is what the kernels receive.  The entry point launches kernels only: no
// hipMalloc, hipFree or allocating library call.
//
// Every call on one GPU shares that GPU's array.  Nothing in it needs to start
// at zero: each call's first kernel writes every partial sum it later reads.
// Calls that overlap in time on one GPU (different streams, or host threads
// without ordering) race on the array and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/jadamilu_amd.h"
#include "kernels.hpp"

namespace {

static __device__ __align__(256) double scratch[jdhip::kGramScratch];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up on first use).
double *scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_ptr[dev] = static_cast<double *>(p);
    }
    return g_ptr[dev];
}

}  // namespace

extern "C" int jdhip_gram_noalloc(void *stream, long n, int k, int m, const double *v_dev,
                                  long ldv, const double *w_dev, long ldw, double *c_dev,
                                  int ldc) {
    if (n < 0 || k < 0 || m < 0) return -1;
    if (k > 64 || m > 64) {
        std::fprintf(stderr,
                     "jdhip_gram_noalloc: k = %d, m = %d: at most 64 columns each (the "
                     "static scratch holds %ld doubles of partial sums)\n",
                     k, m, (long)jdhip::kGramScratch);
        return -1;
    }
    if (k == 0 || m == 0) return 0;
    if (!c_dev || ldc < k || (n > 0 && (!v_dev || !w_dev || ldv < n || ldw < n))) return -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -2;
    double *ws = scratch_ptr(dev);
    if (!ws) return -2;
    if (jdhip::gram(static_cast<hipStream_t>(stream), n, k, m, v_dev, ldv, w_dev, ldw, c_dev, ldc,
                    ws) != 0)
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
