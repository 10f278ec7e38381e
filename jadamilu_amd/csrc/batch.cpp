// batch.cpp -- many independent eigenproblems on one GPU.
//
// No reference counterpart: the reference solves one problem per DPJD call.
// jdhip_pjd_batch() runs DPJD (jdhip_pjd, internal device SpMV and
// multilevel ILDL preconditioner) on every item of a batch of problems of
// uneven size.  A pool of worker threads, each owning one solver context and
// therefore one HIP stream, takes items from a shared queue ordered by
// decreasing n (longest job first); the host factorisations of different
// items run in parallel on the CPU cores and the device work of different
// items overlaps on the GPU through the workers' streams, which keeps the GPU
// busy even when a single small problem cannot fill it.  Every item starts
// from the state of a fresh process (jdhip_reset).  The choices that a
// single jdhip_pjd call makes by timing (kernel form of each factor product,
// sweeps against exact solves, re-factorisation by PJDRVCOM's cost model) are
// made here from the inputs alone (forms by row length, sweep counts by
// accuracy, the cost model with every interval read as 0 s), because timings
// taken while other workers share the GPU are not reproducible: an item's
// result therefore depends only on its inputs, and equals that of a jdhip_pjd
// call with the options ildl_autotune = 0, ildl_tune_time = 0 and
// cost_clock = 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jadamilu_amd.h"

extern "C" int jdhip_pjd_batch_opts(int device, int nitems, jdhip_batch_item *items,
                                    int nworkers, int nopts, const char *const *names,
                                    const double *values) {
    if (nitems < 0 || (nitems > 0 && !items)) return -1;
    if (nopts < 0 || (nopts > 0 && (!names || !values))) return -1;
    // graph capture of the preconditioner apply is not offered here: while
    // one worker captures, another worker's allocations and frees (factor
    // uploads, workspaces) may synchronise the device and invalidate the
    // capture
    for (int o = 0; o < nopts; ++o)
        if (names[o] && std::string(names[o]) == "ildl_graph" && values[o] != 0.0) return -1;
    if (nitems == 0) return 0;
    for (int i = 0; i < nitems; ++i) {
        const jdhip_batch_item &it = items[i];
        if (it.n <= 0 || !it.ia || !it.ja || !it.a || !it.eigs || !it.res || it.neig <= 0)
            return -1;
    }
    for (int i = 0; i < nitems; ++i) {  // until a worker has solved the item
        items[i].info = JDHIP_INFO_DEVICE_ERROR;
        items[i].neig_out = 0;
        items[i].iter = 0;
    }
    nworkers = std::max(1, std::min(nworkers, nitems));
    std::vector<int> order(nitems);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return items[x].n > items[y].n; });
    std::atomic<int> next{0};

    auto worker = [&]() {
        jdhip_ctx *ctx = jdhip_create(device);
        if (!ctx) return;  // this worker cannot run; the others drain the queue
        // the workers already run factorisations in parallel
        jdhip_set_option(ctx, "ildl_threads", 1);
        // set-up choices from the inputs only (see the header comment)
        jdhip_set_option(ctx, "ildl_autotune", 0);
        jdhip_set_option(ctx, "ildl_tune_time", 0);
        jdhip_set_option(ctx, "cost_clock", 1);
        // the caller's options, applied to every worker context; an unknown
        // name or a value out of range leaves the items unsolved
        for (int o = 0; o < nopts; ++o)
            if (jdhip_set_option(ctx, names[o], values[o]) != 0) {
                jdhip_destroy(ctx);
                return;
            }
        double *x = nullptr;
        long xcap = 0;
        std::vector<int> ia, ja;
        std::vector<double> a;
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= nitems) break;
            jdhip_batch_item &it = items[order[k]];
            try {
                const int n = it.n;
                const int nnz = it.ia[n] - 1;
                ia.assign(it.ia, it.ia + n + 1);
                ja.assign(it.ja, it.ja + nnz);
                a.assign(it.a, it.a + nnz);
                const int msp = it.madspace;
                // optimal workspace of PJD.F:99-103
                const long lx = (long)n * (3L * msp + it.neig + 1) + 3L * msp * msp +
                                std::max<long>((long)msp * msp, it.neig);
                if (lx > xcap) {
                    if (x) hipFree(x);
                    x = nullptr;
                    if (hipMalloc(&x, sizeof(double) * lx) != hipSuccess) {
                        x = nullptr;
                        xcap = 0;
                        continue;  // the item keeps INFO = JDHIP_INFO_DEVICE_ERROR
                    }
                    xcap = lx;
                }
                jdhip_reset(ctx);
                int neig = it.neig, iter = it.maxiter, info = 0;
                int icntl[5];
                for (int c = 0; c < 5; ++c) icntl[c] = it.icntl[c];
                double shift = it.shift, droptol = it.droptol, mem = it.mem, gap = 0.0;
                jdhip_pjd(ctx, n, a.data(), ja.data(), ia.data(), it.eigs, it.res, x, lx, &neig,
                          it.sigma, it.isearch, 0, msp, &iter, it.tol, &shift, &droptol, &mem, icntl,
                          0, &info, &gap);
                it.info = info;
                it.iter = iter;
                it.neig_out = info >= 0 ? neig : 0;
                it.gap = gap;
                if (it.x && it.neig_out > 0)
                    if (hipMemcpyAsync(it.x, x, sizeof(double) * (size_t)n * it.neig_out,
                                       hipMemcpyDeviceToHost,
                                       (hipStream_t)jdhip_stream(ctx)) != hipSuccess ||
                        hipStreamSynchronize((hipStream_t)jdhip_stream(ctx)) != hipSuccess)
                        it.info = JDHIP_INFO_DEVICE_ERROR;
            } catch (const std::exception &) {
                it.info = JDHIP_INFO_DEVICE_ERROR;  // host allocation failure
            }
        }
        if (x) hipFree(x);
        jdhip_destroy(ctx);
    };
    std::vector<std::thread> pool;
    pool.reserve(nworkers);
    for (int w = 0; w < nworkers; ++w) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    int failed = 0;
    for (int i = 0; i < nitems; ++i) failed += items[i].info != 0;
    return failed;
}

extern "C" int jdhip_pjd_batch(int device, int nitems, jdhip_batch_item *items, int nworkers) {
    return jdhip_pjd_batch_opts(device, nitems, items, nworkers, 0, nullptr, nullptr);
}
