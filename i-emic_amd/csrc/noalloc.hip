// F(x) for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// iemic_rhs_dev reduces the integral-condition row (coef . x, THCM.C:1013-1027)
// through partial sums in a buffer the context allocated.  This variant keeps
// those partials in a file-scope device array instead, one partial per 256
// unknowns, so a call allocates nothing.  Each GPU holds its own copy of the
// array (NOALLOC_PARTIALS doubles, 128 KiB) until the process ends; its device
// address is looked up once per device with hipGetSymbolAddress and only that
// address is handed to kernels.  Every partial is written before it is read,
// so nothing needs clearing between calls; calls that overlap on one GPU
// (several contexts, streams or threads at once) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.hpp"

namespace iemic {

constexpr int NOALLOC_PARTIALS = IEMIC_NOALLOC_MAX_NDIM / 256;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES] = {};

// device address of this GPU's scratch array (the current device), looked
// up at the first call on that device
static int scratch_address(double **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(IEMIC_EUNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(IEMIC_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return IEMIC_OK;
}

}  // namespace iemic

extern "C" int iemic_rhs_dev_noalloc(iemic_ctx *c, const double *d_x, double *d_F)
{
    using namespace iemic;
    if (!c) return set_error(IEMIC_EINVAL, "null context");
    if (!d_x || !d_F) return set_error(IEMIC_EINVAL, "null vector");
    if ((long long)c->sc.ndim > IEMIC_NOALLOC_MAX_NDIM)
        return set_error(IEMIC_EINVAL, "iemic_rhs_dev_noalloc: ndim " + std::to_string(c->sc.ndim) +
                                           " exceeds the maximum " + std::to_string(IEMIC_NOALLOC_MAX_NDIM));
    double *part = nullptr;
    int rc = scratch_address(&part);
    if (rc) return rc;
    return launch_rhs_partials(c, d_x, d_F, part, NOALLOC_PARTIALS);
}
