// rh_noalloc.hip -- rh_field_dev_noalloc, the field entry point for callers
// that cannot allocate device memory.  Its partial sums live in a file-scope
// device array (one copy per GPU, held until the process ends) whose device
// address is resolved once per device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/radhip.h"
#include "rh_api_internal.hpp"

constexpr int64_t kScratchDoubles = int64_t(RH_NOALLOC_SCRATCH_BYTES) / int64_t(sizeof(double));

// Host code must not pass `scratch` itself to a kernel: on the host the name
// stands for a host-side shadow, not the device copy.  g_scratch holds the
// device address per GPU.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// The calling thread has `device` current.
rh_status scratch_of(int device, double** out)
{
    if (device < 0 || device >= kMaxDevices)
        return rh_internal::fail(RH_ERR_ARG, "rh_field_dev_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return rh_internal::fail(RH_ERR_HIP, std::string("rh_field_dev_noalloc: hipGetSymbolAddress: ") +
                                                     hipGetErrorString(e));
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return RH_OK;
}

}  // namespace

extern "C" rh_status rh_field_dev_noalloc(rh_model* m, const double* d_M, const double* d_pts, int64_t np,
                                          double* d_B, double* d_H, rh_precision prec)
{
    if (!m) return rh_internal::fail(RH_ERR_ARG, "bad argument");
    const hipError_t e = rh_internal::enter_device(rh_internal::model_device(m));
    if (e != hipSuccess) return rh_internal::fail(RH_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    double* work = nullptr;
    const rh_status st = scratch_of(rh_internal::model_device(m), &work);
    if (st != RH_OK) return st;
    return rh_internal::field_dev_into(m, d_M, d_pts, np, d_B, d_H, prec, work, kScratchDoubles);
}
