// scratch.hip -- the static device scratch of the allocation-free entry point
// (convgemm_conv_noalloc): repacked weights and split-K partials live in a
// file-scope device array instead of memory the library allocates, so a call
// makes no hipMalloc / hipFree.  Every GPU the library runs on holds its own
// copy of the array, at its own address, until the process ends.
//
// In host code `scratch` is only a host-side stand-in for the symbol: its
// device address comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address is ever handed to kernels or
// copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../cg_internal.h"

namespace {

static __device__ __align__(256) char scratch[CG_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` per device (nullptr: not looked up yet)

}  // namespace

extern "C" int cg_scratch_get(void **ptr, size_t *bytes) {
    *ptr = nullptr;
    *bytes = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_ptr[dev] = p;
    }
    *ptr = g_ptr[dev];
    *bytes = CG_SCRATCH_BYTES;
    return 0;
}
