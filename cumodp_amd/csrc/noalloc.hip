// noalloc.hip -- the scratch of the non-allocating multiply
// (cumodp_fftmul_uni_dev_noalloc in api.cpp).
//
// The four-step multiply needs two n-word work arrays besides its operands
// and output.  cumodp_fftmul_uni_dev takes them from per-stream buffers that
// the library hipMallocs and grows.  Callers that may not allocate (code
// under graph capture, tight serving loops) use this file-scope device array
// instead: every device that the library runs on holds its own copy,
// reserved when the code object is loaded and kept until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory, so
// its device address is looked up once per device with hipGetSymbolAddress
// and only that address is ever handed to kernels.
#include <mutex>

#include "internal.h"

namespace cm {

namespace {

// two arrays of 2^kNoAllocMaxLog words (F and G work arrays): 32 MiB
constexpr size_t kScratchWords = (size_t)2 << kNoAllocMaxLog;
__device__ __align__(256) u32 scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
u32 *g_scratch_ptr[kMaxDevices];   // scratch's device address per device (null: not looked up)

}  // namespace

size_t noalloc_scratch_words() { return kScratchWords; }

hipError_t noalloc_scratch(u32 **out) {
    *out = nullptr;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = (u32 *)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace cm
