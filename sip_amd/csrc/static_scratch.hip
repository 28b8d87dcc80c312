// The device scratch array behind asdFftExtExecNoAlloc.
//
// `scratch` is a file-scope device array: every GPU the library runs on holds its own copy,
// at its own address, until the process ends.  In host code the name `scratch` is only a
// host-side stand-in for the symbol, never a device address, so host code uses the address
// that hipGetSymbolAddress returns for the current device, looked up once per device and kept
// in `g_addr`.  Calls on one device share that device's array (a call's kernels overwrite what
// the previous call left there), so overlapping calls on one device race.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "static_scratch.h"

namespace asdfft {

namespace {

__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::atomic<void*> g_addr[kMaxDevices];
std::mutex g_mu;

}  // namespace

hipError_t static_scratch(void** ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    void* p = g_addr[dev].load(std::memory_order_acquire);
    if (p == nullptr) {
        std::lock_guard<std::mutex> lock(g_mu);
        p = g_addr[dev].load(std::memory_order_relaxed);
        if (p == nullptr) {
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            if (p == nullptr) return hipErrorInvalidSymbol;
            g_addr[dev].store(p, std::memory_order_release);
        }
    }
    *ptr = p;
    return hipSuccess;
}

}  // namespace asdfft
