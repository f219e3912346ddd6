// The statically allocated device scratch of iclblasSgemmNoAlloc.
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy, at its own address, from the loading of this code object until
// the process ends, and no call ever allocates it. Host code never passes the
// array's name to a kernel (in host code the name is a stand-in, not a device
// address): icl_noalloc_scratch_init resolves the address with
// hipGetSymbolAddress once per device, with that device current, and every later
// use goes through the pointer it keeps.
//
// Every call on one GPU shares that GPU's array and the array keeps the last
// call's contents; the split-BF16 pack kernel writes every byte the GEMM reads
// (padding included), so nothing has to start at zero. Calls that overlap on one
// GPU (two streams, or two handles at once) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "iclBLAS_ext.h"
#include "kernel_abi.h"

static __device__ __align__(256) unsigned char scratch[ICLBLAS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};

} // namespace

extern "C" hipError_t icl_noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_done[device]) return g_err[device];
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_err[device] = e;
    g_done[device] = true;
    return e;
}

extern "C" void* icl_noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_ptr[device];
}
