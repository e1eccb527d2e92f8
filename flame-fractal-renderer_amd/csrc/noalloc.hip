// Staging memory for callers that cannot allocate device memory
// (ffr_add_buffer_noalloc): a file-scope device array that every call on a
// GPU shares.  Host code never passes the array's name to a kernel or a
// copy: in host code the name is a host-side stand-in, not a device
// address.  The device address of each GPU's copy is looked up once per
// device, with that device current, and kept.
#include "util_kernels.hpp"

#include <hip/hip_runtime.h>

#include <mutex>

namespace ffr {

// 16 MiB per GPU, held from the first call on that GPU until the process
// ends (the code object keeps it)
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up on the
// first call for that device)
int scratch_address(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices)
        return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return (int)e;
        if (!p)
            return (int)hipErrorInvalidSymbol;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return 0;
}

} // namespace

int add_cells_staged_noalloc(void *dst, const void *src, unsigned long long bytes, int cw,
                             int word_bytes, hipStream_t s)
{
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess)
        return (int)e;
    void *stage = nullptr;
    int rc = scratch_address(device, &stage);
    if (rc)
        return rc;
    const unsigned long long cell_bytes = (unsigned long long)cw * (unsigned long long)word_bytes;
    const unsigned long long chunk = (kNoAllocScratchBytes / cell_bytes) * cell_bytes;
    if (chunk == 0)
        return (int)hipErrorInvalidValue;
    // the copy of chunk i + 1 into the staging memory is ordered after the
    // add of chunk i on the same stream
    for (unsigned long long off = 0; off < bytes; off += chunk) {
        const unsigned long long n = bytes - off < chunk ? bytes - off : chunk;
        e = hipMemcpyAsync(stage, (const char *)src + off, (size_t)n, hipMemcpyHostToDevice, s);
        if (e != hipSuccess)
            return (int)e;
        rc = launch_add_cells((char *)dst + off, stage, n / (unsigned long long)word_bytes, cw,
                              word_bytes, s);
        if (rc)
            return rc;
    }
    return (int)hipStreamSynchronize(s);
}

} // namespace ffr
