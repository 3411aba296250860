// noalloc.hip -- the batched C2C FFT for callers that cannot allocate device
// memory (code under graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// Transforms above 2^14 points run as two passes through a scratch buffer of
// batch * n complex64 elements.  msosm_k_fft_c2c allocates that buffer per
// call (stream-ordered); msosm_k_fft_c2c_noalloc takes it from the fixed
// device array below instead, so it never allocates.  Every GPU that runs
// the library holds its own copy of the array until the process ends.  Pass
// A writes every scratch element it reads back in pass B, so nothing is
// cleared between calls; two calls that overlap on one GPU (different
// streams, or host threads) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"
#include "kernels.h"
#include "msosm_c.h"

namespace {

constexpr size_t kScratchElems = size_t(1) << 22; // complex64 elements: 32 MiB
constexpr int kMaxDevices = 64;

__device__ __align__(256) float2 scratch[kScratchElems];

std::mutex g_mu;
float2 *g_scratch[kMaxDevices] = {}; // device address of `scratch`, per device

// Device address of this device's copy of `scratch` (looked up once per
// device, with that device current).  The array's name is a host-side
// stand-in and is never passed to a kernel or a copy.
int scratch_ptr(float2 **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return msosm::set_error(MSOSM_ERR_HIP, "fft_c2c_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return msosm::set_error(MSOSM_ERR_UNSUPPORTED, "fft_c2c_noalloc: device %d above %d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return msosm::set_error(MSOSM_ERR_HIP, "fft_c2c_noalloc: hipGetSymbolAddress: %s",
                                    hipGetErrorString(e));
        g_scratch[dev] = static_cast<float2 *>(p);
    }
    *out = g_scratch[dev];
    return MSOSM_OK;
}

int ilog2_pow2(size_t v)
{
    int l = 0;
    while ((size_t(1) << l) < v)
        l++;
    return l;
}

} // namespace

extern "C" int msosm_k_fft_c2c_noalloc(const float *in, float *out, size_t n, size_t batch, size_t idist,
                                       size_t odist, int direction, void *stream)
{
    if (!in || !out || n < 2 || (n & (n - 1)) || n > (size_t(1) << 28) || batch < 1 ||
        (direction != -1 && direction != 1))
        return msosm::set_error(MSOSM_ERR_ARG,
                                "fft_c2c_noalloc: n must be a power of two in [2, 2^28], batch >= 1, direction -1 "
                                "or +1");
    const int log2n = ilog2_pow2(n);
    float2 *sc = nullptr;
    if (log2n > msosm::kMaxLog2SinglePass) {
        const size_t need = msosm::fft_c2c_scratch_elems(log2n, (long long)batch);
        if (batch > kScratchElems || need > kScratchElems)
            return msosm::set_error(MSOSM_ERR_ARG,
                                    "fft_c2c_noalloc: batch * n = %zu * %zu exceeds the fixed scratch of %zu complex "
                                    "samples (sizes up to 2^14 points need none)",
                                    batch, n, kScratchElems);
        int rc = scratch_ptr(&sc);
        if (rc != MSOSM_OK)
            return rc;
    }
    hipError_t e = msosm::launch_fft_c2c(reinterpret_cast<const float2 *>(in), reinterpret_cast<float2 *>(out),
                                         log2n, (long long)batch, (long long)idist, (long long)odist, direction == 1,
                                         1.0f, sc, static_cast<hipStream_t>(stream));
    if (e != hipSuccess)
        return msosm::set_error(MSOSM_ERR_HIP, "fft_c2c_noalloc: %s", hipGetErrorString(e));
    return MSOSM_OK;
}
