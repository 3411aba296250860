// elems.h -- element conversions shared by the filter kernels (device side).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include "fft_device.h"

namespace msosm {

// gpu_kernel.cu:276-299 -- offset binary to float, code 0 ("no data") -> 0.
__device__ __forceinline__ float2 unpack_u16pair(uint32_t u)
{
    uint32_t x = u & 0xFFFFu, y = u >> 16;
    float fx = (float)x - (x != 0 ? 32768.0f : 0.0f);
    float fy = (float)y - (y != 0 ? 32768.0f : 0.0f);
    return make_float2(fx, fy);
}
// The same scaled by a power of two s (exact: (x - 32768) s as one FMA of
// exact terms), so that a transform of the scaled samples equals the scaled
// transform bit for bit and the scale costs nothing at the store.
__device__ __forceinline__ float2 unpack_u16pair_scaled(uint32_t u, float s)
{
    uint32_t x = u & 0xFFFFu, y = u >> 16;
    float fx = __builtin_fmaf((float)x, s, x != 0 ? -32768.0f * s : 0.0f);
    float fy = __builtin_fmaf((float)y, s, y != 0 ? -32768.0f * s : 0.0f);
    return make_float2(fx, fy);
}

template <bool F16>
struct RingElem;
template <>
struct RingElem<true> {
    using T = __half2;
    __device__ static float2 f2(T x) { return __half22float2(x); }
    __device__ static T zero() { return __floats2half2_rn(0.f, 0.f); }
    // x * h in FP32 from FP16 operands: four mixed-precision FMAs that read
    // the halves directly (v_fma_mix_f32), no separate conversions; the
    // partial products xi*hi and xi*hr are rounded once each, as in an FP32
    // multiply.
    __device__ __forceinline__ static float2 mul(T x, T h)
    {
        const float xr = __low2float(x), xi = __high2float(x), hr = __low2float(h), hi = __high2float(h);
        const float t1 = __builtin_fmaf(xi, hi, 0.0f), t2 = __builtin_fmaf(xi, hr, 0.0f);
        return make_float2(__builtin_fmaf(xr, hr, -t1), __builtin_fmaf(xr, hi, t2));
    }
};
template <>
struct RingElem<false> {
    using T = float2;
    __device__ static float2 f2(T x) { return x; }
    __device__ static T zero() { return make_float2(0.f, 0.f); }
    __device__ __forceinline__ static float2 mul(T x, T h) { return cmul(x, h); }
};

} // namespace msosm
