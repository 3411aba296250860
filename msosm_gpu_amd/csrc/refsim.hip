// refsim.hip -- the simulated test signal the reference's drivers feed their
// engines (SimulatedComplexSignal + AWGN, simulated_pulsar/
// simulated_complex_signal.cpp:12-236, simulated_pulsar/AWGN.cpp:21-58), so
// that its benchmark scenarios (e.g. check_msosm.cpp:101-105) can be rerun
// with the same inputs.
//
// One period of Np samples holds two unit impulses (at Np/2 and Np/2 + 2),
// dispersed with the inverse of the dedispersion chirp: H1 = exp(+i phi) on
// the float-rounded frequency grid of `order` points (the planner's grid),
// H1 <- IDFT(H1)/order placed in the second half of a 2*order-point
// buffer, H = DFT of that.  The impulses are shifted by 0 .. range-1 periods,
// filtered with H by FFT, and the first Np samples of each result summed
// (the dispersion spreads a pulse over range periods).  Noise: complex
// Gaussian from std::default_random_engine (default seed) and two
// std::normal_distribution<float>(0, 1), one per component, scaled to the
// requested SNR against the measured signal power.  uint16 conversion:
// v == 0 -> 0, else (uint16_t)(v * 32767 + 32768).
//
// The transforms run on the GPU in FP32 (the reference used FFTW in single
// precision), so values agree with the reference's to FFT rounding; the
// noise is the same sequence whenever the C++ standard library implements
// those two classes the same way (libstdc++ here and in the reference).
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "msosm_c.h"

namespace {

#define SHIP(expr)                                                                                               \
    do {                                                                                                         \
        hipError_t e__ = (expr);                                                                                 \
        if (e__ != hipSuccess) {                                                                                 \
            cleanup();                                                                                           \
            return msosm::set_error(MSOSM_ERR_HIP, "refsim: %s failed: %s", #expr, hipGetErrorString(e__));      \
        }                                                                                                        \
    } while (0)

struct Geometry {
    unsigned long Np, Nd, range, order, fftpoint;
};

// simulated_complex_signal.cpp:21-25, 66-68: float members, double expressions
Geometry geometry(float bw, float dm, float f0, float period)
{
    const float fs = bw;
    const float kdm = (float)(4.15 * 1e15);
    Geometry g;
    g.Np = (unsigned long)(period * fs);
    g.Nd = (unsigned long)std::floor((double)kdm * dm * (1 / std::pow((double)f0, 2) - 1 / std::pow((double)(f0 + bw), 2)) /
                                     (1 / (double)fs));
    g.range = g.Np ? (unsigned long)std::ceil((float)g.Nd / (float)g.Np) : 0;
    g.order = (unsigned long)std::pow(2.0, std::ceil(std::log2((double)(g.range * g.Np))));
    g.fftpoint = 2 * g.order;
    return g;
}

__global__ void k_conj_scale(float2 *x, unsigned long n, float s)
{
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long)gridDim.x * blockDim.x)
        x[i] = make_float2(x[i].x * s, -x[i].y * s);
}

__global__ void k_scale(float2 *x, unsigned long n, float s)
{
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long)gridDim.x * blockDim.x)
        x[i] = make_float2(x[i].x * s, x[i].y * s);
}

__global__ void k_fill(float2 *x, unsigned long n)
{
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long)gridDim.x * blockDim.x)
        x[i] = make_float2(0.f, 0.f);
}

__global__ void k_two_impulses(float2 *x, unsigned long a, unsigned long b)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        x[a] = make_float2(1.f, 0.f);
        x[b] = make_float2(1.f, 0.f);
    }
}

// y = x * h (float products and sums, as the reference's loop), then
// accumulate the first Np samples' IDFT later
__global__ void k_mul(float2 *x, const float2 *h, unsigned long n)
{
#pragma clang fp contract(off)
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long)gridDim.x * blockDim.x) {
        const float2 a = x[i], b = h[i];
        x[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
}

__global__ void k_accumulate(float2 *acc, const float2 *x, unsigned long n, float s)
{
#pragma clang fp contract(off)
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long)gridDim.x * blockDim.x) {
        const float2 v = x[i];
        acc[i] = make_float2(acc[i].x + v.x * s, acc[i].y + v.y * s);
    }
}

unsigned grid(unsigned long n)
{
    unsigned long g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : g > 65536 ? 65536 : g);
}

int ilog2u(unsigned long v)
{
    int l = 0;
    while ((1ul << (l + 1)) <= v)
        l++;
    return l;
}

} // namespace

extern "C" {

int msosm_refsim_info(float bw, float dm, float f0, float period, unsigned long *Np, unsigned long *Nd,
                      unsigned long *range, unsigned long *fftpoint)
{
    if (!(bw > 0) || !(f0 > 0) || !(period > 0) || dm < 0)
        return msosm::set_error(MSOSM_ERR_ARG, "refsim_info: bad parameters");
    const Geometry g = geometry(bw, dm, f0, period);
    if (Np)
        *Np = g.Np;
    if (Nd)
        *Nd = g.Nd;
    if (range)
        *range = g.range;
    if (fftpoint)
        *fftpoint = g.fftpoint;
    return MSOSM_OK;
}

int msosm_refsim_awgn(float snr_db, unsigned long n, float *out)
{
    if (!out)
        return msosm::set_error(MSOSM_ERR_ARG, "refsim_awgn: NULL output");
    // AWGN.cpp:21-28, 29-45
    const float sigma = (float)std::sqrt(std::pow(10, (-snr_db / 10)));
    std::default_random_engine engine;
    std::normal_distribution<float> re(0.0f, 1.0f), im(0.0f, 1.0f);
    for (unsigned long i = 0; i < n; i++) {
        out[2 * i] = (float)(sigma / std::sqrt(2) * re(engine));
        out[2 * i + 1] = (float)(sigma / std::sqrt(2) * im(engine));
    }
    return MSOSM_OK;
}

int msosm_refsim_generate(float bw, float dm, float f0, float period, unsigned long repeat, int add_noise,
                          float snr_db, float *signal, void *signal_u16, float *signal_power_db)
{
    if (!(bw > 0) || !(f0 > 0) || !(period > 0) || dm < 0 || repeat < 1 || (!signal && !signal_u16))
        return msosm::set_error(MSOSM_ERR_ARG, "refsim_generate: bad argument");
    const Geometry g = geometry(bw, dm, f0, period);
    if (g.Np < 3 || g.range < 1 || g.fftpoint > (1ul << 28))
        return msosm::set_error(MSOSM_ERR_UNSUPPORTED, "refsim_generate: period of %lu samples with %lu-point "
                                                       "transforms is outside the supported range",
                                g.Np, g.fftpoint);
    const unsigned long Np = g.Np, order = g.order, N = g.fftpoint;
    float2 *c = nullptr, *H = nullptr, *x = nullptr, *y = nullptr, *acc = nullptr, *scratch = nullptr;
    msosm::ChirpParams *prm_d = nullptr;
    hipStream_t s = nullptr;
    auto cleanup = [&]() {
        if (s)
            (void)hipStreamSynchronize(s);
        for (void *p : {(void *)c, (void *)H, (void *)x, (void *)y, (void *)acc, (void *)scratch, (void *)prm_d})
            if (p)
                (void)hipFree(p);
        if (s)
            (void)hipStreamDestroy(s);
    };
    SHIP(hipStreamCreate(&s));
    SHIP(hipMalloc(&c, sizeof(float2) * order));
    SHIP(hipMalloc(&H, sizeof(float2) * N));
    SHIP(hipMalloc(&x, sizeof(float2) * N));
    SHIP(hipMalloc(&y, sizeof(float2) * N));
    SHIP(hipMalloc(&acc, sizeof(float2) * Np));
    SHIP(hipMalloc(&prm_d, sizeof(msosm::ChirpParams)));
    const size_t sc = msosm::fft_c2c_scratch_elems(ilog2u(N), 1);
    if (sc)
        SHIP(hipMalloc(&scratch, sizeof(float2) * sc));
    // H1 = exp(+i phi) on the float grid (simulated_complex_signal.cpp:82-101):
    // the planner's chirp, conjugated
    msosm::ChirpParams prm;
    prm.fs = bw;
    prm.step = (float)(2 * 3.14159265358979323846 * (double)bw / (double)order);
    prm.w0 = (float)(2 * 3.14159265358979323846 * (double)f0);
    prm.kdm = (float)(4.15 * 1e15);
    prm.dm = dm;
    prm.f0 = f0;
    SHIP(hipMemcpyAsync(prm_d, &prm, sizeof prm, hipMemcpyHostToDevice, s));
    SHIP(msosm::launch_chirp(c, order, MSOSM_CHIRP_REF_F32W, prm_d, 1, s));
    hipLaunchKernelGGL(k_conj_scale, dim3(grid(order)), dim3(256), 0, s, c, order, 1.0f);
    SHIP(hipGetLastError());
    // IDFT / order into the second half of H's input, DFT (:102-120)
    hipLaunchKernelGGL(k_fill, dim3(grid(N)), dim3(256), 0, s, x, N);
    SHIP(msosm::launch_fft_c2c(c, x + order, ilog2u(order), 1, order, order, true, 1.0f / (float)order, scratch, s));
    SHIP(msosm::launch_fft_c2c(x, H, ilog2u(N), 1, N, N, false, 1.0f, scratch, s));
    // shifted impulses through H, first Np samples summed (:128-158)
    hipLaunchKernelGGL(k_fill, dim3(grid(Np)), dim3(256), 0, s, acc, Np);
    for (unsigned long i = 0; i < g.range; i++) {
        hipLaunchKernelGGL(k_fill, dim3(grid(N)), dim3(256), 0, s, x, N);
        hipLaunchKernelGGL(k_two_impulses, dim3(1), dim3(64), 0, s, x, i * Np + Np / 2, i * Np + Np / 2 + 2);
        SHIP(msosm::launch_fft_c2c(x, y, ilog2u(N), 1, N, N, false, 1.0f, scratch, s));
        hipLaunchKernelGGL(k_mul, dim3(grid(N)), dim3(256), 0, s, y, H, N);
        SHIP(hipGetLastError());
        SHIP(msosm::launch_fft_c2c(y, x, ilog2u(N), 1, N, N, true, 1.0f, scratch, s));
        hipLaunchKernelGGL(k_accumulate, dim3(grid(Np)), dim3(256), 0, s, acc, x, Np, 1.0f / (float)N);
        SHIP(hipGetLastError());
    }
    std::vector<float> period_sig(2 * Np);
    SHIP(hipMemcpyAsync(period_sig.data(), acc, sizeof(float2) * Np, hipMemcpyDeviceToHost, s));
    SHIP(hipStreamSynchronize(s));
    cleanup();

    // measured power in dB (:165-170, float accumulation)
    float power = 0;
    for (unsigned long i = 0; i < Np; i++)
        power = power + period_sig[2 * i] * period_sig[2 * i] + period_sig[2 * i + 1] * period_sig[2 * i + 1];
    power = 10 * std::log10(power / Np);
    if (signal_power_db)
        *signal_power_db = power;

    const unsigned long total = repeat * Np;
    std::vector<float> noise;
    if (add_noise) {
        noise.resize(2 * total);
        msosm_refsim_awgn(snr_db - power, total, noise.data());
    }
    uint16_t *u = static_cast<uint16_t *>(signal_u16);
    auto to_u16 = [](float v) -> uint16_t { return v == 0 ? 0 : static_cast<uint16_t>(v * 32767.0f + 32768.0f); };
    for (unsigned long r = 0; r < repeat; r++) {
        for (unsigned long j = 0; j < Np; j++) {
            const unsigned long k = r * Np + j;
            float re = period_sig[2 * j], im = period_sig[2 * j + 1];
            if (add_noise) {
                re = re + noise[2 * k];
                im = im + noise[2 * k + 1];
            }
            if (signal) {
                signal[2 * k] = re;
                signal[2 * k + 1] = im;
            }
            if (u) {
                u[2 * k] = to_u16(re);
                u[2 * k + 1] = to_u16(im);
            }
        }
    }
    return MSOSM_OK;
}

} // extern "C"
