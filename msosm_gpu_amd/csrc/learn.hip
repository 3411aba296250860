// learn.hip -- the learned DM-proposal path: synthetic pulsar generation,
// the incoherent spectrogram front end, and forward/backward kernels of a
// small 1D-convolution network (FP16 weights, FP32 arithmetic and master
// weights), plus its optimiser.
//
// The reference has no counterpart (it searches a DM grid exhaustively with
// one engine per trial set, test_processing_speed_msosm_dmtrial.cpp); this
// component ranks DM trials from a cheap incoherent spectrogram so that only
// the top-k get coherent dedispersion.
//
// Tensor layouts: activations [B][C][T] row-major FP32; conv weights
// [Cout][Cin][KW] FP16 (plus an FP32 master copy for the optimiser); linear
// weights [O][I] FP16.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cmath>

#include "common.h"
#include "fft_device.h"
#include "kernels.h"
#include "msosm_c.h"

namespace {

using msosm::cmul;

#define LHIP(expr)                                                                                               \
    do {                                                                                                         \
        hipError_t e__ = (expr);                                                                                 \
        if (e__ != hipSuccess)                                                                                   \
            return msosm::set_error(MSOSM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e__));              \
    } while (0)

int gridn(size_t n, int block = 256)
{
    size_t g = (n + block - 1) / block;
    if (g < 1)
        g = 1;
    if (g > 65536)
        g = 65536;
    return (int)g;
}

// ---------------------------------------------------------------------------
// Counter-based normal deviates (hash of (seed, index) -> two uniforms ->
// Box-Muller); reproducible for any launch shape.
__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float2 normal2(uint32_t seed, uint64_t i)
{
    uint32_t a = hash32((uint32_t)i ^ hash32(seed ^ (uint32_t)(i >> 32) * 0x9E3779B9u));
    uint32_t b = hash32(a ^ 0x85ebca6bu);
    float u1 = ((a >> 8) + 1) * (1.0f / 16777217.0f); // (0, 1]
    float u2 = (b >> 8) * (1.0f / 16777216.0f);
    float r = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincospif(2.0f * u2, &s, &c);
    return make_float2(r * c, r * s);
}

// s_n = sqrt(amp * g(phase_n)) * z_n, z ~ CN(0,1), phase in FP64.
__global__ void k_sim_pulsar(float2 *__restrict__ x, size_t n, double inv_period, double two_sigma2, float amp,
                             uint32_t seed)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double ph = (double)i * inv_period;
        ph -= floor(ph);
        double dphi = ph - 0.5;
        float g = (float)exp(-dphi * dphi / two_sigma2);
        float2 z = normal2(seed, i);
        float a = sqrtf(amp * g) * 0.70710678f;
        x[i] = make_float2(a * z.x, a * z.y);
    }
}

// X[b] *= exp(sign * i * phi(f_b)), phi(f) = 2 pi D DM f^2 / (f0^2 (f0 + f)),
// f = offset of bin b from the band bottom; the phase is reduced in FP64.
__global__ void k_sim_disperse(float2 *__restrict__ X, size_t n, double fs, double f0, double dm, int sign)
{
    const double PI = 3.14159265358979323846;
    for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < n; b += (size_t)gridDim.x * blockDim.x) {
        double nu = (b < n / 2 ? (double)b : (double)b - (double)n) * (fs / (double)n);
        double f = nu + fs / 2;
        double ph = 2 * PI * 4.15e15 * dm * f * f / (f0 * f0 * (f0 + f));
        double s, c;
        sincos(ph, &s, &c);
        float2 w = make_float2((float)c, (float)(sign * s));
        X[b] = cmul(X[b], w);
    }
}

// u16 = round(scale * (x / n_scale + noise)) + 32768, clipped to [1, 65535]
__global__ void k_sim_quantize(const float2 *__restrict__ x, size_t n, float x_scale, uint32_t seed, float scale,
                               uint32_t *__restrict__ out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float2 z = normal2(seed ^ 0xA5A5A5A5u, i);
        float2 v = x[i];
        float re = (v.x * x_scale + z.x * 0.70710678f) * scale;
        float im = (v.y * x_scale + z.y * 0.70710678f) * scale;
        float a = fminf(fmaxf(rintf(re) + 32768.0f, 1.0f), 65535.0f);
        float b = fminf(fmaxf(rintf(im) + 32768.0f, 1.0f), 65535.0f);
        out[i] = (uint32_t)a | ((uint32_t)b << 16);
    }
}

// ---------------------------------------------------------------------------
// Spectrogram: per time bin, navg consecutive nchan-point spectra of the
// unpacked uint16 input, |X|^2 summed; channel 0 = lowest frequency (FFT bin
// nchan/2).  One workgroup per time bin, nchan = 256, navg = 16: the 16
// transforms of 256 points run as one FFT workgroup of 16 x 16 threads.
__device__ __forceinline__ float2 unpack1(uint32_t u)
{
    uint32_t x = u & 0xFFFFu, y = u >> 16;
    return make_float2((float)x - (x != 0 ? 32768.0f : 0.0f), (float)y - (y != 0 ? 32768.0f : 0.0f));
}

__global__ __launch_bounds__(256) void k_spectrogram256(const uint32_t *__restrict__ in, int T,
                                                         float *__restrict__ out)
{
    constexpr int LOG2N = 8, E = 16;
    using S = msosm::FftShape<LOG2N, E>;
    static_assert(S::NT == 16, "16 threads per 256-point transform");
    __shared__ float2 smem[16 * S::NPAD + S::TWLO + S::TWHI];
    __shared__ float pw[16][257];
    float2 *twlo = smem + 16 * S::NPAD, *twhi = twlo + S::TWLO;
    msosm::fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const int tl = threadIdx.x / 16, t = threadIdx.x % 16;
    const int tb = blockIdx.x;
    const uint32_t *src = in + ((size_t)tb * 16 + tl) * 256;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++)
        v[m] = unpack1(src[t + m * 16]);
    msosm::fft_block<LOG2N, E, false>(v, smem + tl * S::NPAD, twlo, twhi, t);
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int k = t + m * 16;
        pw[tl][(k + 128) & 255] = v[m].x * v[m].x + v[m].y * v[m].y;
    }
    __syncthreads();
    const int c = threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 16; j++)
        s += pw[j][c];
    out[(size_t)c * T + tb] = s;
}

// rows of length T -> (x - mean) / std (one workgroup per row)
__global__ void k_rows_standardize(float *__restrict__ x, int T)
{
    __shared__ double red[2][256];
    float *row = x + (size_t)blockIdx.x * T;
    double s = 0, ss = 0;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        double v = row[t];
        s += v;
        ss += v * v;
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = ss;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    const double mean = red[0][0] / T;
    const double var = red[1][0] / T - mean * mean;
    const float inv = (float)(var > 0 ? 1.0 / sqrt(var) : 0.0);
    const float mu = (float)mean;
    for (int t = threadIdx.x; t < T; t += blockDim.x)
        row[t] = (row[t] - mu) * inv;
}

// out[k][t] = sum_c S[c][t + delay[k][c]] / sqrt(nchan), t < Tout; a
// delayed index outside [0, T) contributes nothing (the host keeps every
// delay within T - Tout, so this only guards against a bad table).
__global__ void k_incoherent(const float *__restrict__ S, int nchan, int T, const int *__restrict__ delay, int Tout,
                             float *__restrict__ out)
{
    const int k = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Tout)
        return;
    const int *dk = delay + (size_t)k * nchan;
    float s = 0.f;
    for (int c = 0; c < nchan; c++) {
        const int ts = t + dk[c];
        if ((unsigned)ts < (unsigned)T)
            s += S[(size_t)c * T + ts];
    }
    out[(size_t)k * Tout + t] = s * rsqrtf((float)nchan);
}

// out[r][b] = scale * mean of x[r][t] over the t whose phase frac(t / period)
// falls in bin b of nbins (FP64 phase), 0 for a bin no sample reaches.  One
// workgroup per row, sums and counts in LDS.
__global__ __launch_bounds__(256) void k_rows_fold(const float *__restrict__ x, int T, double inv_period, int nbins,
                                                   float scale, float *__restrict__ out)
{
    extern __shared__ float fold_lds[];
    float *sum = fold_lds, *cnt = fold_lds + nbins;
    for (int b = threadIdx.x; b < 2 * nbins; b += blockDim.x)
        fold_lds[b] = 0.f;
    __syncthreads();
    const float *row = x + (size_t)blockIdx.x * T;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        double ph = (double)t * inv_period;
        ph -= floor(ph);
        int b = (int)(ph * nbins);
        b = b < nbins ? b : nbins - 1;
        atomicAdd(&sum[b], row[t]);
        atomicAdd(&cnt[b], 1.0f);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        out[(size_t)blockIdx.x * nbins + b] = cnt[b] > 0.f ? scale * sum[b] / cnt[b] : 0.f;
}

// Incoherent dedispersion of folded channels: with Sf[c][b] channel c folded
// at the period into nbins phase bins, trial k's folded profile is
// out[k][b] = sum_c Sf[c][(b + s) mod nbins] / sqrt(nchan), s = the channel's
// delay delay_s[k][c] (seconds) as a whole number of phase bins,
// rint(delay / period * nbins) (FP64).  Folding is linear, so this is the
// fold of the incoherently dedispersed series up to that rounding.  One
// thread per (trial, bin).
__global__ void k_incoherent_fold(const float *__restrict__ Sf, int nchan, int nbins,
                                  const float *__restrict__ delay_s, double inv_period, float *__restrict__ out)
{
    const int k = blockIdx.y;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins)
        return;
    const float *dk = delay_s + (size_t)k * nchan;
    float s = 0.f;
    for (int c = 0; c < nchan; c++) {
        const double sh = rint((double)dk[c] * inv_period * nbins);
        long long q = (long long)sh % nbins;
        q += q < 0 ? nbins : 0;
        int i = b + (int)q;
        i -= i >= nbins ? nbins : 0;
        s += Sf[(size_t)c * nbins + i];
    }
    out[(size_t)k * nbins + b] = s * rsqrtf((float)nchan);
}

// out[k][r] = |sum_b p[r][b] exp(-2 pi i (k+1) b / nbins)| / sqrt(nbins) for
// k < K: amplitudes of the first K harmonics of each row's folded profile
// (independent of where in phase the pulse sits), harmonic-major so that the
// rows become the time axis of a convolution over trials.  One thread per
// (row, harmonic); the angle is reduced exactly as ((k+1) b mod nbins).
__global__ void k_rows_harmonics(const float *__restrict__ p, int R, int nbins, int K, float *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (r >= R)
        return;
    const float *row = p + (size_t)r * nbins;
    float re = 0.f, im = 0.f;
    const int h = k + 1;
    for (int b = 0; b < nbins; b++) {
        const int e = (h * b) % nbins;
        float s, c;
        sincospif(2.0f * (float)e / (float)nbins, &s, &c);
        re += row[b] * c;
        im -= row[b] * s;
    }
    out[(size_t)k * R + r] = sqrtf(re * re + im * im) * rsqrtf((float)nbins);
}

// ---------------------------------------------------------------------------
// 1D convolution, same padding: y[b][co][t] = bias[co] +
//   sum_ci sum_k w[co][ci][k] x[b][ci][t + dil (k - KW/2)]   (ReLU optional).
// With TRANSPOSE the roles swap for the data gradient:
//   gx[b][ci][t] = sum_co sum_k w[co][ci][k] g[b][co][t - dil (k - KW/2)],
// g = gy * relu'(y) read from gy and y.
// Tile: 16 output channels x 64 time steps per workgroup of 256 threads,
// input rows staged in LDS 16 channels at a time, weights converted to FP32
// in LDS.
template <bool TRANSPOSE>
__global__ __launch_bounds__(256) void k_conv1d(const float *__restrict__ x, const float *__restrict__ ymask,
                                                const __half *__restrict__ w, const float *__restrict__ bias,
                                                float *__restrict__ y, int Cin, int Cout, int T, int KW, int dil,
                                                int relu)
{
    // in this kernel "in" channels are the ones summed over, "out" the ones produced
    const int CI = TRANSPOSE ? Cout : Cin; // summed channels
    const int CO = TRANSPOSE ? Cin : Cout; // produced channels
    const int b = blockIdx.z;
    const int co0 = blockIdx.y * 16;
    const int t0 = blockIdx.x * 64;
    const int half = KW / 2;
    const int halo = half * dil;
    __shared__ float xs[16][64 + 2 * 64];     // 16 channels x (64 + 2*halo), halo <= 64
    __shared__ float ws[16][16][16];          // [co][ci][k], KW <= 16
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16; // ty: output channel in tile
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    const int width = 64 + 2 * halo;
    for (int ci0 = 0; ci0 < CI; ci0 += 16) {
        __syncthreads();
        for (int i = threadIdx.x; i < 16 * width; i += 256) {
            const int c = i / width, tt = i % width;
            const int ci = ci0 + c, ts = t0 + tt - halo;
            float v = 0.f;
            if (ci < CI && ts >= 0 && ts < T) {
                const size_t off = ((size_t)b * CI + ci) * T + ts;
                v = x[off];
                if (TRANSPOSE && relu && ymask[off] <= 0.f)
                    v = 0.f;
            }
            xs[c][tt] = v;
        }
        for (int i = threadIdx.x; i < 16 * 16 * KW; i += 256) {
            const int o = i / (16 * KW), r = i % (16 * KW), c = r / KW, k = r % KW;
            const int oc = co0 + o, ic = ci0 + c;
            float v = 0.f;
            if (oc < CO && ic < CI) {
                // forward weight index [co][ci][k]; transposed: produced = ci of
                // the forward layer, summed = co, tap mirrored
                v = TRANSPOSE ? __half2float(w[((size_t)ic * Cin + oc) * KW + (KW - 1 - k)])
                              : __half2float(w[((size_t)oc * Cin + ic) * KW + k]);
            }
            ws[o][c][k] = v;
        }
        __syncthreads();
        const int nci = min(16, CI - ci0);
        for (int c = 0; c < nci; c++) {
            for (int k = 0; k < KW; k++) {
                const float wv = ws[ty][c][k];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[j] += wv * xs[c][tx + 16 * j + k * dil];
            }
        }
    }
    const int co = co0 + ty;
    if (co >= CO)
        return;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int t = t0 + tx + 16 * j;
        if (t < T) {
            float v = acc[j];
            if (!TRANSPOSE) {
                v += bias[co];
                if (relu)
                    v = fmaxf(v, 0.f);
            }
            y[((size_t)b * CO + co) * T + t] = v;
        }
    }
}

// Weight and bias gradients: gw[co][ci][k] += sum_b sum_t g[b][co][t]
// x[b][ci][t + dil (k - KW/2)], gb[co] += sum g, g = gy * relu'(y).
// One workgroup per (co, ci); 256 threads reduce over (b, t).
__global__ __launch_bounds__(256) void k_conv1d_wgrad(const float *__restrict__ x, const float *__restrict__ y,
                                                      const float *__restrict__ gy, float *__restrict__ gw,
                                                      float *__restrict__ gb, int B, int Cin, int Cout, int T,
                                                      int KW, int dil, int relu)
{
    const int co = blockIdx.y, ci = blockIdx.x;
    const int half = KW / 2;
    float acc[16];
#pragma unroll
    for (int k = 0; k < 16; k++)
        acc[k] = 0.f;
    float accb = 0.f;
    for (int b = 0; b < B; b++) {
        const float *grow = gy + ((size_t)b * Cout + co) * T;
        const float *yrow = y + ((size_t)b * Cout + co) * T;
        const float *xrow = x + ((size_t)b * Cin + ci) * T;
        for (int t = threadIdx.x; t < T; t += blockDim.x) {
            float g = grow[t];
            if (relu && yrow[t] <= 0.f)
                g = 0.f;
            accb += g;
#pragma unroll
            for (int k = 0; k < 16; k++) {
                if (k < KW) {
                    const int ts = t + dil * (k - half);
                    if (ts >= 0 && ts < T)
                        acc[k] += g * xrow[ts];
                }
            }
        }
    }
    __shared__ float red[256];
    for (int k = 0; k < KW + 1; k++) {
        red[threadIdx.x] = k < KW ? acc[k] : accb;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
                red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (k < KW)
                gw[((size_t)co * Cin + ci) * KW + k] += red[0];
            else if (ci == 0)
                gb[co] += red[0];
        }
        __syncthreads();
    }
}

// feat[b][c] = max_t x, feat[b][C + c] = mean_t x; argmax kept for backward
__global__ void k_pool_fwd(const float *__restrict__ x, int C, int T, float *__restrict__ feat, int *__restrict__ amax)
{
    const int b = blockIdx.y, c = blockIdx.x;
    const float *row = x + ((size_t)b * C + c) * T;
    __shared__ float mv[256], sv[256];
    __shared__ int mi[256];
    float m = -INFINITY, s = 0.f;
    int im = 0;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        float v = row[t];
        s += v;
        if (v > m) {
            m = v;
            im = t;
        }
    }
    mv[threadIdx.x] = m;
    sv[threadIdx.x] = s;
    mi[threadIdx.x] = im;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sv[threadIdx.x] += sv[threadIdx.x + w];
            const float o = mv[threadIdx.x + w];
            const int oi = mi[threadIdx.x + w];
            if (o > mv[threadIdx.x] || (o == mv[threadIdx.x] && oi < mi[threadIdx.x])) {
                mv[threadIdx.x] = o;
                mi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        feat[(size_t)b * 2 * C + c] = mv[0];
        feat[(size_t)b * 2 * C + C + c] = sv[0] / T;
        amax[(size_t)b * C + c] = mi[0];
    }
}

__global__ void k_pool_bwd(const float *__restrict__ gfeat, const int *__restrict__ amax, int C, int T,
                           float *__restrict__ gx)
{
    const int b = blockIdx.y, c = blockIdx.x;
    const float gmax = gfeat[(size_t)b * 2 * C + c];
    const float gmean = gfeat[(size_t)b * 2 * C + C + c] / T;
    const int am = amax[(size_t)b * C + c];
    float *row = gx + ((size_t)b * C + c) * T;
    for (int t = threadIdx.x; t < T; t += blockDim.x)
        row[t] = gmean + (t == am ? gmax : 0.f);
}

// y[b][o] = bias[o] + sum_i w[o][i] x[b][i]
__global__ void k_linear_fwd(const float *__restrict__ x, const __half *__restrict__ w, const float *__restrict__ bias,
                             float *__restrict__ y, int B, int I, int O)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * O)
        return;
    const int b = idx / O, o = idx % O;
    float s = bias[o];
    for (int i = 0; i < I; i++)
        s += __half2float(w[(size_t)o * I + i]) * x[(size_t)b * I + i];
    y[idx] = s;
}

__global__ void k_linear_bwd_x(const float *__restrict__ gy, const __half *__restrict__ w, float *__restrict__ gx,
                               int B, int I, int O)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * I)
        return;
    const int b = idx / I, i = idx % I;
    float s = 0.f;
    for (int o = 0; o < O; o++)
        s += __half2float(w[(size_t)o * I + i]) * gy[(size_t)b * O + o];
    gx[idx] = s;
}

__global__ void k_linear_bwd_w(const float *__restrict__ x, const float *__restrict__ gy, float *__restrict__ gw,
                               float *__restrict__ gb, int B, int I, int O)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= O * (I + 1))
        return;
    const int o = idx / (I + 1), i = idx % (I + 1);
    float s = 0.f;
    for (int b = 0; b < B; b++)
        s += gy[(size_t)b * O + o] * (i < I ? x[(size_t)b * I + i] : 1.0f);
    if (i < I)
        gw[(size_t)o * I + i] += s;
    else
        gb[o] += s;
}

// mean over the batch of -sum_o target[b][o] log softmax(logits[b])[o];
// glogits = (softmax - target) / B
__global__ void k_softmax_xent(const float *__restrict__ logits, const float *__restrict__ target, int O,
                               float *__restrict__ loss, float *__restrict__ glogits, int B)
{
    const int b = blockIdx.x;
    const float *lr = logits + (size_t)b * O;
    const float *tr = target + (size_t)b * O;
    __shared__ float red[256];
    float m = -INFINITY;
    for (int o = threadIdx.x; o < O; o += blockDim.x)
        m = fmaxf(m, lr[o]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    m = red[0];
    __syncthreads();
    float s = 0.f;
    for (int o = threadIdx.x; o < O; o += blockDim.x)
        s += expf(lr[o] - m);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const float lse = m + logf(red[0]);
    __syncthreads();
    float l = 0.f;
    for (int o = threadIdx.x; o < O; o += blockDim.x) {
        const float lp = lr[o] - lse;
        l -= tr[o] * lp;
        glogits[(size_t)b * O + o] = (expf(lp) - tr[o]) / (float)B;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(loss, red[0] / (float)B);
}

// Adam on FP32 master weights; refreshes the FP16 copy the kernels read.
__global__ void k_adam(float *__restrict__ p, __half *__restrict__ p16, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        if (p16)
            p16[i] = __float2half_rn(pi);
    }
}

__global__ void k_f32_to_f16(const float *__restrict__ a, __half *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = __float2half_rn(a[i]);
}

} // namespace

extern "C" {

int msosm_sim_pulsar(void *x_c64, size_t n, double period_samples, double duty, float amp, unsigned int seed,
                     void *stream)
{
    if (!x_c64 || n == 0 || !(period_samples > 0) || !(duty > 0))
        return msosm::set_error(MSOSM_ERR_ARG, "sim_pulsar: bad argument");
    const double sigma = duty / 2.3548;
    hipLaunchKernelGGL(k_sim_pulsar, dim3(gridn(n)), dim3(256), 0, (hipStream_t)stream, (float2 *)x_c64, n,
                       1.0 / period_samples, 2 * sigma * sigma, amp, (uint32_t)seed);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_sim_disperse(void *X_c64, size_t n, double fs, double f0, double dm, int sign, void *stream)
{
    if (!X_c64 || n == 0 || !(fs > 0) || !(f0 > 0) || (sign != 1 && sign != -1))
        return msosm::set_error(MSOSM_ERR_ARG, "sim_disperse: bad argument");
    hipLaunchKernelGGL(k_sim_disperse, dim3(gridn(n)), dim3(256), 0, (hipStream_t)stream, (float2 *)X_c64, n, fs, f0,
                       dm, sign);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_sim_quantize(const void *x_c64, size_t n, float x_scale, unsigned int seed, float scale, void *u16_out,
                       void *stream)
{
    if (!x_c64 || !u16_out || n == 0)
        return msosm::set_error(MSOSM_ERR_ARG, "sim_quantize: bad argument");
    hipLaunchKernelGGL(k_sim_quantize, dim3(gridn(n)), dim3(256), 0, (hipStream_t)stream, (const float2 *)x_c64, n,
                       x_scale, (uint32_t)seed, scale, (uint32_t *)u16_out);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_spectrogram_u16(const void *u16, size_t n, int nchan, int navg, float *out, void *stream)
{
    if (!u16 || !out || nchan != 256 || navg != 16 || n < (size_t)nchan * navg)
        return msosm::set_error(MSOSM_ERR_UNSUPPORTED, "spectrogram: the library supports 256 channels x 16 spectra");
    const int T = (int)(n / ((size_t)nchan * navg));
    hipLaunchKernelGGL(k_spectrogram256, dim3(T), dim3(256), 0, (hipStream_t)stream, (const uint32_t *)u16, T, out);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_rows_standardize(float *x, int rows, int T, void *stream)
{
    if (!x || rows < 1 || T < 2)
        return msosm::set_error(MSOSM_ERR_ARG, "rows_standardize: bad argument");
    hipLaunchKernelGGL(k_rows_standardize, dim3(rows), dim3(256), 0, (hipStream_t)stream, x, T);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_incoherent_dedisperse(const float *S, int nchan, int T, const int *delay, int K, int Tout, float *out,
                                void *stream)
{
    if (!S || !delay || !out || nchan < 1 || K < 1 || Tout < 1 || Tout > T || K > 65535)
        return msosm::set_error(MSOSM_ERR_ARG, "incoherent_dedisperse: bad argument");
    hipLaunchKernelGGL(k_incoherent, dim3((Tout + 255) / 256, K), dim3(256), 0, (hipStream_t)stream, S, nchan, T,
                       delay, Tout, out);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_rows_fold(const float *x, int rows, int T, double period, int nbins, float scale, float *out,
                    void *stream)
{
    if (!x || !out || rows < 1 || T < 1 || !(period > 0) || nbins < 1 || nbins > 4096)
        return msosm::set_error(MSOSM_ERR_ARG, "rows_fold: bad argument (period > 0, 1 <= nbins <= 4096)");
    hipLaunchKernelGGL(k_rows_fold, dim3(rows), dim3(256), sizeof(float) * 2 * nbins, (hipStream_t)stream, x, T,
                       1.0 / period, nbins, scale, out);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_incoherent_fold(const float *Sf, int nchan, int nbins, const float *delay_s, int K, double period,
                          float *out, void *stream)
{
    if (!Sf || !delay_s || !out || nchan < 1 || nbins < 1 || K < 1 || K > 65535 || !(period > 0))
        return msosm::set_error(MSOSM_ERR_ARG, "incoherent_fold: bad argument");
    hipLaunchKernelGGL(k_incoherent_fold, dim3((nbins + 63) / 64, K), dim3(64), 0, (hipStream_t)stream, Sf, nchan,
                       nbins, delay_s, 1.0 / period, out);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_rows_harmonics(const float *prof, int rows, int nbins, int K, float *out, void *stream)
{
    if (!prof || !out || rows < 1 || nbins < 2 || K < 1 || K > nbins / 2 || K > 65535)
        return msosm::set_error(MSOSM_ERR_ARG, "rows_harmonics: bad argument (1 <= K <= nbins / 2)");
    hipLaunchKernelGGL(k_rows_harmonics, dim3((rows + 255) / 256, K), dim3(256), 0, (hipStream_t)stream, prof, rows,
                       nbins, K, out);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_conv1d_fwd(const float *x, const void *w16, const float *bias, float *y, int B, int Cin, int Cout, int T,
                     int KW, int dil, int relu, void *stream)
{
    if (!x || !w16 || !bias || !y || B < 1 || Cin < 1 || Cout < 1 || T < 1 || KW < 1 || KW > 16 || (KW & 1) == 0 ||
        dil < 1 || (KW / 2) * dil > 64 || B > 65535)
        return msosm::set_error(MSOSM_ERR_ARG, "conv1d_fwd: bad shape (odd KW <= 15, halo <= 64)");
    dim3 grid((T + 63) / 64, (Cout + 15) / 16, B);
    hipLaunchKernelGGL((k_conv1d<false>), grid, dim3(256), 0, (hipStream_t)stream, x, nullptr, (const __half *)w16,
                       bias, y, Cin, Cout, T, KW, dil, relu);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_conv1d_bwd(const float *x, const void *w16, const float *y, const float *gy, float *gx, float *gw,
                     float *gb, int B, int Cin, int Cout, int T, int KW, int dil, int relu, void *stream)
{
    if (!x || !w16 || !y || !gy || !gw || !gb || B < 1 || Cin < 1 || Cout < 1 || KW > 16 || (KW & 1) == 0 ||
        (KW / 2) * dil > 64 || B > 65535)
        return msosm::set_error(MSOSM_ERR_ARG, "conv1d_bwd: bad shape");
    if (gx) {
        dim3 grid((T + 63) / 64, (Cin + 15) / 16, B);
        hipLaunchKernelGGL((k_conv1d<true>), grid, dim3(256), 0, (hipStream_t)stream, gy, y, (const __half *)w16,
                           nullptr, gx, Cin, Cout, T, KW, dil, relu);
        LHIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_conv1d_wgrad, dim3(Cin, Cout), dim3(256), 0, (hipStream_t)stream, x, y, gy, gw, gb, B, Cin,
                       Cout, T, KW, dil, relu);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_pool_fwd(const float *x, int B, int C, int T, float *feat, int *amax, void *stream)
{
    if (!x || !feat || !amax || B < 1 || C < 1 || T < 1 || B > 65535)
        return msosm::set_error(MSOSM_ERR_ARG, "pool_fwd: bad argument");
    hipLaunchKernelGGL(k_pool_fwd, dim3(C, B), dim3(256), 0, (hipStream_t)stream, x, C, T, feat, amax);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_pool_bwd(const float *gfeat, const int *amax, int B, int C, int T, float *gx, void *stream)
{
    if (!gfeat || !amax || !gx || B < 1 || C < 1 || T < 1 || B > 65535)
        return msosm::set_error(MSOSM_ERR_ARG, "pool_bwd: bad argument");
    hipLaunchKernelGGL(k_pool_bwd, dim3(C, B), dim3(256), 0, (hipStream_t)stream, gfeat, amax, C, T, gx);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_linear_fwd(const float *x, const void *w16, const float *bias, float *y, int B, int I, int O, void *stream)
{
    if (!x || !w16 || !bias || !y || B < 1 || I < 1 || O < 1)
        return msosm::set_error(MSOSM_ERR_ARG, "linear_fwd: bad argument");
    hipLaunchKernelGGL(k_linear_fwd, dim3((B * O + 255) / 256), dim3(256), 0, (hipStream_t)stream, x,
                       (const __half *)w16, bias, y, B, I, O);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_linear_bwd(const float *x, const void *w16, const float *gy, float *gx, float *gw, float *gb, int B, int I,
                     int O, void *stream)
{
    if (!x || !w16 || !gy || !gw || !gb || B < 1 || I < 1 || O < 1)
        return msosm::set_error(MSOSM_ERR_ARG, "linear_bwd: bad argument");
    if (gx) {
        hipLaunchKernelGGL(k_linear_bwd_x, dim3((B * I + 255) / 256), dim3(256), 0, (hipStream_t)stream, gy,
                           (const __half *)w16, gx, B, I, O);
        LHIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_linear_bwd_w, dim3((O * (I + 1) + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, gy, gw,
                       gb, B, I, O);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_softmax_xent(const float *logits, const float *target, int B, int O, float *loss, float *glogits,
                       void *stream)
{
    if (!logits || !target || !loss || !glogits || B < 1 || O < 1)
        return msosm::set_error(MSOSM_ERR_ARG, "softmax_xent: bad argument");
    hipLaunchKernelGGL(k_softmax_xent, dim3(B), dim3(256), 0, (hipStream_t)stream, logits, target, O, loss, glogits, B);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_adam(float *p, void *p16, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
               float eps, int step, void *stream)
{
    if (!p || !g || !m || !v || n == 0 || step < 1)
        return msosm::set_error(MSOSM_ERR_ARG, "adam: bad argument");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(gridn(n)), dim3(256), 0, (hipStream_t)stream, p, (__half *)p16, g, m, v, n, lr, b1,
                       b2, eps, c1, c2);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_f32_to_f16(const float *a, void *b, size_t n, void *stream)
{
    if (!a || !b || n == 0)
        return msosm::set_error(MSOSM_ERR_ARG, "f32_to_f16: bad argument");
    hipLaunchKernelGGL(k_f32_to_f16, dim3(gridn(n)), dim3(256), 0, (hipStream_t)stream, a, (__half *)b, n);
    LHIP(hipGetLastError());
    return MSOSM_OK;
}

} // extern "C"
