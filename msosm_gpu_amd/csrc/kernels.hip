// kernels.hip -- hand-written CDNA4 (gfx950) kernels of the MS-OSM engine.
//
//  * element kernels with the semantics of the reference's gpu/gpu_kernel.cu
//    (unpack, pack, intensity, multiply, multiply-delay scatter, discard);
//  * batched C2C FFTs: one LDS pass up to 2^14 points, a two-pass four-step
//    (columns then rows, inter-pass twiddle in FP64) above;
//  * the fused forms: unpack + forward FFT of overlapped windows straight into
//    a shared spectra ring, and gather X_{g-d(n)}[n]*H[n] + inverse FFT +
//    discard + scale in one kernel;
//  * the FP64 chirp generator for the filter tables.
#include "kernels.h"
#include "mfma_fft.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>

#include "elems.h"
#include "fft_device.h"

namespace msosm {

static constexpr int kBlock = 256;

static inline int grid_for(size_t n, int per_thread = 1)
{
    size_t g = (n + (size_t)kBlock * per_thread - 1) / ((size_t)kBlock * per_thread);
    if (g < 1)
        g = 1;
    if (g > 1048576)
        g = 1048576; // grid-stride loops cover the rest
    return (int)g;
}

// ---------------------------------------------------------------------------
// Element kernels

// gpu_kernel.cu:317-329 -- truncating conversion after adding 32768 to
// non-zero values; saturates to [0, 65535], NaN -> 0.
__device__ __forceinline__ uint32_t pack_one(float v)
{
#pragma clang fp contract(off)
    float t = v + (v != 0.0f ? 32768.0f : 0.0f);
    uint32_t u;
    if (!(t > 0.0f))
        u = 0;
    else if (t >= 65535.0f)
        u = 65535;
    else
        u = (uint32_t)t;
    return u;
}

__global__ void k_unpack(float2 *__restrict__ dst, const uint32_t *__restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = unpack_u16pair(src[i]);
}

__global__ void k_pack(uint32_t *__restrict__ dst, const float2 *__restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float2 v = src[i];
        dst[i] = pack_one(v.x) | (pack_one(v.y) << 16);
    }
}

// gpu_kernel.cu:347-353 -- sqrt(|a|^2 + |b|^2), summed left to right, no
// fused multiply-add (so a CPU restatement reproduces it bit for bit).
__global__ void k_intensity(const float2 *__restrict__ a, const float2 *__restrict__ b, float *__restrict__ out,
                            size_t n)
{
#pragma clang fp contract(off)
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float2 p = a[i];
        float s = p.x * p.x;
        s = s + p.y * p.y;
        if (b) {
            float2 q = b[i];
            s = s + q.x * q.x;
            s = s + q.y * q.y;
        }
        out[i] = sqrtf(s); // correctly rounded (v_sqrt_f32 alone is not)
    }
}

// gpu_kernel.cu:56-66 -- out[dm][c][n] = (a[c][n]/N) * b[dm][n]
__global__ void k_multiply(const float2 *__restrict__ a, const float2 *__restrict__ b, float2 *__restrict__ out,
                           int N, int count, int ndm)
{
    const size_t per = (size_t)N * count, total = per * ndm;
    const float invN = 1.0f / (float)N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        size_t dm = i / per, r = i % per;
        float2 x = a[r];
        float2 h = b[dm * N + (r % N)];
        x = make_float2(x.x * invN, x.y * invN);
        out[i] = cmul(x, h);
    }
}

// gpu_kernel.cu:160-197 -- scatter the product of block c into ring slot
// j = (index + c + d[dm][n]) % dbs, layout [j/count][dm][j%count][N].
__global__ void k_multiply_delay(const float2 *__restrict__ a, const float2 *__restrict__ b, float2 *__restrict__ ring,
                                 const int32_t *__restrict__ delay, int dbs, int index, int N, int count, int ndm)
{
    const float invN = 1.0f / (float)N;
    const int dm = blockIdx.x % ndm;
    const int c = blockIdx.x / ndm;
    const float2 *arow = a + (size_t)c * N;
    const float2 *brow = b + (size_t)dm * N;
    const int32_t *drow = delay + (size_t)dm * N;
    for (int n = blockIdx.y * blockDim.x + threadIdx.x; n < N; n += gridDim.y * blockDim.x) {
        int j = (index + c + drow[n]) % dbs;
        size_t off = ((size_t)(j / count) * ndm + dm) * ((size_t)count * N) + (size_t)(j % count) * N + n;
        float2 x = arow[n], h = brow[n];
        float2 r = make_float2((x.x * h.x - x.y * h.y) * invN, (x.x * h.y + x.y * h.x) * invN);
        store_stream(&ring[off], r); // read back calls later, from HBM
    }
}

// gpu_kernel.cu:236-258 -- keep samples [M, 2M) of each inverse FFT, * 1/(2M)
__global__ void k_discard(const float2 *__restrict__ a, float2 *__restrict__ b, int M, size_t groups)
{
    const float scale = 1.0f / (float)(2 * M);
    const size_t total = groups * (size_t)M;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        size_t g = i / M, o = i % M;
        float2 v = a[g * 2 * M + M + o];
        store_stream(&b[i], make_float2(v.x * scale, v.y * scale));
    }
}

// ---------------------------------------------------------------------------
// Filter-table generation.
//
// Chirp C(w_i) in IDFT input order (msosm_prepare.cpp:50-63): the phase
// 4 pi^2 kdm dm w^2 / ((w + w0) w0^2) is evaluated in FP64.  mode 0 follows
// the reference's float-rounded grid (w_i and w_i + w0 rounded to float as
// its float arrays do); mode 1 uses an exact FP64 grid f = i*fs/M and
// phi = 2 pi D DM f^2 / (f0^2 (f0 + f)).
__global__ void k_chirp(float2 *__restrict__ c, unsigned long M, int mode, const ChirpParams *__restrict__ prm)
{
#pragma clang fp contract(off)
    const double PI = 3.14159265358979323846;
    const int d = blockIdx.y;
    const ChirpParams p = prm[d];
    float2 *cd = c + (size_t)d * M;
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < M;
         i += (unsigned long)gridDim.x * blockDim.x) {
        double ph;
        if (mode == 0) {
            float sw = p.step * (float)i;                       // float product
            float wi = (float)(-PI * (double)p.fs + (double)sw);
            wi = (float)((double)wi + PI * (double)p.fs);
            float wsum = wi + p.w0;                             // float sum
            ph = 4 * PI;
            ph = ph * PI;
            ph = ph * (double)p.kdm;
            ph = ph * (double)p.dm;
            ph = ph * (double)wi;
            ph = ph * (double)wi;
            ph = ph / (double)wsum;
            ph = ph / (double)p.w0;
            ph = ph / (double)p.w0;
        } else {
            double f = (double)p.fs * (double)i / (double)M;
            double F0 = (double)p.f0;
            ph = 2 * PI * 4.15e15 * (double)p.dm * f * f / (F0 * F0 * (F0 + f));
        }
        double s, co;
        sincos(ph, &s, &co);
        unsigned long idx = (i < M / 2) ? (M / 2 + i) : (i - M / 2);
        cd[idx] = make_float2((float)co, (float)(-s));
    }
}

__global__ void k_zero_pad(float2 *__restrict__ dst, const float2 *__restrict__ src, unsigned long M)
{
    const int d = blockIdx.y;
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < 2 * M;
         i += (unsigned long)gridDim.x * blockDim.x)
        dst[(size_t)d * 2 * M + i] = (i < M) ? src[(size_t)d * M + i] : make_float2(0.f, 0.f);
}

__global__ void k_to_half(__half2 *__restrict__ dst, const float2 *__restrict__ src, size_t n, float scale)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float2 v = src[i];
        dst[i] = __floats2half2_rn(v.x * scale, v.y * scale);
    }
}

__global__ void k_i32_to_u16(uint16_t *__restrict__ dst, const int32_t *__restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (uint16_t)src[i];
}

__global__ void k_fill_u32(uint32_t *__restrict__ dst, size_t n, uint32_t v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

__global__ void k_copy_u32(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill_random_u16(uint32_t *__restrict__ dst, size_t n, uint32_t seed)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 2654435761u ^ seed;
        h ^= h >> 16;
        h *= 0x7feb352du;
        h ^= h >> 15;
        h *= 0x846ca68bu;
        h ^= h >> 16;
        uint32_t a = 32768u - 2048u + (h & 4095u), b = 32768u - 2048u + ((h >> 12) & 4095u);
        dst[i] = a | (b << 16);
    }
}

// ---------------------------------------------------------------------------
// Single-pass batched C2C: TPB transforms per workgroup, thread t of
// transform tl handles points t + m*NT (consecutive t -> consecutive
// addresses, so every load and store is coalesced).
template <int LOG2N, bool INV>
__global__ __launch_bounds__(fft_threads(LOG2N)) void k_fft_c2c(
    const float2 *__restrict__ in, float2 *__restrict__ out, long long idist, long long odist, long long batch,
    float scale)
{
    constexpr int E = fft_e(LOG2N);
    using S = FftShape<LOG2N, E>;
    constexpr int TPB = fft_tpb(LOG2N);
    __shared__ float2 smem[TPB * S::NPAD + S::TWLO + S::TWHI];
    float2 *twlo = smem + TPB * S::NPAD, *twhi = twlo + S::TWLO;
    const int tl = threadIdx.x / S::NT, t = threadIdx.x % S::NT;
    if constexpr (S::NPASS > 1)
        fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const long long b = (long long)blockIdx.x * TPB + tl;
    const bool valid = b < batch;
    float2 v[E];
    const float2 *src = in + b * idist;
#pragma unroll
    for (int m = 0; m < E; m++)
        v[m] = valid ? src[t + m * S::NT] : make_float2(0.f, 0.f);
    fft_block<LOG2N, E, INV>(v, smem + tl * S::NPAD, twlo, twhi, t);
    if (valid) {
        float2 *dst = out + b * odist;
#pragma unroll
        for (int m = 0; m < E; m++)
            dst[t + m * S::NT] = cscale(v[m], scale);
    }
}

// ---------------------------------------------------------------------------
// Two-pass (four-step) C2C for N = N1 * N2 > 2^14:
//   pass A: for every column n2, DFT_N1 over x[N2*n1 + n2], times
//           W_N^(n2*k1) (FP32 sincospi and digit products), stored to
//           T[k1*N2 + n2];
//   pass B: for every row k1, DFT_N2 over T[k1*N2 + n2], stored to
//           X[k1 + N1*k2].
// Pass A runs C columns per workgroup with the column index fastest across
// lanes (coalesced strided access); pass B loads rows with the point index
// fastest and transposes through LDS so its strided stores are coalesced.
__host__ __device__ constexpr int cols_per_wg(int log2n)
{
    int nt = (1 << log2n) / fft_e(log2n);
    int npad = (1 << log2n) + ((1 << log2n) >> ilog2c(fft_e(log2n)));
    int c = 64;
    while (c > 1 && (c * nt > 1024 || c * npad * 8 > 140 * 1024))
        c >>= 1;
    return c;
}

__host__ __device__ constexpr int cols_threads(int log2n) { return ((1 << log2n) / fft_e(log2n)) * cols_per_wg(log2n); }

template <int LOG2N, bool INV>
__global__ __launch_bounds__(cols_threads(LOG2N)) void k_fft_passA(
    const float2 *__restrict__ in, float2 *__restrict__ out, long long idist, long long batch, int log2n2,
    int log2ntot)
{
    constexpr int E = fft_e(LOG2N);
    using S = FftShape<LOG2N, E>;
    constexpr int C = cols_per_wg(LOG2N);
    __shared__ float2 smem[C * S::NPAD + S::TWLO + S::TWHI];
    float2 *twlo = smem + C * S::NPAD, *twhi = twlo + S::TWLO;
    if constexpr (S::NPASS > 1)
        fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const long long N2 = 1ll << log2n2;
    const long long ncols = batch * N2;
    const long long col = (long long)blockIdx.x * C + c;
    const bool valid = col < ncols;
    const long long b = col >> log2n2, n2 = col & (N2 - 1);
    float2 v[E];
    const float2 *src = in + b * idist + n2;
#pragma unroll
    for (int m = 0; m < E; m++)
        v[m] = valid ? src[(long long)(t + m * S::NT) * N2] : make_float2(0.f, 0.f);
    fft_block<LOG2N, E, INV>(v, smem + c * S::NPAD, twlo, twhi, t);
    if (valid) {
        const long long ntot = 1ll << log2ntot;
        float2 *dst = out + (b << log2ntot) + n2;
        // W_ntot^(n2 k1) for k1 = t + m NT
        const ColumnTwiddles<E, INV> tw(S::NT > 1 ? n2 * t : 0, n2 * S::NT, ntot);
#pragma unroll
        for (int m = 0; m < E; m++) {
            const long long k1 = t + m * S::NT;
            dst[k1 * N2] = (m == 0 && !tw.has_a) ? v[m] : cmul(v[m], tw(m));
        }
    }
}

template <int LOG2N, bool INV>
__global__ __launch_bounds__(cols_threads(LOG2N)) void k_fft_passB(
    const float2 *__restrict__ in, float2 *__restrict__ out, long long odist, long long batch, int log2n1,
    float scale)
{
    constexpr int E = fft_e(LOG2N);
    using S = FftShape<LOG2N, E>;
    constexpr int C = cols_per_wg(LOG2N);
    __shared__ float2 smem[C * S::NPAD + S::TWLO + S::TWHI];
    float2 *twlo = smem + C * S::NPAD, *twhi = twlo + S::TWLO;
    if constexpr (S::NPASS > 1)
        fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const long long N1 = 1ll << log2n1;
    const long long nrows = batch * N1;
    // load phase: transform tl, point index fastest
    {
        const int tl = threadIdx.x / S::NT, t = threadIdx.x % S::NT;
        const long long row = (long long)blockIdx.x * C + tl;
        const bool valid = row < nrows;
        float2 v[E];
        const float2 *src = in + row * S::N;
#pragma unroll
        for (int m = 0; m < E; m++)
            v[m] = valid ? src[t + m * S::NT] : make_float2(0.f, 0.f);
        fft_block<LOG2N, E, INV>(v, smem + tl * S::NPAD, twlo, twhi, t);
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; m++)
            smem[tl * S::NPAD + lidx<S::LOGE>(t + m * S::NT)] = v[m];
    }
    __syncthreads();
    // store phase: column (row index) fastest across lanes
    {
        const int c = threadIdx.x % C, t = threadIdx.x / C;
        const long long row = (long long)blockIdx.x * C + c;
        if (row < nrows) {
            const long long b = row >> log2n1, k1 = row & (N1 - 1);
            float2 *dst = out + b * odist + k1;
#pragma unroll
            for (int m = 0; m < E; m++) {
                int k2 = t + m * S::NT;
                dst[(long long)k2 * N1] = cscale(smem[c * S::NPAD + lidx<S::LOGE>(k2)], scale);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused forward: window w of this call covers stream samples
// [(w-1)M, (w+1)M) relative to the call start; samples before the call come
// from the previous call's tail.  The spectrum goes to ring slot
// (g0 + w) mod rx, either FP32 or FP16 scaled by 1/N (keeps |X| in range:
// |x| <= 32768*sqrt(2) bounds |X|/N by 46341 < 65504).
template <int LOG2N, bool F16, bool MFMA_ORDER = false>
__global__ __launch_bounds__(fft_threads(LOG2N)) void k_fwd_u16(
    const uint32_t *__restrict__ cur, const uint32_t *__restrict__ tail, uint32_t *__restrict__ tail_out,
    void *__restrict__ ring, int count, long long g0, int rx)
{
    constexpr int E = fft_e(LOG2N);
    using S = FftShape<LOG2N, E>;
    constexpr int TPB = fft_tpb(LOG2N);
    constexpr long long N = S::N, M = S::N / 2;
    __shared__ float2 smem[TPB * S::NPAD + S::TWLO + S::TWHI];
    float2 *twlo = smem + TPB * S::NPAD, *twhi = twlo + S::TWLO;
    const int tl = threadIdx.x / S::NT, t = threadIdx.x % S::NT;
    if constexpr (S::NPASS > 1)
        fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const int w = blockIdx.x * TPB + tl;
    const bool valid = w < count;
    // window w covers input samples [(w - 1) M, (w + 1) M) of this call; the
    // first M of window 0 are the previous call's last M (tail)
    uint32_t u[E];
    if (valid) {
#pragma unroll
        for (int m = 0; m < E; m++) {
            const long long rel = (long long)(w - 1) * M + t + m * S::NT;
            u[m] = *(rel < 0 ? tail + (M + rel) : cur + rel);
        }
        // the last window's second half is the next call's history
        if (w == count - 1) {
#pragma unroll
            for (int m = E / 2; m < E; m++)
                tail_out[t + m * S::NT - M] = u[m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; m++)
            u[m] = 0;
    }
    // the FP16 ring holds X/N: the samples are scaled by 1/N (a power of two)
    // on unpacking, which gives the same values as scaling the transform
    constexpr float invN = 1.0f / (float)N;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++)
        v[m] = F16 ? unpack_u16pair_scaled(u[m], invN) : unpack_u16pair(u[m]);
    fft_block<LOG2N, E, false>(v, smem + tl * S::NPAD, twlo, twhi, t);
    if (valid) {
        const long long slot = (g0 + w) % rx;
        if constexpr (F16) {
            __half2 *dst = reinterpret_cast<__half2 *>(ring) + slot * N;
#pragma unroll
            for (int m = 0; m < E; m++) {
                // MFMA_ORDER: bin n at position mfft::ring_pos(n), the gather
                // order of the matrix-core inverse (2^14 points)
                const int n = t + m * S::NT;
                dst[MFMA_ORDER ? mfft::ring_pos(n) : n] = __floats2half2_rn(v[m].x, v[m].y);
            }
        } else {
            float2 *dst = reinterpret_cast<float2 *>(ring) + slot * N;
#pragma unroll
            for (int m = 0; m < E; m++)
                dst[t + m * S::NT] = v[m];
        }
    }
}

// Fused forward for calls with many 2^14-point windows (more than the GPU
// runs at once): each workgroup transforms a contiguous run of nwpw windows,
// so the twiddle tables are copied into LDS once per run, the next window's codes are
// loaded into registers while the current window is transformed, and the
// half of each window that the previous one also read is an L2 hit on the
// same XCD.  The middle pass reads its twiddles from the row table.  Same
// outputs as k_fwd_u16 (ring slot (g0 + w) mod rx, FP16 scaled by 1/N or
// FP32; the last window's second half to tail_out).  Needs count >= 2.
template <bool F16>
__global__ __launch_bounds__((1 << 14) / 32) void k_fwd_u16_nb(
    const uint32_t *__restrict__ cur, const uint32_t *__restrict__ tail, uint32_t *__restrict__ tail_out,
    void *__restrict__ ring, int count, long long g0, int rx, int nwpw)
{
    constexpr int LOG2N = 14, E = 32;
    using S = FftShape<LOG2N, E>;
    constexpr long long N = S::N, M = S::N / 2;
    constexpr int NT = S::NT;
    __shared__ float2 smem[S::NPAD + S::TWLO + S::TWHI + kTwRowElems];
    float2 *twlo = smem + S::NPAD, *twhi = twlo + S::TWLO, *twrows = twhi + S::TWHI;
    fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    fill_twrows(twrows, threadIdx.x, blockDim.x);
    const int t = threadIdx.x;
    const int w0 = blockIdx.x * nwpw;
    const int nw = min(nwpw, count - w0);
    if (nw <= 0)
        return; // uniform: the whole workgroup leaves
    uint32_t u[E];
    // window w covers input samples [(w - 1) M, (w + 1) M) of this call; the
    // first M of window 0 are the previous call's last M (tail).  Only the
    // run's first load can start at window 0.
#pragma unroll
    for (int m = 0; m < E; m++) {
        const long long rel = (long long)(w0 - 1) * M + t + m * NT;
        u[m] = *(rel < 0 ? tail + (M + rel) : cur + rel);
    }
    // Wait for the first window here: the loop's first iteration then has
    // nothing in flight, and the compiler's waits at the top of the loop
    // follow the steady state alone (vector-memory operations retire in
    // issue order; the next window's loads are issued before this window's
    // ring stores, so they are waited for without the stores).  The loop
    // body has no branches for the same reason; the last window of a run
    // loads itself again, which leaves the call's last window in u for the
    // tail.
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    for (int i = 0; i < nw; i++) {
        const int w = w0 + i;
        // FP16 ring: X/N from samples scaled by 1/N on unpacking (exact)
        constexpr float invN = 1.0f / (float)N;
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; m++)
            v[m] = F16 ? unpack_u16pair_scaled(u[m], invN) : unpack_u16pair(u[m]);
        // window min(w + 1, count - 1) starts at (min(w + 1, count - 1) - 1) M;
        // with count >= 2 that is in [0, (count - 2) M] and the window ends by
        // count M, inside this call's input
        const uint32_t *next = cur + (long long)min(w, count - 2) * M + t;
#pragma unroll
        for (int m = 0; m < E; m++)
            u[m] = next[m * NT];
        fft_block<LOG2N, E, false, true>(v, smem, twlo, twhi, t, twrows);
        const long long slot = (g0 + w) % rx;
        if constexpr (F16) {
            __half2 *dst = reinterpret_cast<__half2 *>(ring) + slot * N;
#pragma unroll
            for (int m = 0; m < E; m++)
                dst[t + m * NT] = __floats2half2_rn(v[m].x, v[m].y);
        } else {
            float2 *dst = reinterpret_cast<float2 *>(ring) + slot * N;
#pragma unroll
            for (int m = 0; m < E; m++)
                dst[t + m * NT] = v[m];
        }
    }
    // the call's last window's second half is the next call's history
    if (w0 + nw == count) {
#pragma unroll
        for (int m = E / 2; m < E; m++)
            tail_out[t + m * NT - M] = u[m];
    }
}

// Fused inverse: unit u = (dm, block b) with blocks of one DM consecutive.
// Y[n] = X_{g-d(n)}[n] * H[n] (zero before the stream start), inverse FFT,
// keep points [M, 2M), scale: FP32 ring/H -> 1/N^2 (the reference's two 1/N
// factors, gpu_kernel.cu:169,251); FP16 ring (X/N) and H (H/(2M)) carry the
// whole scale already.
// This variant serves the small sizes (several transforms per workgroup).
template <int LOG2N, bool F16>
__global__ __launch_bounds__(fft_threads(LOG2N)) void k_inv_fused(
    const void *__restrict__ ring, const void *__restrict__ H, const uint16_t *__restrict__ delay,
    float2 *__restrict__ out, int count, int ndm, long long g0, int rx)
{
    constexpr int E = fft_e(LOG2N);
    using S = FftShape<LOG2N, E>;
    constexpr int TPB = fft_tpb(LOG2N);
    constexpr long long N = S::N, M = S::N / 2;
    __shared__ float2 smem[TPB * S::NPAD + S::TWLO + S::TWHI];
    float2 *twlo = smem + TPB * S::NPAD, *twhi = twlo + S::TWLO;
    const int tl = threadIdx.x / S::NT, t = threadIdx.x % S::NT;
    if constexpr (S::NPASS > 1)
        fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const int u = blockIdx.x * TPB + tl;
    const bool valid = u < ndm * count;
    const int dm = valid ? u / count : 0;
    const int b = valid ? u - dm * count : 0;
    const long long g = g0 + b;
    const int gslot = (int)(g % rx);
    const uint16_t *drow = delay + (long long)dm * N;
    // three batches of independent loads (delays, then ring and H values) so
    // each batch is in flight at once; a ring index is clamped to slot 0 when
    // its block precedes the stream start and the value is dropped after
    using R = RingElem<F16>;
    using T = typename R::T;
    int d[E];
#pragma unroll
    for (int m = 0; m < E; m++)
        d[m] = drow[t + m * S::NT];
    T xr[E], hr[E];
    const T *ringT = reinterpret_cast<const T *>(ring);
    const T *hrow = reinterpret_cast<const T *>(H) + (long long)dm * N;
#pragma unroll
    for (int m = 0; m < E; m++) {
        int sl = gslot - d[m];
        sl += sl < 0 ? rx : 0;
        xr[m] = ringT[(long long)sl * N + t + m * S::NT];
        hr[m] = hrow[t + m * S::NT];
    }
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++)
        v[m] = (valid && g >= d[m]) ? R::mul(xr[m], hr[m]) : make_float2(0.f, 0.f);
    fft_block<LOG2N, E, true>(v, smem + tl * S::NPAD, twlo, twhi, t);
    if (valid) {
        const float scale = 1.0f / ((float)N * (float)N);
        float2 *dst = out + ((long long)dm * count + b) * M;
#pragma unroll
        for (int m = E / 2; m < E; m++)
            store_stream(&dst[t + m * S::NT - M], F16 ? v[m] : cscale(v[m], scale));
    }
}

// Fused inverse for one transform per workgroup (2^12..2^14 points): the
// workgroup runs nbpw consecutive blocks of one DM trial (H and the delays
// are re-read per block, L2 hits).  The delays
// come in a thread-major packed layout dpk[dm][j][t] = d(t + 2j NT) |
// d(t + (2j+1) NT) << 16 so each thread reads E/2 coalesced words.  Work
// items are ordered so that each XCD (workgroups are dealt to the 8 XCDs
// round robin) gets a contiguous range of DM trials and keeps their tables in
// its L2.
template <int LOG2N, int E, bool F16>
__global__ __launch_bounds__((1 << LOG2N) / E) void k_inv_fused_nb(
    const void *__restrict__ ring, const void *__restrict__ H, const uint32_t *__restrict__ dpk,
    float2 *__restrict__ out, int count, int ndm, long long g0, int rx, int nbpw)
{
    using S = FftShape<LOG2N, E>;
    static_assert(S::NT >= 256, "one transform per workgroup");
    constexpr long long N = S::N, M = S::N / 2;
    constexpr int NT = S::NT;
    using R = RingElem<F16>;
    using T = typename R::T;
    // the middle pass of 2^14 points with 32 per thread reads its twiddles
    // from a row table (one LDS read per twiddle instead of digit products)
    constexpr bool ROWS = (LOG2N == 14 && E == 32);
    __shared__ float2 smem[S::NPAD + S::TWLO + S::TWHI + (ROWS ? kTwRowElems : 0)];
    float2 *twlo = smem + S::NPAD, *twhi = twlo + S::TWLO, *twrows = twhi + S::TWHI;
    fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    if constexpr (ROWS)
        fill_twrows(twrows, threadIdx.x, blockDim.x);
    const int t = threadIdx.x;
    const int chunks = (count + nbpw - 1) / nbpw;
    const int G = gridDim.x;
    int w = blockIdx.x;
    if ((G & 7) == 0)
        w = (w & 7) * (G >> 3) + (w >> 3);
    if (w >= ndm * chunks)
        return; // the whole workgroup leaves together (w is uniform)
    const int dm = w / chunks;
    const int b0 = (w - dm * chunks) * nbpw;
    const int nb = min(nbpw, count - b0);
    const char *ringB = reinterpret_cast<const char *>(ring);
    const T *hrow = reinterpret_cast<const T *>(H) + (long long)dm * N;
    const uint32_t *drow = dpk + (long long)dm * N / 2;
    int slot = (int)((g0 + b0) % rx);
    // A block before the stream start maps to slot rx + g - d, the slot of a
    // block (> g0 + count - 1) that no call has written yet: the ring is
    // zeroed at creation and reset, so those gathers read zero without a test.
    T raw[E];
    // byte offsets fit 32 bits (the ring is < 4 GiB): loads use a scalar base
    // plus one 32-bit VGPR offset
    auto gather = [&]() {
#pragma unroll
        for (int j = 0; j < E / 2; j++) {
            const uint32_t pair = drow[j * NT + t];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int m = 2 * j + q;
                const int d = q ? (int)(pair >> 16) : (int)(pair & 0xFFFFu);
                int sl = slot - d;
                sl += sl < 0 ? rx : 0;
                if constexpr (F16) {
                    const uint32_t off =
                        ((uint32_t)sl * (uint32_t)N + (uint32_t)(t + m * NT)) * (uint32_t)sizeof(T);
                    raw[m] = *reinterpret_cast<const T *>(ringB + off);
                } else {
                    // the point's column offset in the scalar base (one
                    // instruction per address; measured faster for FP32 rings
                    // only)
                    const uint32_t off = ((uint32_t)sl * (uint32_t)N + (uint32_t)t) * (uint32_t)sizeof(T);
                    raw[m] = *reinterpret_cast<const T *>(ringB + (long long)m * NT * (long long)sizeof(T) + off);
                }
            }
        }
    };
    const float scale = 1.0f / ((float)N * (float)N);
    for (int i = 0; i < nb; i++) {
        gather();
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; m++)
            v[m] = R::mul(raw[m], hrow[t + m * NT]);
        slot = slot + 1 == rx ? 0 : slot + 1;
        fft_block<LOG2N, E, true, ROWS>(v, smem, twlo, twhi, t, twrows);
        float2 *dst = out + ((long long)dm * count + b0 + i) * M;
#pragma unroll
        for (int m = E / 2; m < E; m++)
            store_stream(&dst[t + m * NT - M], F16 ? v[m] : cscale(v[m], scale));
    }
}

// ---------------------------------------------------------------------------
// Launchers

#define MSOSM_LAUNCH_CHECK()                                                                                     \
    do {                                                                                                         \
        hipError_t e__ = hipGetLastError();                                                                      \
        if (e__ != hipSuccess)                                                                                   \
            return e__;                                                                                          \
    } while (0)

template <int LOG2N>
static hipError_t fft_single(const float2 *in, float2 *out, long long batch, long long idist, long long odist,
                             bool inv, float scale, hipStream_t s)
{
    constexpr int E = fft_e(LOG2N);
    constexpr int threads = FftShape<LOG2N, E>::NT * fft_tpb(LOG2N);
    const long long grid = (batch + fft_tpb(LOG2N) - 1) / fft_tpb(LOG2N);
    if (inv)
        hipLaunchKernelGGL((k_fft_c2c<LOG2N, true>), dim3((unsigned)grid), dim3(threads), 0, s, in, out, idist,
                           odist, batch, scale);
    else
        hipLaunchKernelGGL((k_fft_c2c<LOG2N, false>), dim3((unsigned)grid), dim3(threads), 0, s, in, out, idist,
                           odist, batch, scale);
    MSOSM_LAUNCH_CHECK();
    return hipSuccess;
}

template <int LOG2N>
static hipError_t fft_passA(const float2 *in, float2 *out, long long idist, long long batch, int log2n2,
                            int log2ntot, bool inv, hipStream_t s)
{
    constexpr int E = fft_e(LOG2N);
    constexpr int C = cols_per_wg(LOG2N);
    constexpr int threads = FftShape<LOG2N, E>::NT * C;
    const long long ncols = batch << log2n2;
    const long long grid = (ncols + C - 1) / C;
    if (inv)
        hipLaunchKernelGGL((k_fft_passA<LOG2N, true>), dim3((unsigned)grid), dim3(threads), 0, s, in, out, idist,
                           batch, log2n2, log2ntot);
    else
        hipLaunchKernelGGL((k_fft_passA<LOG2N, false>), dim3((unsigned)grid), dim3(threads), 0, s, in, out, idist,
                           batch, log2n2, log2ntot);
    MSOSM_LAUNCH_CHECK();
    return hipSuccess;
}

template <int LOG2N>
static hipError_t fft_passB(const float2 *in, float2 *out, long long odist, long long batch, int log2n1, bool inv,
                            float scale, hipStream_t s)
{
    constexpr int E = fft_e(LOG2N);
    constexpr int C = cols_per_wg(LOG2N);
    constexpr int threads = FftShape<LOG2N, E>::NT * C;
    const long long nrows = batch << log2n1;
    const long long grid = (nrows + C - 1) / C;
    if (inv)
        hipLaunchKernelGGL((k_fft_passB<LOG2N, true>), dim3((unsigned)grid), dim3(threads), 0, s, in, out, odist,
                           batch, log2n1, scale);
    else
        hipLaunchKernelGGL((k_fft_passB<LOG2N, false>), dim3((unsigned)grid), dim3(threads), 0, s, in, out, odist,
                           batch, log2n1, scale);
    MSOSM_LAUNCH_CHECK();
    return hipSuccess;
}

#define MSOSM_DISPATCH_1_14(L, CALL)                                                                             \
    switch (L) {                                                                                                 \
    case 1: return CALL(1);                                                                                      \
    case 2: return CALL(2);                                                                                      \
    case 3: return CALL(3);                                                                                      \
    case 4: return CALL(4);                                                                                      \
    case 5: return CALL(5);                                                                                      \
    case 6: return CALL(6);                                                                                      \
    case 7: return CALL(7);                                                                                      \
    case 8: return CALL(8);                                                                                      \
    case 9: return CALL(9);                                                                                      \
    case 10: return CALL(10);                                                                                    \
    case 11: return CALL(11);                                                                                    \
    case 12: return CALL(12);                                                                                    \
    case 13: return CALL(13);                                                                                    \
    case 14: return CALL(14);                                                                                    \
    default: return hipErrorInvalidValue;                                                                        \
    }

static void split_two_pass(int log2n, int &a, int &b)
{
    a = log2n - kMaxLog2SinglePass;
    if (a < log2n / 2)
        a = log2n / 2;
    b = log2n - a;
}

size_t fft_c2c_scratch_elems(int log2n, long long batch)
{
    if (log2n <= kMaxLog2SinglePass)
        return 0;
    return (size_t)batch << log2n;
}

hipError_t launch_fft_c2c(const float2 *in, float2 *out, int log2n, long long batch, long long idist,
                          long long odist, bool inverse, float scale, float2 *scratch, hipStream_t s)
{
    if (log2n < 1 || log2n > 28 || batch < 1)
        return hipErrorInvalidValue;
    if (log2n <= kMaxLog2SinglePass) {
#define CALL_SINGLE(L) fft_single<L>(in, out, batch, idist, odist, inverse, scale, s)
        MSOSM_DISPATCH_1_14(log2n, CALL_SINGLE)
#undef CALL_SINGLE
    }
    int a, b;
    split_two_pass(log2n, a, b);
    bool own = false;
    if (!scratch) {
        hipError_t e = hipMallocAsync((void **)&scratch, fft_c2c_scratch_elems(log2n, batch) * sizeof(float2), s);
        if (e != hipSuccess)
            return e;
        own = true;
    }
    hipError_t e;
    {
#define CALL_A(L) fft_passA<L>(in, scratch, idist, batch, b, log2n, inverse, s)
        e = [&]() -> hipError_t { MSOSM_DISPATCH_1_14(a, CALL_A) }();
#undef CALL_A
    }
    if (e == hipSuccess) {
#define CALL_B(L) fft_passB<L>(scratch, out, odist, batch, a, inverse, scale, s)
        e = [&]() -> hipError_t { MSOSM_DISPATCH_1_14(b, CALL_B) }();
#undef CALL_B
    }
    if (own) {
        hipError_t e2 = hipFreeAsync(scratch, s);
        if (e == hipSuccess)
            e = e2;
    }
    return e;
}

bool fused_supported(int log2n) { return log2n >= 2 && log2n <= kMaxLog2SinglePass; }

int device_cu_count()
{
    constexpr int kMaxDevices = 64;
    static std::atomic<int> cache[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0)
        return 256;
    int n = dev < kMaxDevices ? cache[dev].load(std::memory_order_relaxed) : 0;
    if (n > 0)
        return n;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    if (dev < kMaxDevices)
        cache[dev].store(n, std::memory_order_relaxed);
    return n;
}

int runs_per_wg(long long items)
{
    const long long cus = device_cu_count();
    return items <= cus ? 1 : (int)((items + cus - 1) / cus);
}

template <int LOG2N>
static hipError_t fwd_u16(const uint32_t *cur, const uint32_t *tail, uint32_t *tail_out, void *ring, bool f16,
                          int count, long long g0, int rx, bool mfma_order, hipStream_t s)
{
    constexpr int E = fft_e(LOG2N);
    constexpr int threads = FftShape<LOG2N, E>::NT * fft_tpb(LOG2N);
    const int grid = (count + fft_tpb(LOG2N) - 1) / fft_tpb(LOG2N);
    if constexpr (LOG2N == mfft::kLog2N) {
        if (mfma_order) {
            if (!f16)
                return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_fwd_u16<LOG2N, true, true>), dim3(grid), dim3(threads), 0, s, cur, tail, tail_out,
                               ring, count, g0, rx);
            MSOSM_LAUNCH_CHECK();
            return hipSuccess;
        }
    } else if (mfma_order) {
        return hipErrorInvalidValue;
    }
    if constexpr (LOG2N == 14) {
        // more windows than run at once: runs of windows per workgroup (the
        // kernel reloads the call's last window, which needs count >= 2)
        const int nwpw = runs_per_wg(count);
        if (nwpw > 1 && count >= 2) {
            const int g = (count + nwpw - 1) / nwpw;
            if (f16)
                hipLaunchKernelGGL((k_fwd_u16_nb<true>), dim3(g), dim3(threads), 0, s, cur, tail, tail_out, ring,
                                   count, g0, rx, nwpw);
            else
                hipLaunchKernelGGL((k_fwd_u16_nb<false>), dim3(g), dim3(threads), 0, s, cur, tail, tail_out, ring,
                                   count, g0, rx, nwpw);
            MSOSM_LAUNCH_CHECK();
            return hipSuccess;
        }
    }
    if (f16)
        hipLaunchKernelGGL((k_fwd_u16<LOG2N, true>), dim3(grid), dim3(threads), 0, s, cur, tail, tail_out, ring, count,
                           g0, rx);
    else
        hipLaunchKernelGGL((k_fwd_u16<LOG2N, false>), dim3(grid), dim3(threads), 0, s, cur, tail, tail_out, ring,
                           count, g0, rx);
    MSOSM_LAUNCH_CHECK();
    return hipSuccess;
}

// Blocks per workgroup of k_inv_fused_nb: the largest of 32, 16, ..., 1 that
// still gives at least two workgroups per CU (kInvMinGrid), else 1.  A
// trial's chunks share its filter and delay rows in L2, neighbouring trials
// their ring entries, and fewer, longer workgroups turn over less often; few
// trials need short chunks to fill the GPU at all.  For 256 trials x 64
// blocks of 2^14 points with streaming output stores: 0.463 ms per launch at
// 32 (FP32 ring 0.578), 0.474 (0.591) at 16, 0.463 at 64.
constexpr long long kInvMinGrid = 512;
int blocks_per_wg(int count, long long rows, int max_nbpw)
{
    for (int nbpw = max_nbpw; nbpw > 1; nbpw /= 2)
        if ((long long)((count + nbpw - 1) / nbpw) * rows >= kInvMinGrid)
            return nbpw;
    return 1;
}
static int inv_blocks_per_wg(int count, int ndm) { return blocks_per_wg(count, ndm, 32); }

template <int LOG2N, int E, bool F16>
static void launch_nb(const void *ring, const void *H, const uint32_t *dpk, float2 *out, int count, int ndm,
                      long long g0, int rx, int nbpw_req, hipStream_t s)
{
    constexpr int threads = (1 << LOG2N) / E;
    const int nbpw = nbpw_req > 0 ? std::min(nbpw_req, count) : inv_blocks_per_wg(count, ndm);
    const long long chunks = (count + nbpw - 1) / nbpw;
    const long long grid = chunks * ndm;
    hipLaunchKernelGGL((k_inv_fused_nb<LOG2N, E, F16>), dim3((unsigned)grid), dim3(threads), 0, s, ring, H, dpk, out,
                       count, ndm, g0, rx, nbpw);
}

template <int LOG2N>
static hipError_t inv_fused(const void *ring, const void *H, const uint16_t *delay, const uint32_t *dpk, float2 *out,
                            bool f16, int count, int ndm, long long g0, int rx, int nbpw, hipStream_t s)
{
    constexpr int E = fft_e(LOG2N);
    constexpr int threads = FftShape<LOG2N, E>::NT * fft_tpb(LOG2N);
    if constexpr (fft_tpb(LOG2N) == 1) {
        if (f16)
            launch_nb<LOG2N, E, true>(ring, H, dpk, out, count, ndm, g0, rx, nbpw, s);
        else
            launch_nb<LOG2N, E, false>(ring, H, dpk, out, count, ndm, g0, rx, nbpw, s);
    } else {
        const long long units = (long long)count * ndm;
        const long long grid = (units + fft_tpb(LOG2N) - 1) / fft_tpb(LOG2N);
        if (f16)
            hipLaunchKernelGGL((k_inv_fused<LOG2N, true>), dim3((unsigned)grid), dim3(threads), 0, s, ring, H, delay,
                               out, count, ndm, g0, rx);
        else
            hipLaunchKernelGGL((k_inv_fused<LOG2N, false>), dim3((unsigned)grid), dim3(threads), 0, s, ring, H,
                               delay, out, count, ndm, g0, rx);
    }
    MSOSM_LAUNCH_CHECK();
    return hipSuccess;
}

#define MSOSM_DISPATCH_2_14(L, CALL)                                                                             \
    switch (L) {                                                                                                 \
    case 2: return CALL(2);                                                                                      \
    case 3: return CALL(3);                                                                                      \
    case 4: return CALL(4);                                                                                      \
    case 5: return CALL(5);                                                                                      \
    case 6: return CALL(6);                                                                                      \
    case 7: return CALL(7);                                                                                      \
    case 8: return CALL(8);                                                                                      \
    case 9: return CALL(9);                                                                                      \
    case 10: return CALL(10);                                                                                    \
    case 11: return CALL(11);                                                                                    \
    case 12: return CALL(12);                                                                                    \
    case 13: return CALL(13);                                                                                    \
    case 14: return CALL(14);                                                                                    \
    default: return hipErrorInvalidValue;                                                                        \
    }

hipError_t launch_fwd_u16(const uint32_t *cur, const uint32_t *tail, uint32_t *tail_out, void *ring, bool f16,
                          int log2n, int count, long long g0, int rx, bool mfma_order, hipStream_t s)
{
#define CALL_F(L) fwd_u16<L>(cur, tail, tail_out, ring, f16, count, g0, rx, mfma_order, s)
    MSOSM_DISPATCH_2_14(log2n, CALL_F)
#undef CALL_F
}

int fused_delay_threads(int log2n) { return (1 << log2n) / fft_e(log2n); }
bool fused_uses_packed_delay(int log2n) { return log2n >= 2 && log2n <= kMaxLog2SinglePass && fft_tpb(log2n) == 1; }

hipError_t launch_inv_fused(const void *ring, const void *H, const uint16_t *delay, const uint32_t *dpk, float2 *out,
                            bool f16, int log2n, int count, int ndm, long long g0, int rx, int nbpw, hipStream_t s)
{
#define CALL_I(L) inv_fused<L>(ring, H, delay, dpk, out, f16, count, ndm, g0, rx, nbpw, s)
    MSOSM_DISPATCH_2_14(log2n, CALL_I)
#undef CALL_I
}

hipError_t launch_unpack(float2 *dst, const uint32_t *src, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_pack(uint32_t *dst, const float2 *src, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_intensity(const float2 *a, const float2 *b, float *out, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_intensity, dim3(grid_for(n)), dim3(kBlock), 0, s, a, b, out, n);
    return hipGetLastError();
}

hipError_t launch_multiply(const float2 *a, const float2 *b, int N, float2 *out, int count, int ndm, hipStream_t s)
{
    size_t total = (size_t)N * count * ndm;
    hipLaunchKernelGGL(k_multiply, dim3(grid_for(total)), dim3(kBlock), 0, s, a, b, out, N, count, ndm);
    return hipGetLastError();
}

hipError_t launch_multiply_delay(const float2 *a, const float2 *b, int N, float2 *ring, const int32_t *delay,
                                 int dbs, int index, int count, int ndm, hipStream_t s)
{
    int gy = (N + kBlock - 1) / kBlock;
    if (gy > 64)
        gy = 64;
    hipLaunchKernelGGL(k_multiply_delay, dim3((unsigned)((long long)count * ndm), gy), dim3(kBlock), 0, s, a, b,
                       ring, delay, dbs, index, N, count, ndm);
    return hipGetLastError();
}

hipError_t launch_discard(const float2 *a, float2 *b, int M, int count, int ndm, hipStream_t s)
{
    size_t groups = (size_t)count * ndm;
    hipLaunchKernelGGL(k_discard, dim3(grid_for(groups * M)), dim3(kBlock), 0, s, a, b, M, groups);
    return hipGetLastError();
}

hipError_t launch_chirp(float2 *c, unsigned long M, int mode, const ChirpParams *params_dev, int ndm,
                        hipStream_t s)
{
    int gx = grid_for(M);
    if (gx > 4096)
        gx = 4096;
    hipLaunchKernelGGL(k_chirp, dim3(gx, ndm), dim3(kBlock), 0, s, c, M, mode, params_dev);
    return hipGetLastError();
}

hipError_t launch_zero_pad(float2 *dst, const float2 *src, unsigned long M, int ndm, hipStream_t s)
{
    int gx = grid_for(2 * M);
    if (gx > 4096)
        gx = 4096;
    hipLaunchKernelGGL(k_zero_pad, dim3(gx, ndm), dim3(kBlock), 0, s, dst, src, M);
    return hipGetLastError();
}

hipError_t launch_to_half(__half2 *dst, const float2 *src, size_t n, float scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_half, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n, scale);
    return hipGetLastError();
}

hipError_t launch_fill_u32(void *dst, size_t words, uint32_t value, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(words, 4)), dim3(kBlock), 0, s, (uint32_t *)dst, words, value);
    return hipGetLastError();
}

hipError_t launch_copy_u32(void *dst, const void *src, size_t words, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(words)), dim3(kBlock), 0, s, (uint32_t *)dst,
                       (const uint32_t *)src, words);
    return hipGetLastError();
}

hipError_t launch_fill_random_u16(uint32_t *dst, size_t n, uint32_t seed, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_random_u16, dim3(grid_for(n, 4)), dim3(kBlock), 0, s, dst, n, seed);
    return hipGetLastError();
}

hipError_t launch_i32_to_u16(uint16_t *dst, const int32_t *src, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_i32_to_u16, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
    return hipGetLastError();
}

} // namespace msosm
