// bigfft.hip -- the shared-ring forms (FUSED_F32, FUSED_F16) for FFT sizes
// 2^15 .. 2^21, as four-step transforms N = N1 * N2 with rows of N2 = 2^14
// points (one workgroup, the 2^14-point LDS transform) and columns of N1 =
// 2 .. 128 points (many columns per workgroup, coalesced across columns).
//
// Index maps (n = input sample of a window, k = frequency bin, m = output):
//   forward  X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) [W_N^(n2 k1) sum_n1 W_N1^(n1 k1) x[n2 + N2 n1]]
//   inverse  y[m2 + N2 m1] = sum_k1 W_N1^(-k1 m1) [W_N^(-k1 m2) sum_k2 W_N2^(-k2 m2) Z[k1 + N1 k2]]
// The spectra ring stores bin k1 + N1 k2 at position p = k1 N2 + k2 (row k1
// of the forward's second pass, written as it comes out), and the filter
// table H and the delay table are kept in the same position order, so the
// inverse's first pass reads whole rows: Z[p] = X_{g - d(p)}[p] H[p] is a
// row-wise gather.  Only y[M .. 2M) is kept: m1 >= N1/2 in the second pass.
//
// Kernels per call (count windows, u = (trial, block) units):
//   fwd_cols  (window, 256-thread tile of columns n2): unpack the overlapped
//             uint16 window (previous call's tail for negative offsets; the
//             last window also writes this call's tail), DFT_N1, twiddle
//             W_N^(n2 k1) (ColumnTwiddles), store T[w][k1][n2] (FP32 scratch; in the
//             FP16 form FP16 scaled by 1/N1, which keeps |T| <= 46341)
//   fwd_rows  (window, k1): DFT_N2 of row k1, store to the ring slot of the
//             window (FP16 scaled by 1/N, or FP32)
//   inv_rows  (unit, k1), over a chunk of units: gather * H (FP32 products),
//             inverse DFT_N2, twiddle W_N^(-k1 m2), store U[unit][k1][m2]
//             (FP32; in the FP16 form FP16 scaled by 2^-7, see below)
//   inv_cols  (unit, tile of columns m2): inverse DFT_N1 of column m2, keep
//             m1 >= N1/2, scale, streaming store to out[trial][block][.]
// The inverse runs chunk by chunk so its scratch (about 128 MiB in FP32) is
// re-used while it is still in the last-level cache.
//
// FP16 row scratch: a row of U is an inverse DFT of N2 products Z = (X/N)
// (H/(2M)).  By Parseval sum |X/N|^2 <= max|x|^2 <= 46341^2 over all N bins,
// so over the N2 bins of one row sum |X/N| <= sqrt(N2) * 46341 (Cauchy-Schwarz)
// and |U| <= 128 * 46341 * max|H/(2M)|.  H[2k] = M C[k] with |C| = 1, and an
// odd bin is the zero-padded interpolation sum_k C[k] 2 / (1 - W_2M^(j-2k)),
// so |H/(2M)| <= ln(M)/pi + 0.4 (4.8 at M = 2^20; even bins 1/2, the plans
// measured 0.7-1.5 from M = 2^9 to 2^20).  Stored times 2^-9 (exact), |U|
// stays below 65504 for any uint16 input and any plan here; the column pass
// multiplies 2^9 back into its output scale.  Typical values (noise inputs)
// sit near 0.4 after the scale, far above FP16's normal minimum.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>

#include "elems.h"
#include "fft_device.h"
#include "kernels.h"

namespace msosm {
namespace {

constexpr int kRowLog2 = 14;
constexpr int kRowE = 32;
using RowS = FftShape<kRowLog2, kRowE>; // 512 threads, radix 32-32-16
constexpr long long kN2 = 1ll << kRowLog2;
constexpr int kColThreads = 256;
// windows (units) per forward (inverse) launch pair: about this many
// scratch bytes by default
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr float kRowScratchScale = 1.0f / 512.0f; // FP16 row scratch of the inverse (below)

template <int L1>
struct ColShape {
    static constexpr int E = fft_e(L1);
    using S = FftShape<L1, E>;
    static constexpr int NT = S::NT;                  // threads per column
    static constexpr int C = kColThreads / NT;        // columns per workgroup
    static constexpr int LDS = S::NPASS > 1 ? C * S::NPAD + S::TWLO + S::TWHI : 1;
    static_assert(C >= 32 && kN2 % C == 0, "column tiles must be at least 32 wide");
};

// W_N^(-e) (inverse sign) in FP64, rounded to FP32
__device__ __forceinline__ float2 twiddle_f64(long long e, long long N, bool inv)
{
    double s, c;
    sincospi(2.0 * (double)e / (double)N, &s, &c);
    return make_float2((float)c, inv ? (float)s : (float)(-s));
}

template <int L1, bool F16>
__global__ __launch_bounds__(kColThreads) void k_big_fwd_cols(const uint32_t *__restrict__ cur,
                                                              const uint32_t *__restrict__ tail,
                                                              uint32_t *__restrict__ tail_out,
                                                              void *__restrict__ T, int w0, int count)
{
    using CS = ColShape<L1>;
    using S = typename CS::S;
    constexpr int E = CS::E, NT = CS::NT, C = CS::C;
    constexpr long long N1 = 1ll << L1, N = N1 * kN2, M = N / 2;
    __shared__ float2 smem[CS::LDS];
    float2 *twlo = smem + (S::NPASS > 1 ? C * S::NPAD : 0), *twhi = twlo + S::TWLO;
    if constexpr (S::NPASS > 1)
        fill_twiddles<L1, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    constexpr int tiles = (int)(kN2 / C);
    const int wl = blockIdx.x / tiles, w = w0 + wl; // window within the chunk / the call
    const long long n2 = (long long)(blockIdx.x % tiles) * C + c;
    // window w covers samples [(w - 1) M, (w + 1) M) of this call; negative
    // offsets are the previous call's last M samples (tail), which only the
    // call's first window reads (w is uniform over the workgroup)
    uint32_t u[E];
    if (w > 0) {
        const uint32_t *src = cur + ((long long)(w - 1) * M + n2);
#pragma unroll
        for (int m = 0; m < E; m++)
            u[m] = src[kN2 * (t + m * NT)];
    } else {
#pragma unroll
        for (int m = 0; m < E; m++) {
            const long long rel = kN2 * (t + m * NT) + n2 - M;
            u[m] = *(rel < 0 ? tail + (M + rel) : cur + rel);
        }
    }
    if (w == count - 1) {
#pragma unroll
        for (int m = 0; m < E; m++) {
            const long long j = kN2 * (t + m * NT) + n2;
            if (j >= M)
                tail_out[j - M] = u[m]; // the next call's history
        }
    }
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++)
        v[m] = unpack_u16pair(u[m]);
    fft_block<L1, E, false>(v, smem + c * S::NPAD, twlo, twhi, t);
    // W_N^(n2 k1) for this thread's k1 = t + m NT
    const ColumnTwiddles<E, false> tw(NT > 1 ? n2 * t : 0, n2 * NT, N);
    const long long off = (long long)wl * N + n2;
#pragma unroll
    for (int m = 0; m < E; m++) {
        const long long k1 = t + m * NT;
        const float2 y = (m == 0 && !tw.has_a) ? v[m] : cmul(v[m], tw(m));
        if constexpr (F16) {
            constexpr float s = 1.0f / (float)N1;
            reinterpret_cast<__half2 *>(T)[off + k1 * kN2] = __floats2half2_rn(y.x * s, y.y * s);
        } else {
            reinterpret_cast<float2 *>(T)[off + k1 * kN2] = y;
        }
    }
}

template <bool F16>
__global__ __launch_bounds__(RowS::NT) void k_big_fwd_rows(const void *__restrict__ T, void *__restrict__ ring,
                                                           int w0, int l1, long long g0, int rx)
{
    constexpr int E = kRowE, NT = RowS::NT;
    __shared__ float2 smem[RowS::NPAD + RowS::TWLO + RowS::TWHI + kTwRowElems];
    float2 *twlo = smem + RowS::NPAD, *twhi = twlo + RowS::TWLO, *twrows = twhi + RowS::TWHI;
    fill_twiddles<kRowLog2, E>(twlo, twhi, threadIdx.x, blockDim.x);
    fill_twrows(twrows, threadIdx.x, blockDim.x);
    const int t = threadIdx.x;
    const long long N1 = 1ll << l1, N = N1 * kN2;
    const int wl = blockIdx.x >> l1, k1 = blockIdx.x & (int)(N1 - 1);
    const int w = w0 + wl;
    const long long roff = (long long)wl * N + k1 * kN2;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++)
        v[m] = F16 ? __half22float2(reinterpret_cast<const __half2 *>(T)[roff + t + m * NT])
                   : reinterpret_cast<const float2 *>(T)[roff + t + m * NT];
    fft_block<kRowLog2, E, false, true>(v, smem, twlo, twhi, t, twrows);
    const long long slot = (g0 + w) % rx;
    if constexpr (F16) {
        // the scratch carried 1/N1: the ring's X/N is this row's output / N2
        __half2 *dst = reinterpret_cast<__half2 *>(ring) + slot * N + k1 * kN2;
        const float invN = 1.0f / (float)kN2;
#pragma unroll
        for (int m = 0; m < E; m++)
            dst[t + m * NT] = __floats2half2_rn(v[m].x * invN, v[m].y * invN);
    } else {
        float2 *dst = reinterpret_cast<float2 *>(ring) + slot * N + k1 * kN2;
#pragma unroll
        for (int m = 0; m < E; m++)
            dst[t + m * NT] = v[m];
    }
}

// k_big_fwd_rows for calls with more rows than CUs: each workgroup transforms
// a run of nrpw consecutive rows r = wl N1 + k1 of the column scratch (row r
// starts at r N2), builds its twiddle tables once, and loads the next row
// while the current one is transformed (issued before the current row's ring
// stores; the loop body has no branches, so the compiler's waits at its top
// wait for those loads only).  The run's last row loads itself again.
template <bool F16>
__global__ __launch_bounds__(RowS::NT) void k_big_fwd_rows_nb(const void *__restrict__ T, void *__restrict__ ring,
                                                              int w0, int l1, long long g0, int rx, int nrows,
                                                              int nrpw)
{
    using TT = typename RingElem<F16>::T;
    constexpr int E = kRowE, NT = RowS::NT;
    __shared__ float2 smem[RowS::NPAD + RowS::TWLO + RowS::TWHI + kTwRowElems];
    float2 *twlo = smem + RowS::NPAD, *twhi = twlo + RowS::TWLO, *twrows = twhi + RowS::TWHI;
    fill_twiddles<kRowLog2, E>(twlo, twhi, threadIdx.x, blockDim.x);
    fill_twrows(twrows, threadIdx.x, blockDim.x);
    const int t = threadIdx.x;
    const long long N1 = 1ll << l1, N = N1 * kN2;
    const int r0 = blockIdx.x * nrpw;
    const int nr = min(nrpw, nrows - r0);
    if (nr <= 0)
        return; // uniform: the whole workgroup leaves
    const TT *src = reinterpret_cast<const TT *>(T);
    TT raw[E];
#pragma unroll
    for (int m = 0; m < E; m++)
        raw[m] = src[(long long)r0 * kN2 + t + m * NT];
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0); // the loop's first iteration has nothing in flight
    for (int i = 0; i < nr; i++) {
        const int r = r0 + i;
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; m++)
            v[m] = RingElem<F16>::f2(raw[m]);
        const TT *next = src + (long long)min(r + 1, nrows - 1) * kN2 + t;
#pragma unroll
        for (int m = 0; m < E; m++)
            raw[m] = next[m * NT];
        fft_block<kRowLog2, E, false, true>(v, smem, twlo, twhi, t, twrows);
        const int w = w0 + (r >> l1), k1 = r & (int)(N1 - 1);
        const long long slot = (g0 + w) % rx;
        if constexpr (F16) {
            // the scratch carried 1/N1: the ring's X/N is this row's output / N2
            __half2 *dst = reinterpret_cast<__half2 *>(ring) + slot * N + (long long)k1 * kN2;
            const float invN = 1.0f / (float)kN2;
#pragma unroll
            for (int m = 0; m < E; m++)
                dst[t + m * NT] = __floats2half2_rn(v[m].x * invN, v[m].y * invN);
        } else {
            float2 *dst = reinterpret_cast<float2 *>(ring) + slot * N + (long long)k1 * kN2;
#pragma unroll
            for (int m = 0; m < E; m++)
                dst[t + m * NT] = v[m];
        }
    }
}

// A block before the stream start gathers from slot rx + g - d, a slot no
// call has written yet (blocks newer than the call's last): the ring is
// zeroed at creation and reset, so those reads are zero.
template <bool F16>
__global__ __launch_bounds__(RowS::NT) void k_big_inv_rows(const void *__restrict__ ring, const void *__restrict__ Hp,
                                                           const uint16_t *__restrict__ dp, float2 *__restrict__ U,
                                                           int u0, int count, int l1, long long g0, int rx)
{
    using R = RingElem<F16>;
    using TT = typename R::T;
    constexpr int E = kRowE, NT = RowS::NT;
    __shared__ float2 smem[RowS::NPAD + RowS::TWLO + RowS::TWHI + kTwRowElems + E];
    float2 *twlo = smem + RowS::NPAD, *twhi = twlo + RowS::TWLO, *twrows = twhi + RowS::TWHI;
    float2 *step = twrows + kTwRowElems;
    fill_twiddles<kRowLog2, E>(twlo, twhi, threadIdx.x, blockDim.x);
    fill_twrows(twrows, threadIdx.x, blockDim.x);
    const int t = threadIdx.x;
    const long long N1 = 1ll << l1, N = N1 * kN2;
    const int ul = blockIdx.x >> l1, k1 = blockIdx.x & (int)(N1 - 1);
    const int u = u0 + ul;
    const int dm = u / count, b = u - dm * count;
    // twiddle W_N^(-k1 m2) for m2 = t + m NT: (W_N^(-k1 t)) (W_N^(-k1 NT m))
    if (t < E)
        step[t] = twiddle_f64(((long long)k1 * NT * t) & (N - 1), N, true);
    const float2 base = twiddle_f64(((long long)k1 * t) & (N - 1), N, true);
    const long long p0 = (long long)k1 * kN2;
    // the ring's row k1 of slot sl starts at byte (sl N + p0) sizeof(TT): the
    // engine keeps fused rings below 4 GiB, so 32-bit offsets from a scalar
    // base, with N = 2^(l1 + 14) a shift
    const char *rgB = reinterpret_cast<const char *>(ring) + p0 * (long long)sizeof(TT);
    const int lshift = l1 + kRowLog2;
    const TT *hrow = reinterpret_cast<const TT *>(Hp) + (long long)dm * N + p0;
    const uint16_t *drow = dp + (long long)dm * N + p0;
    const int s0 = (int)((g0 + b) % rx);
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int q = t + m * NT;
        int sl = s0 - (int)drow[q];
        sl += sl < 0 ? rx : 0;
        const uint32_t off = (((uint32_t)sl << lshift) + (uint32_t)q) * (uint32_t)sizeof(TT);
        v[m] = R::mul(*reinterpret_cast<const TT *>(rgB + off), hrow[q]);
    }
    fft_block<kRowLog2, E, true, true>(v, smem, twlo, twhi, t, twrows); // barriers: step[] is visible after it
    const long long roff = ((long long)ul * N1 + k1) * kN2;
    if constexpr (F16) {
        __half2 *dst = reinterpret_cast<__half2 *>(U) + roff;
#pragma unroll
        for (int m = 0; m < E; m++) {
            const float2 r = cmul(v[m], cmul(base, step[m]));
            dst[t + m * NT] = __floats2half2_rn(r.x * kRowScratchScale, r.y * kRowScratchScale);
        }
    } else {
        float2 *dst = U + roff;
#pragma unroll
        for (int m = 0; m < E; m++)
            dst[t + m * NT] = cmul(v[m], cmul(base, step[m]));
    }
}

template <int L1, bool F16>
__global__ __launch_bounds__(kColThreads) void k_big_inv_cols(const float2 *__restrict__ U, float2 *__restrict__ out,
                                                              int u0, float scale)
{
    using CS = ColShape<L1>;
    using S = typename CS::S;
    constexpr int E = CS::E, NT = CS::NT, C = CS::C;
    constexpr long long N1 = 1ll << L1, N = N1 * kN2, M = N / 2;
    __shared__ float2 smem[CS::LDS];
    float2 *twlo = smem + (S::NPASS > 1 ? C * S::NPAD : 0), *twhi = twlo + S::TWLO;
    if constexpr (S::NPASS > 1)
        fill_twiddles<L1, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    constexpr int tiles = (int)(kN2 / C);
    const int ul = blockIdx.x / tiles;
    const long long m2 = (long long)(blockIdx.x % tiles) * C + c;
    float2 v[E];
    if constexpr (F16) {
        const __half2 *src = reinterpret_cast<const __half2 *>(U) + (long long)ul * N + m2;
#pragma unroll
        for (int m = 0; m < E; m++)
            v[m] = __half22float2(src[(long long)(t + m * NT) * kN2]);
    } else {
        const float2 *src = U + (long long)ul * N + m2;
#pragma unroll
        for (int m = 0; m < E; m++)
            v[m] = src[(long long)(t + m * NT) * kN2];
    }
    fft_block<L1, E, true>(v, smem + c * S::NPAD, twlo, twhi, t);
    // unit u = trial * count + block: its output row starts at u * M
    float2 *dst = out + (long long)(u0 + ul) * M + m2;
#pragma unroll
    for (int m = 0; m < E; m++) {
        const long long m1 = t + m * NT;
        if (m1 >= N1 / 2)
            store_stream(&dst[kN2 * (m1 - N1 / 2)], cscale(v[m], scale));
    }
}

__global__ void k_big_permute(const float2 *__restrict__ src, void *__restrict__ dst, bool f16, float scale, int l1,
                              long long rows)
{
    const long long N1 = 1ll << l1, N = N1 * kN2;
    const long long total = rows * N;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long r = i / N, p = i - r * N;
        const long long n = (p >> kRowLog2) + N1 * (p & (kN2 - 1)); // bin at position p
        const float2 v = src[r * N + n];
        if (f16)
            reinterpret_cast<__half2 *>(dst)[i] = __floats2half2_rn(v.x * scale, v.y * scale);
        else
            reinterpret_cast<float2 *>(dst)[i] = make_float2(v.x * scale, v.y * scale);
    }
}

#define MSOSM_BIG_DISPATCH(L1, CALL)                                                                             \
    switch (L1) {                                                                                                \
    case 1: CALL(1); break;                                                                                      \
    case 2: CALL(2); break;                                                                                      \
    case 3: CALL(3); break;                                                                                      \
    case 4: CALL(4); break;                                                                                      \
    case 5: CALL(5); break;                                                                                      \
    case 6: CALL(6); break;                                                                                      \
    case 7: CALL(7); break;                                                                                      \
    default: return hipErrorInvalidValue;                                                                        \
    }

} // namespace

bool big_supported(int log2n) { return log2n > kRowLog2 && log2n <= kRowLog2 + 7; }

long long big_position_bin(int log2n, long long p)
{
    const long long N1 = 1ll << (log2n - kRowLog2);
    return (p >> kRowLog2) + N1 * (p & (kN2 - 1));
}

int big_chunk(int log2n, long long items, size_t bytes)
{
    long long c = (long long)((bytes ? bytes : kScratchBytes) / (sizeof(float2) << log2n));
    c = std::max(1ll, std::min(c, items));
    return (int)c;
}

hipError_t launch_big_permute(const float2 *src, void *dst, bool f16, float scale, int log2n, long long rows,
                              hipStream_t s)
{
    if (!big_supported(log2n))
        return hipErrorInvalidValue;
    const long long total = rows << log2n;
    long long grid = (total + 255) / 256;
    grid = std::min(grid, 1ll << 20);
    hipLaunchKernelGGL(k_big_permute, dim3((unsigned)grid), dim3(256), 0, s, src, dst, f16, scale,
                       log2n - kRowLog2, rows);
    return hipGetLastError();
}

hipError_t launch_big_forward(const uint32_t *cur, const uint32_t *tail, uint32_t *tail_out, void *ring,
                              float2 *scratch, int chunk, bool f16, int log2n, int count, long long g0, int rx,
                              hipStream_t s)
{
    if (!big_supported(log2n) || count < 1 || chunk < 1)
        return hipErrorInvalidValue;
    const int l1 = log2n - kRowLog2;
    for (int w0 = 0; w0 < count; w0 += chunk) {
        const int nw = std::min(chunk, count - w0);
#define CALL_C(L)                                                                                                \
    if (f16)                                                                                                     \
        hipLaunchKernelGGL((k_big_fwd_cols<L, true>), dim3((unsigned)((long long)nw * (kN2 / ColShape<L>::C))),   \
                           dim3(kColThreads), 0, s, cur, tail, tail_out, scratch, w0, count);                    \
    else                                                                                                         \
        hipLaunchKernelGGL((k_big_fwd_cols<L, false>), dim3((unsigned)((long long)nw * (kN2 / ColShape<L>::C))),  \
                           dim3(kColThreads), 0, s, cur, tail, tail_out, scratch, w0, count)
        MSOSM_BIG_DISPATCH(l1, CALL_C)
#undef CALL_C
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
        const int nrows = nw << l1;
        const int nrpw = runs_per_wg(nrows);
        if (nrpw > 1) {
            const unsigned grid = (unsigned)((nrows + nrpw - 1) / nrpw);
            if (f16)
                hipLaunchKernelGGL((k_big_fwd_rows_nb<true>), dim3(grid), dim3(RowS::NT), 0, s, scratch, ring, w0,
                                   l1, g0, rx, nrows, nrpw);
            else
                hipLaunchKernelGGL((k_big_fwd_rows_nb<false>), dim3(grid), dim3(RowS::NT), 0, s, scratch, ring, w0,
                                   l1, g0, rx, nrows, nrpw);
        } else {
            const unsigned grid = (unsigned)nrows;
            if (f16)
                hipLaunchKernelGGL((k_big_fwd_rows<true>), dim3(grid), dim3(RowS::NT), 0, s, scratch, ring, w0, l1,
                                   g0, rx);
            else
                hipLaunchKernelGGL((k_big_fwd_rows<false>), dim3(grid), dim3(RowS::NT), 0, s, scratch, ring, w0, l1,
                                   g0, rx);
        }
        e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_big_inverse(const void *ring, const void *Hp, const uint16_t *dp, float2 *scratch, int chunk,
                              float2 *out, bool f16, int log2n, int count, int ndm, long long g0, int rx,
                              hipStream_t s)
{
    if (!big_supported(log2n) || count < 1 || ndm < 1 || chunk < 1)
        return hipErrorInvalidValue;
    const int l1 = log2n - kRowLog2;
    const long long N = 1ll << log2n;
    const long long units = (long long)ndm * count;
    // FP16: the products carry the whole scale, the row scratch 2^-9 of it
    const float scale = f16 ? 1.0f / kRowScratchScale : 1.0f / ((float)N * (float)N);
    for (long long u0 = 0; u0 < units; u0 += chunk) {
        const int nu = (int)std::min<long long>(chunk, units - u0);
        const unsigned grid_r = (unsigned)((long long)nu << l1);
        if (f16)
            hipLaunchKernelGGL((k_big_inv_rows<true>), dim3(grid_r), dim3(RowS::NT), 0, s, ring, Hp, dp, scratch,
                               (int)u0, count, l1, g0, rx);
        else
            hipLaunchKernelGGL((k_big_inv_rows<false>), dim3(grid_r), dim3(RowS::NT), 0, s, ring, Hp, dp, scratch,
                               (int)u0, count, l1, g0, rx);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
#define CALL_I(L)                                                                                                \
    if (f16)                                                                                                     \
        hipLaunchKernelGGL((k_big_inv_cols<L, true>), dim3((unsigned)((long long)nu * (kN2 / ColShape<L>::C))),   \
                           dim3(kColThreads), 0, s, scratch, out, (int)u0, scale);                               \
    else                                                                                                         \
        hipLaunchKernelGGL((k_big_inv_cols<L, false>), dim3((unsigned)((long long)nu * (kN2 / ColShape<L>::C))),  \
                           dim3(kColThreads), 0, s, scratch, out, (int)u0, scale)
        MSOSM_BIG_DISPATCH(l1, CALL_I)
#undef CALL_I
        e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

} // namespace msosm
