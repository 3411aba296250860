// hankel.hip -- the delay-line accumulate recast as per-bin complex GEMMs on
// the gfx950 matrix cores (MFMA), plus the inverse transform that consumes
// their dense output.
//
// For frequency bin n, with X_a[n] the ring of forward spectra (block a) and
// d_t(n) the segment delay of DM trial t, the reference's multiply-delay
// (gpu_kernel.cu:84-100, 176-196) writes Y_t,g[n] = H_t[n] X_{g-d_t(n)}[n]
// for every output block g.  As a product of matrices:
//
//     Y_n[t][g] = sum_k A_n[t][k] B_n[k][g],   A_n[t][k] = H_t[n] [k == d_t(n)],
//                                            B_n[k][g] = X_{g-k}[n]  (Hankel in g, k)
//
// A_n has one non-zero per row, so most MACs of a dense product are wasted.
// The kernel runs the product over the band of k that a tile of 16 trials
// actually uses at bin n ([min_t d_t(n), max_t d_t(n)], one or two 16-wide
// k steps when the trials of a tile are close in DM) and skips the rest, so
// issued MACs stay within a small factor of the useful ones.
//
// Complex arithmetic uses the real embedding along k:
//     Re Y = [Re A | -Im A] [Re B ; Im B],   Im Y = [Im A | Re A] [Re B ; Im B]
// so one 32-deep MFMA step covers 16 complex k.  Operand forms:
//   MODE_F16     FP16 ring X/N and FP16 H/(2M), v_mfma_f32_16x16x32_f16,
//                Y stored FP16 as their product
//   MODE_BF16    FP32 ring X and FP32 H rounded to BF16, one product,
//                Y stored FP32 as X H
//   MODE_BF16X3  as MODE_BF16 with split operands x = hi + lo (both BF16)
//                and three products hi*hi + hi*lo + lo*hi (FP32 accumulate)
//
// MFMA 16x16x32 operand layout (64 lanes, lane l, element j = 0..7):
//   A[i][k]: i = l % 16, k = 8 (l / 16) + j
//   B[k][c]: c = l % 16, k = 8 (l / 16) + j
//   C[i][c] (4 floats): c = l % 16, i = 4 (l / 16) + j
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "fft_device.h"
#include "kernels.h"

namespace msosm {

namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kDmTile = 16; // trials per MFMA tile (rows)
constexpr size_t kHankelThreePerCu = 160 * 1024 / 3;
constexpr int kGTile = 16;  // output blocks per workgroup (columns)
constexpr int kWaves = 4;

template <int MODE>
struct HankelMode;

template <>
struct HankelMode<HANKEL_F16> {
    using RingT = __half2;
    using YT = __half2;
    using Op = _Float16;
    using OpV = f16x8;
    static constexpr int NBIN = 32; // 32 bins x 4 B = one 128 B line per output row
    static constexpr int TERMS = 1;
    static constexpr bool SPLIT = false;
    __device__ static float2 load(const RingT *p) { return __half22float2(*p); }
    __device__ static f32x4 mma(OpV a, OpV b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
    __device__ static YT out(float re, float im) { return __floats2half2_rn(re, im); }
};

template <>
struct HankelMode<HANKEL_BF16> {
    using RingT = float2;
    using YT = float2;
    using Op = __bf16;
    using OpV = bf16x8;
    static constexpr int NBIN = 16; // 16 bins x 8 B = one 128 B line per output row
    static constexpr int TERMS = 1;
    static constexpr bool SPLIT = false;
    __device__ static float2 load(const RingT *p) { return *p; }
    __device__ static f32x4 mma(OpV a, OpV b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
    __device__ static YT out(float re, float im) { return make_float2(re, im); }
};

template <>
struct HankelMode<HANKEL_BF16X3> : HankelMode<HANKEL_BF16> {
    static constexpr int TERMS = 3;
    static constexpr bool SPLIT = true;
};

// Operand parts kept in LDS per (bin, u): real and imaginary, and for the
// split modes their low parts.
template <int MODE>
struct Parts {
    static constexpr int P = HankelMode<MODE>::SPLIT ? 4 : 2;
};

template <typename Op>
__device__ __forceinline__ Op to_op(float x)
{
    return (Op)x;
}

template <typename Op>
__device__ __forceinline__ float from_op(Op x)
{
    return (float)x;
}

template <int MODE>
__device__ __forceinline__ void split(float x, typename HankelMode<MODE>::Op &hi, typename HankelMode<MODE>::Op &lo)
{
    using Op = typename HankelMode<MODE>::Op;
    hi = to_op<Op>(x);
    lo = HankelMode<MODE>::SPLIT ? to_op<Op>(x - from_op<Op>(hi)) : to_op<Op>(0.f);
}

// Product rows are read back only by the next kernel, from HBM (they are far
// larger than the caches): streaming stores keep this kernel's ring columns
// and filter rows in L2.
__device__ __forceinline__ void store_row(__half2 *p, __half2 v)
{
    __builtin_nontemporal_store(__builtin_bit_cast(uint32_t, v), reinterpret_cast<uint32_t *>(p));
}
__device__ __forceinline__ void store_row(float2 *p, float2 v) { store_stream(p, v); }

// One workgroup: NBIN consecutive bins x kGTile consecutive output blocks,
// looping over all DM tiles.  LDS (dynamic):
//   z[P][NBIN][U]   ring column parts, u = kGTile-1 - (g - gb) + k
//   dd[kDmTile][NBIN], hh[kDmTile][NBIN] (float2)   one DM tile's tables
//   ys[kDmTile][kGTile][NBIN+1]                      staged output
template <int MODE>
__global__ __launch_bounds__(64 * kWaves) void k_hankel(const void *__restrict__ ring_, const void *__restrict__ H_,
                                                        const uint16_t *__restrict__ delay, void *__restrict__ Y_,
                                                        int N, int count, int ndm, long long g0, int rx, int kmax,
                                                        int U, int tpg, const int *__restrict__ band_lo)
{
    using Md = HankelMode<MODE>;
    using RingT = typename Md::RingT;
    using YT = typename Md::YT;
    using Op = typename Md::Op;
    using OpV = typename Md::OpV;
    constexpr int NBIN = Md::NBIN;
    constexpr int P = Parts<MODE>::P;
    const RingT *ring = reinterpret_cast<const RingT *>(ring_);
    const RingT *H = reinterpret_cast<const RingT *>(H_);
    YT *Y = reinterpret_cast<YT *>(Y_);

    extern __shared__ __align__(16) unsigned char lds[];
    Op *z = reinterpret_cast<Op *>(lds);                                  // [P][NBIN][U]
    size_t off = ((size_t)P * NBIN * U * sizeof(Op) + 15) & ~(size_t)15;
    int *dd = reinterpret_cast<int *>(lds + off);                         // [kDmTile][NBIN]
    off += sizeof(int) * kDmTile * NBIN;
    float2 *hh = reinterpret_cast<float2 *>(lds + off);                   // [kDmTile][NBIN]
    off += sizeof(float2) * kDmTile * NBIN;
    YT *ys = reinterpret_cast<YT *>(lds + off);                           // [kDmTile][kGTile][NBIN+1]
    constexpr int YS_ROW = NBIN + 1;

    const int n0 = blockIdx.x * NBIN;
    const int gb = blockIdx.y * kGTile;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;

    // This workgroup's trials: DM tiles [tg0, tg1) (grid z), and the band of
    // delays [lo_w, hi_w] they take at its bins; only ring columns u >= lo_w
    // are staged (LDS column u' = u - lo_w, U >= hi_w - lo_w + 31 columns)
    const int ntiles = (ndm + kDmTile - 1) / kDmTile;
    const int tg0 = blockIdx.z * tpg, tg1 = min(ntiles, tg0 + tpg);
    // (band_lo[group][bin group] from the host's delay table; none: 0)
    const int lo_w = band_lo ? band_lo[(size_t)blockIdx.z * gridDim.x + blockIdx.x] : 0;

    // ring columns of this workgroup's bins: u -> absolute block
    // a = g0 + gb + kGTile - 1 - u; zero before the stream start and past the
    // oldest block any (g, k) of the tile uses
    // (batches of kLoadBatch independent loads per thread, issued before any
    // is converted, so they are in flight together)
    // 32-bit slot arithmetic: slot(u) = (s0 - u) mod rx with s0 the slot of
    // block g0 + gb + kGTile - 1 (a true modulo: with few blocks per call and
    // a short delay line, u can exceed s0 by several times rx).  Blocks past
    // the newest one of this call (u < ufirst: they only feed output columns
    // beyond count) and before the stream start (u > g0 + gb + kGTile - 1)
    // read as zero.
    constexpr int kLoadBatch = 8;
    const long long top = g0 + gb + kGTile - 1;
    const int s0 = (int)(top % rx);
    const int ufirst = max(0, gb + kGTile - count);                     // block top - u <= g0 + count - 1
    const int ulim = (int)min(top + 1, (long long)(kGTile - 1 + kmax)); // loaded u < ulim
    for (int base = 0; base < NBIN * U; base += kLoadBatch * 64 * kWaves) {
        RingT raw[kLoadBatch];
        bool use[kLoadBatch];
#pragma unroll
        for (int r = 0; r < kLoadBatch; r++) {
            const int idx = base + r * 64 * kWaves + tid;
            const int nl = idx % NBIN, u = lo_w + idx / NBIN;
            use[r] = idx < NBIN * U && u >= ufirst && u < ulim;
            int sl = (s0 - (use[r] ? u : 0)) % rx;
            sl += sl < 0 ? rx : 0;
            raw[r] = ring[(size_t)sl * N + n0 + nl];
        }
#pragma unroll
        for (int r = 0; r < kLoadBatch; r++) {
            const int idx = base + r * 64 * kWaves + tid;
            if (idx >= NBIN * U)
                break;
            const int nl = idx % NBIN, u = idx / NBIN; // LDS column
            const float2 x = use[r] ? Md::load(&raw[r]) : make_float2(0.f, 0.f);
            Op rh, rl, ih, il;
            split<MODE>(x.x, rh, rl);
            split<MODE>(x.y, ih, il);
            z[(0 * NBIN + nl) * U + u] = rh;
            z[(1 * NBIN + nl) * U + u] = ih;
            if constexpr (P == 4) {
                z[(2 * NBIN + nl) * U + u] = rl;
                z[(3 * NBIN + nl) * U + u] = il;
            }
        }
    }

    const int row = lane & 15;      // A row (trial in tile) / B column (block in tile)
    const int q = lane >> 4;        // k group: 0,1 real rows, 2,3 imaginary rows
    const int kq = 8 * (q & 1);
    const bool imag_half = q >= 2;
    // one DM tile's delays and filter values, kDmTile * NBIN entries: at most
    // PER per thread, loaded one tile ahead so the loads overlap the MFMA
    // work of the current tile
    constexpr int PER = (kDmTile * NBIN + 64 * kWaves - 1) / (64 * kWaves);
    uint16_t dr[PER];
    RingT hr[PER];
    auto load_tables = [&](int tl) {
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int idx = tid + r * 64 * kWaves;
            const int i = idx / NBIN, nl = idx % NBIN;
            const int dm = min(tl * kDmTile + i, ndm - 1);
            dr[r] = delay[(long long)dm * N + n0 + nl];
            hr[r] = H[(long long)dm * N + n0 + nl];
        }
    };
    load_tables(tg0);
    for (int tile = tg0; tile < tg1; tile++) {
        __syncthreads(); // z ready (first tile); tables and staging free
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int idx = tid + r * 64 * kWaves;
            if (idx < kDmTile * NBIN) {
                const bool ok = tile * kDmTile + idx / NBIN < ndm;
                dd[idx] = ok ? (int)dr[r] : -1;
                hh[idx] = ok ? Md::load(&hr[r]) : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
        if (tile + 1 < tg1)
            load_tables(tile + 1);
        for (int nl = wave; nl < NBIN; nl += kWaves) {
            const int d = dd[row * NBIN + nl];
            const float2 h = hh[row * NBIN + nl];
            int lo = d >= 0 ? d : INT_MAX, hi = d;
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) {
                lo = min(lo, __shfl_xor(lo, m));
                hi = max(hi, __shfl_xor(hi, m));
            }
            f32x4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
            // A entries: Re rows [Re h | -Im h], Im rows [Im h | Re h]
            Op ar_h, ar_l, ai_h, ai_l;
            split<MODE>(imag_half ? -h.y : h.x, ar_h, ar_l);
            split<MODE>(imag_half ? h.x : h.y, ai_h, ai_l);
            const Op zero = to_op<Op>(0.f);
            const Op *zb = z + ((imag_half ? 1 : 0) * NBIN + nl) * U + (kGTile - 1 - row) + kq - lo_w;
            const Op *zl = z + ((imag_half ? 3 : 2) * NBIN + nl) * U + (kGTile - 1 - row) + kq - lo_w;
            for (int k0 = lo; k0 <= hi; k0 += 16) { // wave-uniform bounds
                OpV a_r, a_i, b;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const bool hit = (k0 + kq + j) == d;
                    a_r[j] = hit ? ar_h : zero;
                    a_i[j] = hit ? ai_h : zero;
                    b[j] = zb[k0 + j];
                }
                cr = Md::mma(a_r, b, cr);
                ci = Md::mma(a_i, b, ci);
                if constexpr (Md::TERMS == 3) {
                    OpV al_r, al_i, bl;
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const bool hit = (k0 + kq + j) == d;
                        al_r[j] = hit ? ar_l : zero;
                        al_i[j] = hit ? ai_l : zero;
                        bl[j] = zl[k0 + j];
                    }
                    cr = Md::mma(a_r, bl, cr);
                    ci = Md::mma(a_i, bl, ci);
                    cr = Md::mma(al_r, b, cr);
                    ci = Md::mma(al_i, b, ci);
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int i = 4 * q + j; // trial in tile
                ys[(i * kGTile + row) * YS_ROW + nl] = Md::out(cr[j], ci[j]);
            }
        }
        __syncthreads();
        // write the tile: rows (trial, block), NBIN consecutive bins each
        for (int idx = tid; idx < kDmTile * kGTile * NBIN; idx += blockDim.x) {
            const int nl = idx % NBIN, r = idx / NBIN;
            const int i = r / kGTile, g = r % kGTile;
            const int dm = tile * kDmTile + i;
            if (dm < ndm && gb + g < count)
                store_row(&Y[((long long)dm * count + gb + g) * N + n0 + nl], ys[r * YS_ROW + nl]);
        }
    }
}

// Inverse transform of dense rows Y[dm][b][N] (one per (trial, block)), keep
// the last M points, scale; out layout [dm][count][M].
template <int LOG2N, bool F16Y>
__global__ __launch_bounds__(fft_threads(LOG2N)) void k_inv_dense(const void *__restrict__ Y_, float2 *__restrict__ out,
                                                                   long long units, float scale)
{
    constexpr int E = fft_e(LOG2N);
    using S = FftShape<LOG2N, E>;
    constexpr int TPB = fft_tpb(LOG2N);
    constexpr long long N = S::N, M = S::N / 2;
    __shared__ float2 smem[TPB * S::NPAD + S::TWLO + S::TWHI];
    float2 *twlo = smem + TPB * S::NPAD, *twhi = twlo + S::TWLO;
    if constexpr (S::NPASS > 1)
        fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    const int tl = threadIdx.x / S::NT, t = threadIdx.x % S::NT;
    const long long u = (long long)blockIdx.x * TPB + tl;
    const bool valid = u < units;
    // all E loads are issued before any is used (one guarded block, so the
    // compiler keeps them in flight together instead of load-wait-convert)
    using Raw = typename std::conditional<F16Y, __half2, float2>::type;
    const Raw *src = reinterpret_cast<const Raw *>(Y_) + (valid ? u : 0) * N + t;
    Raw raw[E];
    if (valid) {
#pragma unroll
        for (int m = 0; m < E; m++)
            raw[m] = src[m * S::NT];
    } else {
#pragma unroll
        for (int m = 0; m < E; m++)
            raw[m] = Raw{};
    }
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
        if constexpr (F16Y)
            v[m] = __half22float2(raw[m]);
        else
            v[m] = raw[m];
    }
    fft_block<LOG2N, E, true>(v, smem + tl * S::NPAD, twlo, twhi, t);
    if (valid) {
        float2 *dst = out + u * M;
#pragma unroll
        for (int m = E / 2; m < E; m++)
            store_stream(&dst[t + m * S::NT - M], cscale(v[m], scale));
    }
}

// The same for sizes with one transform per workgroup: nbpw consecutive
// units (blocks of one trial) per workgroup, so the twiddle tables are filled
// once per nbpw transforms, and 2^14 points take the middle pass's twiddles
// from the row table.  nbpw is the largest power of two up to 32 that leaves
// at least 512 workgroups (256 trials x 64 blocks of 2^14, streaming stores:
// FP16 rows 0.481 ms per launch at 32, 0.487 at 16, 0.512 at 8 before the
// streaming stores).
template <int LOG2N, bool F16Y>
__global__ __launch_bounds__((1 << LOG2N) / fft_e(LOG2N)) void k_inv_dense_nb(const void *__restrict__ Y_,
                                                                            float2 *__restrict__ out,
                                                                            long long units, float scale, int nbpw)
{
    constexpr int E = fft_e(LOG2N);
    using S = FftShape<LOG2N, E>;
    static_assert(S::NT >= 256, "one transform per workgroup");
    constexpr long long N = S::N, M = S::N / 2;
    constexpr int NT = S::NT;
    constexpr bool ROWS = (LOG2N == 14 && E == 32);
    __shared__ float2 smem[S::NPAD + S::TWLO + S::TWHI + (ROWS ? kTwRowElems : 0)];
    float2 *twlo = smem + S::NPAD, *twhi = twlo + S::TWLO, *twrows = twhi + S::TWHI;
    fill_twiddles<LOG2N, E>(twlo, twhi, threadIdx.x, blockDim.x);
    if constexpr (ROWS)
        fill_twrows(twrows, threadIdx.x, blockDim.x);
    const int t = threadIdx.x;
    const long long u0 = (long long)blockIdx.x * nbpw;
    const int nb = (int)(units - u0 < nbpw ? units - u0 : nbpw);
    using Raw = typename std::conditional<F16Y, __half2, float2>::type;
    // 2^14 points: block i + 1 is loaded while block i is transformed (the
    // products were just written, so these loads come from HBM); the
    // workgroup's LDS already limits the CU to one, so the extra registers
    // cost no occupancy (at smaller sizes they would)
    constexpr bool PF = LOG2N == 14;
    Raw raw[E];
    if constexpr (PF) {
        const Raw *src = reinterpret_cast<const Raw *>(Y_) + u0 * N + t;
#pragma unroll
        for (int m = 0; m < E; m++)
            raw[m] = src[m * NT];
    }
    for (int i = 0; i < nb; i++) {
        const long long u = u0 + i;
        if constexpr (!PF) {
            const Raw *src = reinterpret_cast<const Raw *>(Y_) + u * N + t;
#pragma unroll
            for (int m = 0; m < E; m++)
                raw[m] = src[m * NT];
        }
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            if constexpr (F16Y)
                v[m] = __half22float2(raw[m]);
            else
                v[m] = raw[m];
        }
        if (PF && i + 1 < nb) {
            const Raw *src = reinterpret_cast<const Raw *>(Y_) + (u + 1) * N + t;
#pragma unroll
            for (int m = 0; m < E; m++)
                raw[m] = src[m * NT];
        }
        fft_block<LOG2N, E, true, ROWS>(v, smem, twlo, twhi, t, twrows);
        float2 *dst = out + u * M;
#pragma unroll
        for (int m = E / 2; m < E; m++)
            store_stream(&dst[t + m * NT - M], cscale(v[m], scale));
    }
}

template <int MODE>
size_t hankel_lds_bytes(int U)
{
    using Md = HankelMode<MODE>;
    size_t b = ((size_t)Parts<MODE>::P * Md::NBIN * U * sizeof(typename Md::Op) + 15) & ~(size_t)15;
    b += sizeof(int) * kDmTile * Md::NBIN + sizeof(float2) * kDmTile * Md::NBIN;
    b += sizeof(typename Md::YT) * kDmTile * kGTile * (Md::NBIN + 1);
    return b;
}

template <int MODE>
hipError_t hankel_launch(const void *ring, const void *H, const uint16_t *delay, void *Y, int N, int count, int ndm,
                         long long g0, int rx, int kmax, int tpg, int kspan, const int *band_lo, hipStream_t s)
{
    using Md = HankelMode<MODE>;
    const int ntiles = (ndm + kDmTile - 1) / kDmTile;
    if (!band_lo || tpg <= 0 || tpg >= ntiles || kspan <= 0 || kspan > kmax) {
        tpg = ntiles; // one group staging the whole line [0, kmax)
        kspan = kmax;
        band_lo = nullptr;
    }
    if (N % Md::NBIN != 0)
        return hipErrorInvalidValue;
    const int U = kGTile - 1 + kspan + 16;
    const size_t lds = hankel_lds_bytes<MODE>(U);
    if (lds > kHankelMaxLds)
        return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_hankel<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kHankelMaxLds);
        if (e != hipSuccess)
            return e;
        attr_set = true;
    }
    dim3 grid(N / Md::NBIN, (count + kGTile - 1) / kGTile, (ntiles + tpg - 1) / tpg);
    hipLaunchKernelGGL(k_hankel<MODE>, grid, dim3(64 * kWaves), lds, s, ring, H, delay, Y, N, count, ndm, g0, rx,
                       kmax, U, tpg, band_lo);
    return hipGetLastError();
}

template <int LOG2N>
hipError_t inv_dense(const void *Y, float2 *out, bool f16y, long long units, float scale, int nbpw_req, hipStream_t s)
{
    constexpr int E = fft_e(LOG2N);
    constexpr int threads = FftShape<LOG2N, E>::NT * fft_tpb(LOG2N);
    if constexpr (fft_tpb(LOG2N) == 1) {
        const int nbpw = nbpw_req > 0 ? (int)std::min<long long>(nbpw_req, units)
                                      : blocks_per_wg((int)std::min<long long>(units, INT_MAX), 1, 32);
        const long long grid = (units + nbpw - 1) / nbpw;
        if (f16y)
            hipLaunchKernelGGL((k_inv_dense_nb<LOG2N, true>), dim3((unsigned)grid), dim3(threads), 0, s, Y, out,
                               units, scale, nbpw);
        else
            hipLaunchKernelGGL((k_inv_dense_nb<LOG2N, false>), dim3((unsigned)grid), dim3(threads), 0, s, Y, out,
                               units, scale, nbpw);
        return hipGetLastError();
    }
    const long long grid = (units + fft_tpb(LOG2N) - 1) / fft_tpb(LOG2N);
    if (f16y)
        hipLaunchKernelGGL((k_inv_dense<LOG2N, true>), dim3((unsigned)grid), dim3(threads), 0, s, Y, out, units,
                           scale);
    else
        hipLaunchKernelGGL((k_inv_dense<LOG2N, false>), dim3((unsigned)grid), dim3(threads), 0, s, Y, out, units,
                           scale);
    return hipGetLastError();
}

} // namespace

int hankel_max_kmax(int mode)
{
    // largest kmax whose LDS footprint fits kHankelMaxLds
    int lo = 1, hi = 1 << 16;
    auto fits = [&](int k) {
        const int U = kGTile - 1 + k + 16;
        size_t b = mode == HANKEL_F16 ? hankel_lds_bytes<HANKEL_F16>(U)
                   : mode == HANKEL_BF16 ? hankel_lds_bytes<HANKEL_BF16>(U)
                                         : hankel_lds_bytes<HANKEL_BF16X3>(U);
        return b <= kHankelMaxLds;
    };
    if (!fits(lo))
        return 0;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (fits(mid))
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

int hankel_bins_per_wg(int mode) { return mode == HANKEL_F16 ? HankelMode<HANKEL_F16>::NBIN : HankelMode<HANKEL_BF16>::NBIN; }

static size_t hankel_lds_of(int mode, int U)
{
    return mode == HANKEL_F16 ? hankel_lds_bytes<HANKEL_F16>(U)
           : mode == HANKEL_BF16 ? hankel_lds_bytes<HANKEL_BF16>(U)
                                 : hankel_lds_bytes<HANKEL_BF16X3>(U);
}

void hankel_groups(int mode, const int32_t *delay, int ndm, int N, int kmax, int tpg_req, int *tpg, int *kspan,
                   std::vector<int> *band_lo)
{
    const int nbin = hankel_bins_per_wg(mode), ngrp = N / nbin;
    const int ntiles = (ndm + kDmTile - 1) / kDmTile;
    // per (trial, bin group): lowest and highest delay
    std::vector<int> lo((size_t)ndm * ngrp, INT_MAX), hi((size_t)ndm * ngrp, 0);
    for (int d = 0; d < ndm; d++)
        for (int n = 0; n < N; n++) {
            const size_t i = (size_t)d * ngrp + n / nbin;
            const int v = delay[(size_t)d * N + n];
            lo[i] = std::min(lo[i], v);
            hi[i] = std::max(hi[i], v);
        }
    // widest band (hi - lo + 1) over all workgroups when groups hold t tiles
    auto span = [&](int t) {
        int w = 1;
        for (int g0 = 0; g0 < ntiles; g0 += t)
            for (int b = 0; b < ngrp; b++) {
                int l = INT_MAX, h = 0;
                for (int d = g0 * kDmTile; d < std::min(ndm, (g0 + t) * kDmTile); d++) {
                    l = std::min(l, lo[(size_t)d * ngrp + b]);
                    h = std::max(h, hi[(size_t)d * ngrp + b]);
                }
                w = std::max(w, h - l + 1);
            }
        return std::min(w, kmax);
    };
    // the largest group (fewest repeated ring columns) whose LDS leaves room
    // for three workgroups per CU (the staged output alone takes 34 KB for
    // FP16), unless the caller asks for a group size (tuning); a requested
    // size whose band does not fit the LDS limit is reduced like the rule
    int t = ntiles, w = kmax;
    if (tpg_req > 0) {
        t = std::min(tpg_req, ntiles);
        for (;;) {
            w = span(t);
            if (hankel_lds_of(mode, kGTile - 1 + w + 16) <= kHankelMaxLds || t == 1)
                break;
            t = (t + 1) / 2;
        }
    } else {
        for (;;) {
            w = span(t);
            if (hankel_lds_of(mode, kGTile - 1 + w + 16) <= kHankelThreePerCu || t == 1)
                break;
            t = (t + 1) / 2;
        }
    }
    *tpg = t;
    *kspan = w;
    // lowest delay per (group, bin group), the first column a workgroup stages
    band_lo->assign((size_t)((ntiles + t - 1) / t) * ngrp, INT_MAX);
    for (int g = 0; g * t < ntiles; g++)
        for (int d = g * t * kDmTile; d < std::min(ndm, (g + 1) * t * kDmTile); d++)
            for (int b = 0; b < ngrp; b++)
                (*band_lo)[(size_t)g * ngrp + b] = std::min((*band_lo)[(size_t)g * ngrp + b], lo[(size_t)d * ngrp + b]);
}

hipError_t launch_hankel(const void *ring, const void *H, const uint16_t *delay, void *Y, int mode, int log2n,
                         int count, int ndm, long long g0, int rx, int kmax, int tpg, int kspan, const int *band_lo,
                         hipStream_t s)
{
    const int N = 1 << log2n;
    switch (mode) {
    case HANKEL_F16:
        return hankel_launch<HANKEL_F16>(ring, H, delay, Y, N, count, ndm, g0, rx, kmax, tpg, kspan, band_lo, s);
    case HANKEL_BF16:
        return hankel_launch<HANKEL_BF16>(ring, H, delay, Y, N, count, ndm, g0, rx, kmax, tpg, kspan, band_lo, s);
    case HANKEL_BF16X3:
        return hankel_launch<HANKEL_BF16X3>(ring, H, delay, Y, N, count, ndm, g0, rx, kmax, tpg, kspan, band_lo, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_inv_dense(const void *Y, float2 *out, bool f16y, int log2n, long long units, float scale,
                            int nbpw, hipStream_t s)
{
#define CALL_D(L) inv_dense<L>(Y, out, f16y, units, scale, nbpw, s)
    switch (log2n) {
    case 2: return CALL_D(2);
    case 3: return CALL_D(3);
    case 4: return CALL_D(4);
    case 5: return CALL_D(5);
    case 6: return CALL_D(6);
    case 7: return CALL_D(7);
    case 8: return CALL_D(8);
    case 9: return CALL_D(9);
    case 10: return CALL_D(10);
    case 11: return CALL_D(11);
    case 12: return CALL_D(12);
    case 13: return CALL_D(13);
    case 14: return CALL_D(14);
    default: return hipErrorInvalidValue;
    }
#undef CALL_D
}

} // namespace msosm
