// invmfma.hip -- fused inverse of the mixed form with its DFT stages on the
// matrix cores (2^14 points: M = 8192).
//
// Per (DM trial, block): gather Y[n] = X_{g-d(n)}[n] H[n] from the FP16
// spectra ring, inverse FFT, keep points [M, 2M) -- the same result as the
// fused inverse in kernels.hip, with the butterflies moved from the vector
// ALUs to v_mfma_f32_16x16x32_f16.  A Stockham FFT with passes of radix 32,
// 32 and 16 (index maps in mfma_fft.h):
//
//   pass 0: each wave takes 4 groups of 16 columns; lanes gather their
//           column's 32 products straight into B operands (FP16), 8 MFMAs
//           per group (4 M-tiles x 2 K-tiles of the 64 x 64 real embedding
//           of DFT_32) give the outputs in FP32; each is multiplied by its
//           pass 1 twiddle, packed to FP16 and stored in LDS (A1);
//   pass 1: the same from A1 (one 16-byte LDS read per B operand) into A2;
//   pass 2: DFT_16 from A2, only the M-tile whose outputs fall into [M, 2M)
//           (one MFMA per group), stored as complex64.
//
// Precision: data enter every pass rounded to FP16 and the DFT matrices are
// FP16; products, accumulations and twiddles are FP32; the filter is stored
// with a factor 1/64 of headroom against FP16 overflow (below).  The transform adds
// about 4e-4 relative L2 to the FP16 storage of ring and filter (measured
// total against the FP64 oracle about 5e-4, gate 1e-3).
//
// One 512-thread workgroup runs 8 or 16 consecutive blocks of one trial (the
// filter and delay rows stay in L2), XCD-aware as the other fused inverse;
// A1 and A2 share one 64 KiB FP16 buffer and registers are capped at 128,
// so two workgroups run per CU.
//
// Measured on MI355X (256 trials x 64 blocks): 0.66 ms per launch against
// 0.52 ms for the vector-ALU fused inverse.  The butterflies cost far less,
// but the swizzled FP16 exchange layouts need per-element integer address
// arithmetic (about 60% of the remaining vector instructions), pass 0's
// columns put lanes on bins 16 apart (gathers touch more ring slots per
// instruction: 0.18 ms), and the pass 2 twiddles cost two LDS reads each.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>

#include "fft_device.h"
#include "kernels.h"
#include "mfma_fft.h"

namespace msosm {
namespace {

using namespace mfft;

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 512; // 8 waves
// blocks per workgroup: the largest power of two up to 16 that leaves at
// least 512 workgroups (256 trials x 64 blocks: 0.653 ms per launch against
// 0.658 at 8; with streaming output stores 0.614 at 16 against 0.630 at 32)
constexpr int kNBWide = 16;
constexpr int kTiles = 9;     // DFT_32: 4 M-tiles x 2 K-tiles; DFT_16: M-tile 1
constexpr int kTwLo = 128, kTwHi = 128; // W_N^i and W_N^(128 i), i < 128
// FP16 headroom: the filter rows are stored divided by 64 (a power of two,
// exact) and the last DFT matrix multiplied by 64.  Data crossing the FP16
// exchanges are partial transforms of the products P = (X/N)(H/(2M)), whose
// norm over all bins is at most max|H/(2M)| * 46341 (|x| <= 32768 sqrt 2,
// Parseval); an output of pass 1 (pass 0) is at most sqrt(1024) (sqrt(32))
// times that.  Divided by 64 this stays below FP16's 65504 for max|H/(2M)|
// up to 2.2 (the filters of 8-128 MHz plans up to DM 500 reach 1.25), while
// typical values (about 1 after the division) stay far from FP16's
// subnormal range.
constexpr int kHeadroom = 64;

// Constant tables, built once per device buffer by k_mfma_tables:
//   tiles[t][lane]  A operand (8 FP16) of tile t for this lane
//   rows[k][r]      W_1024^(k r), k, r < 32, rows padded to 33 (forward sign)
//   lo[i], hi[i]    W_N^i, W_N^(128 i) (forward sign)
struct Tables {
    f16x8 tiles[kTiles][64];
    float2 rows[kTwRowElems];
    float2 lo[kTwLo];
    float2 hi[kTwHi];
};

// entry (2k + a, 2r + b) of the real embedding of the inverse DFT_R:
// [[cos, -sin], [sin, cos]] of 2 pi k r / R
__device__ double dft_entry(int R, int k, int a, int r, int b)
{
    double s, c;
    sincospi(2.0 * (double)((k * r) % R) / (double)R, &s, &c);
    return a == 0 ? (b == 0 ? c : -s) : (b == 0 ? s : c);
}

__global__ void k_mfma_tables(Tables *t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kTiles * 64) {
        const int tile = i / 64, lane = i % 64, q = lane >> 4, c = lane & 15;
        const int R = tile < 8 ? 32 : 16, mt = tile < 8 ? tile >> 1 : 1, kt = tile < 8 ? tile & 1 : 0;
        // DFT_16 (the last pass) carries the factor kHeadroom back (exact)
        const double scale = tile < 8 ? 1.0 : (double)kHeadroom;
        f16x8 a;
        for (int jj = 0; jj < 8; jj++)
            a[jj] = (_Float16)(scale * dft_entry(R, 8 * mt + (c >> 1), c & 1, 16 * kt + 4 * (jj >> 1) + q, jj & 1));
        t->tiles[tile][lane] = a;
    }
    if (i < kTwRowElems) {
        const int k = i / kTwRowStride, r = i - k * kTwRowStride;
        double s, c;
        sincospi(2.0 * (double)((k * r) & 1023) / 1024.0, &s, &c);
        t->rows[i] = make_float2((float)c, (float)(-s));
    }
    if (i < kTwLo + kTwHi) {
        const double e = i < kTwLo ? (double)i : (double)(i - kTwLo) * kTwLo;
        double s, c;
        sincospi(2.0 * e / (double)kN, &s, &c);
        (i < kTwLo ? t->lo[i] : t->hi[i - kTwLo]) = make_float2((float)c, (float)(-s));
    }
}

__device__ __forceinline__ uint32_t pack(float2 v) { return __builtin_bit_cast(uint32_t, __float22half2_rn(v)); }
__device__ __forceinline__ float2 conj(float2 v) { return make_float2(v.x, -v.y); }
__device__ __forceinline__ f32x4 mma(f16x8 a, f16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// x * h in FP32 from FP16 operands (four mixed-precision FMAs)
__device__ __forceinline__ float2 mul_h(__half2 x, __half2 h)
{
    const float xr = __low2float(x), xi = __high2float(x), hr = __low2float(h), hi = __high2float(h);
    const float t1 = __builtin_fmaf(xi, hi, 0.0f), t2 = __builtin_fmaf(xi, hr, 0.0f);
    return make_float2(__builtin_fmaf(xr, hr, -t1), __builtin_fmaf(xr, hi, t2));
}

__global__ __launch_bounds__(kThreads, 2) void k_inv_mfma(const __half2 *__restrict__ ring,
                                                       const __half2 *__restrict__ Hp,
                                                       const uint32_t *__restrict__ dpk,
                                                       const Tables *__restrict__ tab, float2 *__restrict__ out,
                                                       int count, int ndm, long long g0, int rx, int nbpw)
{
    // one 64 KiB exchange buffer for A1 and A2 (extra barriers below), so two
    // workgroups fit one CU's LDS and hide each other's gather latency
    __shared__ __align__(16) uint32_t a12[kN];
    uint32_t *const a1 = a12, *const a2 = a12;
    __shared__ float2 rows[kTwRowElems];
    __shared__ float2 twlo[kTwLo], twhi[kTwHi];
    for (int i = threadIdx.x; i < kTwRowElems; i += kThreads)
        rows[i] = tab->rows[i];
    for (int i = threadIdx.x; i < kTwLo; i += kThreads) {
        twlo[i] = tab->lo[i];
        twhi[i] = tab->hi[i];
    }
    const int lane0 = threadIdx.x & 63, wave = threadIdx.x >> 6;
    f16x8 A32[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int kt = 0; kt < 2; kt++)
            A32[mt][kt] = tab->tiles[2 * mt + kt][lane0];
    const f16x8 A16 = tab->tiles[8][lane0];

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
    const __half2 *hrow = Hp + (long long)dm * kN;
    const uint32_t *drow = dpk + (long long)dm * (kN / 2);
    const char *ringB = reinterpret_cast<const char *>(ring);
    int slot = (int)((g0 + b0) % rx);
    __syncthreads(); // tables in LDS

    for (int blk = 0; blk < nb; blk++) {
        // The lane's addresses do not change from block to block; an opaque
        // copy of the lane index (an empty asm statement) keeps the compiler
        // from computing all of them once before the loop and holding them in
        // registers (it would spill), so they are recomputed per block.
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const int q = lane >> 4, c = lane & 15;
        // ---- pass 0: gather the products, DFT_32, pass 1 twiddle, FP16 -> A1
        // Every gather of the wave's 4 groups is issued before any product is
        // used (one memory round trip per block).  A block before the stream
        // start maps to the slot of a block that no call has written yet (the
        // ring is zeroed at creation and reset), so those gathers read zero
        // without a test.
        f16x8 b[4][2];
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            const int g = 4 * wave + gi;
            if (gi == 2) // gathers in two halves of two groups each
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int kt = 0; kt < 2; kt++) {
                uint32_t hw[4];
#pragma unroll
                for (int ip = 0; ip < 2; ip++) {
                    const uint32_t pair = drow[((g * 2 + kt) * 2 + ip) * 64 + lane];
#pragma unroll
                    for (int e = 0; e < 2; e++) {
                        const int i = 2 * ip + e;
                        const int p = ((g * 2 + kt) * 4 + i) * 64 + lane;
                        const int d = e ? (int)(pair >> 16) : (int)(pair & 0xFFFFu);
                        int sl = slot - d;
                        sl += sl < 0 ? rx : 0;
                        // byte offsets fit 32 bits (the ring is < 4 GiB)
                        const uint32_t off = ((uint32_t)sl * (uint32_t)kN + (uint32_t)p) * 4u;
                        const __half2 x = *reinterpret_cast<const __half2 *>(ringB + off);
                        hw[i] = pack(mul_h(x, hrow[p]));
                    }
                }
                b[gi][kt] = __builtin_bit_cast(f16x8, u32x4{hw[0], hw[1], hw[2], hw[3]});
            }
        }
        // keep the groups' MFMA and store phases apart: the scheduler would
        // otherwise interleave all of them and spill
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            const int g = 4 * wave + gi;
            const int rp = 16 * (g >> 4) + c;
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mma(A32[mt][0], b[gi][0], acc);
                acc = mma(A32[mt][1], b[gi][1], acc);
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int k = 8 * mt + 2 * q + e;
                    const float2 t = conj(rows[k * kTwRowStride + rp]); // W_1024^(k r'), inverse sign
                    a1[a1_addr(32 * (g & 15) + k, rp)] = pack(cmul(make_float2(acc[2 * e], acc[2 * e + 1]), t));
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        slot = slot + 1 == rx ? 0 : slot + 1;
        __syncthreads();
        // ---- pass 1: DFT_32 from A1, pass 2 twiddle, FP16 -> A2
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            const int jp = 32 * c + 4 * wave + gi;
#pragma unroll
            for (int kt = 0; kt < 2; kt++)
                b[gi][kt] = *reinterpret_cast<const f16x8 *>(&a1[a1_addr(jp, 16 * kt + q)]);
        }
        __syncthreads(); // every wave has read A1 before A2 (the same memory) is written
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
            const int g = 4 * wave + gi;
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mma(A32[mt][0], b[gi][0], acc);
                acc = mma(A32[mt][1], b[gi][1], acc);
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int jpp = g + 32 * (8 * mt + 2 * q + e);
                    // W_N^(j'' r''), j'' r'' < N, from the two-level table
                    const float2 t = conj(cmul(twhi[(jpp * c) >> 7], twlo[(jpp * c) & (kTwLo - 1)]));
                    a2[a2_addr(jpp, c)] = pack(cmul(make_float2(acc[2 * e], acc[2 * e + 1]), t));
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        // ---- pass 2: DFT_16 from A2, outputs 8..15 (points [M, 2M)) only
        float2 *dst = out + ((long long)dm * count + b0 + blk) * kM;
        f16x8 bb[8];
#pragma unroll
        for (int gi = 0; gi < 8; gi++)
            bb[gi] = *reinterpret_cast<const f16x8 *>(&a2[a2_addr(16 * (8 * wave + gi) + c, q)]);
#pragma unroll
        for (int gi = 0; gi < 8; gi++) {
            const int jpp = 16 * (8 * wave + gi) + c;
            const f32x4 acc = mma(A16, bb[gi], f32x4{0.f, 0.f, 0.f, 0.f});
            store_stream(&dst[jpp + 1024 * (2 * q)], make_float2(acc[0], acc[1]));
            store_stream(&dst[jpp + 1024 * (2 * q + 1)], make_float2(acc[2], acc[3]));
        }
        __syncthreads(); // every wave has read A2 before the next block's pass 0 writes A1
    }
}

__global__ void k_permute_h2(__half2 *__restrict__ dst, const __half2 *__restrict__ src, long long rows, float scale)
{
    const long long total = rows * kN;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long r = i / kN;
        const int n = (int)(i - r * kN);
        const float2 v = __half22float2(src[i]);
        dst[r * kN + ring_pos(n)] = __float22half2_rn(make_float2(v.x * scale, v.y * scale));
    }
}

} // namespace

size_t inv_mfma_tables_bytes() { return sizeof(Tables); }

hipError_t launch_inv_mfma_tables(void *tables, hipStream_t s)
{
    const int n = kTiles * 64 > kTwRowElems ? kTiles * 64 : kTwRowElems;
    hipLaunchKernelGGL(k_mfma_tables, dim3((n + 255) / 256), dim3(256), 0, s, static_cast<Tables *>(tables));
    return hipGetLastError();
}

hipError_t launch_inv_mfma_filter(__half2 *dst, const __half2 *src, long long rows, hipStream_t s)
{
    const long long total = rows * kN;
    long long grid = (total + 255) / 256;
    if (grid > 65536)
        grid = 65536;
    hipLaunchKernelGGL(k_permute_h2, dim3((unsigned)grid), dim3(256), 0, s, dst, src, rows,
                       1.0f / (float)kHeadroom);
    return hipGetLastError();
}

hipError_t launch_inv_mfma(const __half2 *ring, const __half2 *Hp, const uint32_t *dpk, const void *tables,
                           float2 *out, int count, int ndm, long long g0, int rx, int nbpw_req, hipStream_t s)
{
    const int nbpw = nbpw_req > 0 ? std::min(nbpw_req, count) : blocks_per_wg(count, ndm, kNBWide);
    const long long chunks = (count + nbpw - 1) / nbpw;
    const long long grid = chunks * ndm;
    hipLaunchKernelGGL(k_inv_mfma, dim3((unsigned)grid), dim3(kThreads), 0, s, ring, Hp, dpk,
                       static_cast<const Tables *>(tables), out, count, ndm, g0, rx, nbpw);
    return hipGetLastError();
}

} // namespace msosm
