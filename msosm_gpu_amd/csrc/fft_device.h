// fft_device.h -- hand-written CDNA4 FFT building blocks (device side).
//
// One transform of N = 2^LOG2N complex points is spread over NT = N/E threads
// of a workgroup; thread t holds the E points t + m*NT (m = 0..E-1) in
// registers ("cyclic" layout) on entry and the E outputs with the same
// indices on exit.  Inside, the transform is a Stockham autosort FFT (the
// radix-R formulation of Govindaraju et al., "High performance discrete
// Fourier transforms on graphics processors", SC08): every pass does radix-R
// butterflies in registers and exchanges through LDS; the last pass may use a
// smaller radix.  Callers fuse their own loads (unpack, gather x filter) and
// stores (discard, scale, FP16 packing) around fft_block().
//
// Twiddles W_N^e = exp(-2 pi i e/N) come from two small LDS tables
// (W = hi[e >> LS] * lo[e & (2^LS-1)]) copied once per workgroup from
// read-only tables evaluated at compile time (twiddle_tables.h).
// LDS addresses are padded by one element every E elements so the strided
// Stockham writes do not all land in the same banks.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include "twiddle_tables.h"

namespace msosm {

__host__ __device__ constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

// s_waitcnt vmcnt(0) as an immediate (gfx9 encoding: vmcnt in bits 3:0 and
// 15:14, expcnt in 6:4, lgkmcnt in 11:8; the counters at their maxima are not
// waited for)
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulconj(float2 a, float2 b) // a * conj(b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// Store of output that nothing reads again before the kernel ends: a
// streaming (non-temporal) store, so that the output stream does not evict
// the tables and spectra that the kernel's gathers re-read from L2.
__device__ __forceinline__ void store_stream(float2 *p, float2 v)
{
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// cos(2 pi k / 64) at compile time (Taylor series on a reduced argument;
// double precision, rounded to float by the caller).
__host__ __device__ constexpr double ccos_reduced(double x)
{
    double term = 1.0, sum = 1.0;
    for (int i = 1; i < 30; i++) {
        term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
__host__ __device__ constexpr double cos64(int k)
{
    k &= 63;
    // reduce to [0, pi/2] with exact symmetries
    return k == 0 ? 1.0
         : k == 16 ? 0.0
         : k == 32 ? -1.0
         : k == 48 ? 0.0
         : k < 16 ? ccos_reduced(2.0 * 3.14159265358979323846 * k / 64.0)
         : k < 32 ? -ccos_reduced(2.0 * 3.14159265358979323846 * (32 - k) / 64.0)
         : k < 48 ? -ccos_reduced(2.0 * 3.14159265358979323846 * (k - 32) / 64.0)
                  : ccos_reduced(2.0 * 3.14159265358979323846 * (64 - k) / 64.0);
}
__host__ __device__ constexpr double sin64(int k) { return cos64(k + 48); } // sin t = cos(t - pi/2)

// Multiply by W_R^e (forward: exp(-2 pi i e/R), inverse: conjugate), R <= 64,
// with the trivial angles resolved at compile time.
template <int R, bool INV, int E_>
__device__ __forceinline__ float2 twiddle_const(float2 a)
{
    constexpr int k = ((E_ % R) * (64 / R)) & 63; // angle index in 64ths
    if constexpr (k == 0) {
        return a;
    } else if constexpr (k == 16) { // -i (fwd) / +i (inv)
        return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
    } else if constexpr (k == 32) {
        return make_float2(-a.x, -a.y);
    } else if constexpr (k == 48) { // +i (fwd) / -i (inv)
        return INV ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
    } else {
        constexpr float c = (float)cos64(k);
        constexpr float s = INV ? (float)sin64(k) : -(float)sin64(k);
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// In-register DFT of size R, in place: natural order in, output X[k] left at
// register pos(k) (a compile-time digit permutation, absorbed by the
// caller's indexing, so no temporary copy of the R values is needed).
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<1, INV> {
    __host__ __device__ static constexpr int pos(int k) { return k; }
    __device__ __forceinline__ static void run(float2 *) {}
};

template <bool INV>
struct Dft<2, INV> {
    __host__ __device__ static constexpr int pos(int k) { return k; }
    __device__ __forceinline__ static void run(float2 *v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<4, INV> {
    __host__ __device__ static constexpr int pos(int k) { return k; }
    __device__ __forceinline__ static void run(float2 *v)
    {
        float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        float2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        // forward: X1 = d02 - i d13, X3 = d02 + i d13
        float2 jd13 = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, jd13);
        v[3] = csub(d02, jd13);
    }
};

// R = A*B with A = 4: n = B*n1 + n2, k = k1 + A*k2.  Stage 1 (DFT_A over n1
// for every n2, then the W_R^(n2 k1) twiddle) leaves Y[n2][k1] at v[B*k1 +
// n2]; stage 2 is an in-place DFT_B on each contiguous run v[B*k1 .. +B).
template <int R, bool INV>
struct Dft {
    static constexpr int A = 4;
    static constexpr int B = R / A;
    __host__ __device__ static constexpr int pos(int k) { return B * (k % A) + Dft<B, INV>::pos(k / A); }
    template <int N2>
    __device__ __forceinline__ static void stage1(float2 *v)
    {
        float2 u[A];
#pragma unroll
        for (int n1 = 0; n1 < A; n1++)
            u[n1] = v[B * n1 + N2];
        Dft<A, INV>::run(u);
        v[B * 0 + N2] = u[0];
        v[B * 1 + N2] = twiddle_const<R, INV, N2 * 1>(u[1]);
        v[B * 2 + N2] = twiddle_const<R, INV, N2 * 2>(u[2]);
        v[B * 3 + N2] = twiddle_const<R, INV, N2 * 3>(u[3]);
    }
    template <int N2>
    __device__ __forceinline__ static void stage1_all(float2 *v)
    {
        if constexpr (N2 < B) {
            stage1<N2>(v);
            stage1_all<N2 + 1>(v);
        }
    }
    __device__ __forceinline__ static void run(float2 *v)
    {
        stage1_all<0>(v);
#pragma unroll
        for (int k1 = 0; k1 < A; k1++)
            Dft<B, INV>::run(v + B * k1);
    }
};

// Compile-time shape of one transform.
template <int LOG2N, int E>
struct FftShape {
    static constexpr int N = 1 << LOG2N;
    static constexpr int LOGE = ilog2c(E);
    static constexpr int NT = N / E;                       // threads per transform
    static constexpr int NPASS = (LOG2N + LOGE - 1) / LOGE; // Stockham passes
    static constexpr int LS = (LOG2N + 1) / 2;              // low twiddle bits
    static constexpr int TWLO = 1 << LS;
    static constexpr int TWHI = 1 << (LOG2N - LS);
    static constexpr int NPAD = N + (N >> LOGE);            // padded LDS length
};

template <int LOGE>
__device__ __forceinline__ int lidx(int i) { return i + (i >> LOGE); }

// The twiddle tables in device memory (read-only, forward sign): the
// two-level table of each transform size and the middle pass's row table.
template <int LOG2N>
__device__ constexpr twc::TwTable<LOG2N> kTwTable = twc::make_twtable<LOG2N>();
__device__ constexpr twc::TwRows kTwRowsTable = twc::make_twrows();

// Copy the two twiddle tables (forward sign) of a transform of 2^LOG2N into
// LDS.  The thread index passes through an empty asm so that the copy's index
// arithmetic is not shared with the caller's loop: otherwise the compiler
// keeps fewer of the caller's loop-invariant offsets in registers (measured:
// 7 more vector instructions per block in the 2^14-point fused inverse).
template <int LOG2N, int E>
__device__ __forceinline__ void fill_twiddles(float2 *twlo, float2 *twhi, int tid, int nthreads)
{
    using S = FftShape<LOG2N, E>;
    using TT = twc::TwTable<LOG2N>;
    static_assert(TT::LO == S::TWLO && TT::HI == S::TWHI, "table shape");
    const float2 *src = reinterpret_cast<const float2 *>(kTwTable<LOG2N>.v);
    asm volatile("" : "+v"(tid));
    for (int i = tid; i < S::TWLO; i += nthreads)
        twlo[i] = src[i];
    for (int i = tid; i < S::TWHI; i += nthreads)
        twhi[i] = src[S::TWLO + i];
}

// Row table of the inter-pass twiddles of a pass with NS = 32 and radix 32
// (the middle pass of the 2^14-point transform with 32 points per thread):
// row k holds W_1024^(k r) for r < 32 (forward sign), rows padded to 33
// entries.  A thread's k is fixed, so its 31 reads are one base address plus
// immediate offsets, and the 32 rows a wave reads start in distinct banks.
using twc::kTwRowElems;
using twc::kTwRowN;
using twc::kTwRowStride;
__device__ __forceinline__ void fill_twrows(float2 *tw, int tid, int nthreads)
{
    const float2 *src = reinterpret_cast<const float2 *>(kTwRowsTable.v);
    asm volatile("" : "+v"(tid));
    for (int i = tid; i < kTwRowElems; i += nthreads)
        tw[i] = src[i];
}

template <int LOG2N, int E, bool INV>
__device__ __forceinline__ float2 twiddle_lookup(const float2 *twlo, const float2 *twhi, int e)
{
    using S = FftShape<LOG2N, E>;
    float2 w = cmul(twhi[e >> S::LS], twlo[e & (S::TWLO - 1)]);
    return INV ? make_float2(w.x, -w.y) : w;
}

template <int LOG2N, int E, bool INV, int P, bool ROWS = false>
__device__ __forceinline__ void fft_pass(float2 (&v)[E], float2 *lds, const float2 *twlo, const float2 *twhi, int t,
                                         const float2 *twrows = nullptr)
{
    using S = FftShape<LOG2N, E>;
    constexpr int N = S::N, NT = S::NT, LOGE = S::LOGE, NPASS = S::NPASS;
    constexpr int LOGR = (P < NPASS - 1) ? LOGE : (LOG2N - (NPASS - 1) * LOGE);
    constexpr int R = 1 << LOGR;
    constexpr int LOGNS = P * LOGE;
    constexpr int NS = 1 << LOGNS;
    constexpr int Q = E / R;
    if constexpr (P > 0) {
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
            for (int r = 0; r < R; r++)
                v[q + r * Q] = lds[lidx<LOGE>(t + q * NT + r * (N / R))];
        if constexpr (P < NPASS - 1)
            __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        float2 u[R];
#pragma unroll
        for (int r = 0; r < R; r++)
            u[r] = v[q + r * Q];
        const int j = t + q * NT;
        const int k = j & (NS - 1);
        if constexpr (P > 0 && ROWS && NS == kTwRowN && R == kTwRowN) {
            // W_(NS R)^(k r) = W_1024^(k r) from row k of the table
            const float2 *row = twrows + k * kTwRowStride;
#pragma unroll
            for (int r = 1; r < R; r++) {
                const float2 w = row[r];
                u[r] = cmul(u[r], INV ? make_float2(w.x, -w.y) : w);
            }
        } else if constexpr (P > 0) {
            // W^r for r = c + 4b + 16a from three lookups (W, W^4, W^16) and
            // products of base-4 digits: at most 4 roundings per twiddle.
            const int e1 = k << (LOG2N - LOGNS - LOGR);
            const float2 w1 = twiddle_lookup<LOG2N, E, INV>(twlo, twhi, e1);
            const float2 w2 = cmul(w1, w1);
            const float2 w3 = cmul(w2, w1);
            float2 w4 = make_float2(1.f, 0.f), w8 = w4, w12 = w4, w16 = w4;
            if constexpr (R > 4) {
                w4 = twiddle_lookup<LOG2N, E, INV>(twlo, twhi, (4 * e1) & (S::N - 1));
                w8 = cmul(w4, w4);
                w12 = cmul(w8, w4);
            }
            if constexpr (R > 16)
                w16 = twiddle_lookup<LOG2N, E, INV>(twlo, twhi, (16 * e1) & (S::N - 1));
#pragma unroll
            for (int r = 1; r < R; r++) {
                const int c = r & 3, b = (r >> 2) & 3, a = r >> 4;
                float2 w = c == 1 ? w1 : c == 2 ? w2 : w3;
                if (c == 0)
                    w = b == 1 ? w4 : b == 2 ? w8 : w12;
                else if (b != 0)
                    w = cmul(w, b == 1 ? w4 : b == 2 ? w8 : w12);
                if (a != 0)
                    w = (c == 0 && b == 0) ? w16 : cmul(w, w16);
                u[r] = cmul(u[r], w);
            }
        }
        Dft<R, INV>::run(u);
        if constexpr (P == NPASS - 1) {
#pragma unroll
            for (int r = 0; r < R; r++)
                v[q + r * Q] = u[Dft<R, INV>::pos(r)];
        } else {
            const int base = ((j >> LOGNS) << (LOGNS + LOGR)) + k;
#pragma unroll
            for (int r = 0; r < R; r++)
                lds[lidx<LOGE>(base + r * NS)] = u[Dft<R, INV>::pos(r)];
        }
    }
    if constexpr (P < NPASS - 1) {
        __syncthreads();
        fft_pass<LOG2N, E, INV, P + 1, ROWS>(v, lds, twlo, twhi, t, twrows);
    }
}

// Twiddles W_N^(e0 + e1 m) (forward: exp(-2 pi i e/N); inverse: conjugate)
// for a thread's points m = 0 .. E-1, as used by the column passes of
// four-step transforms (k1 = t + m NT: e0 = n2 t, e1 = n2 NT).  Evaluated from
// at most four FP32 sincospi -- W^e0, W^e1, W^(4 e1), W^(16 e1); the exponents
// are reduced mod N (a power of two <= 2^28, so 2 e / N is exact in FP32) --
// and base-4 digit products of them: a few roundings per twiddle, against one
// FP64 sincospi per point.
template <int E, bool INV>
struct ColumnTwiddles {
    static_assert(E <= 64, "three base-4 digits");
    float2 a, s1, s2, s3, q1, q2, q3, h1, h2, h3;
    bool has_a;
    __device__ __forceinline__ static float2 w(long long e, long long N)
    {
        float s, c;
        sincospif(2.0f * (float)(e & (N - 1)) / (float)N, &s, &c);
        return make_float2(c, INV ? s : -s);
    }
    __device__ __forceinline__ ColumnTwiddles(long long e0, long long e1, long long N)
    {
        const float2 one = make_float2(1.f, 0.f);
        has_a = (e0 & (N - 1)) != 0;
        a = has_a ? w(e0, N) : one;
        s1 = w(e1, N);
        s2 = cmul(s1, s1);
        s3 = cmul(s2, s1);
        q1 = q2 = q3 = h1 = h2 = h3 = one;
        if constexpr (E > 4) {
            q1 = w(4 * e1, N);
            q2 = cmul(q1, q1);
            q3 = cmul(q2, q1);
        }
        if constexpr (E > 16) {
            h1 = w(16 * e1, N);
            h2 = cmul(h1, h1);
            h3 = cmul(h2, h1);
        }
    }
    // W^(e0 + e1 m); m is a compile-time constant after unrolling
    __device__ __forceinline__ float2 operator()(int m) const
    {
        const int c = m & 3, d = (m >> 2) & 3, f = m >> 4;
        float2 r = a;
        bool unit = !has_a;
        auto mul = [&](float2 x) {
            r = unit ? x : cmul(r, x);
            unit = false;
        };
        if (c)
            mul(c == 1 ? s1 : c == 2 ? s2 : s3);
        if (d)
            mul(d == 1 ? q1 : d == 2 ? q2 : q3);
        if (f)
            mul(f == 1 ? h1 : f == 2 ? h2 : h3);
        return r;
    }
};

// Full transform.  Every thread of the workgroup must call it (it contains
// barriers).  lds must hold FftShape::NPAD float2 for this transform.  On
// entry v[m] = x[t + m*NT]; on exit v[m] = X[t + m*NT] (unnormalised).
template <int LOG2N, int E, bool INV, bool ROWS = false>
__device__ __forceinline__ void fft_block(float2 (&v)[E], float2 *lds, const float2 *twlo, const float2 *twhi, int t,
                                          const float2 *twrows = nullptr)
{
    using S = FftShape<LOG2N, E>;
    if constexpr (S::NPASS == 1) {
        Dft<S::N, INV>::run(v);
        float2 o[E];
#pragma unroll
        for (int k = 0; k < E; k++)
            o[k] = v[Dft<S::N, INV>::pos(k)];
#pragma unroll
        for (int k = 0; k < E; k++)
            v[k] = o[k];
    } else {
        __syncthreads(); // LDS may still be read by a previous transform
        fft_pass<LOG2N, E, INV, 0, ROWS>(v, lds, twlo, twhi, t, twrows);
    }
}

// Elements per thread for a transform size: 16 (radix-16 passes), 32 for
// 2^14 points (512 threads, radix 32-32-16: one LDS exchange fewer and a
// 256-register budget per thread).
__host__ __device__ constexpr int fft_e(int log2n) { return log2n >= 14 ? 32 : log2n >= 4 ? 16 : (1 << log2n); }
// Transforms per workgroup (fill ~256 threads for small sizes).
__host__ __device__ constexpr int fft_tpb(int log2n)
{
    return ((1 << log2n) / fft_e(log2n)) >= 256 ? 1 : 256 / ((1 << log2n) / fft_e(log2n));
}

// Threads of a workgroup running fft_tpb() transforms.
__host__ __device__ constexpr int fft_threads(int log2n) { return ((1 << log2n) / fft_e(log2n)) * fft_tpb(log2n); }

} // namespace msosm
