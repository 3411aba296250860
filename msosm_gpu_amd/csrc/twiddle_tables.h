// twiddle_tables.h -- the FFT twiddle tables as compile-time constants.
//
// The LDS twiddle tables of the Stockham passes (fft_device.h) are copied at
// the start of each workgroup from these read-only arrays instead of being
// evaluated there with FP64 sincospi (several hundred vector instructions per
// wave, a noticeable share of a workgroup that transforms only a few blocks).
// Values are W_N^e = exp(-2 pi i e / N) stored as (cos, -sin) pairs, rounded
// to float from a double-precision evaluation at compile time: the angle is
// reduced to an octant with exact integer arithmetic and the sine and cosine
// of the remaining [0, pi/4] come from their Taylor series.  Plain C++17 (no
// HIP), so a host program can check every entry against the C library.
#pragma once

namespace msosm {
namespace twc {

constexpr double kQuarterPi = 0.785398163397448309615660845819875721;

// Taylor series on [0, pi/4]; the first omitted term is below 1e-30 there
constexpr double sin_oct(double x)
{
    double term = x, sum = x;
    for (int k = 1; k <= 13; k++) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cos_oct(double x)
{
    double term = 1.0, sum = 1.0;
    for (int k = 1; k <= 13; k++) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}

struct CosSin {
    double c, s;
};

// cos and sin of 2 pi e / n for a power of two n: with 8 e = oct n + rem the
// angle is (pi/4) (oct + rem / n); odd octants use the complement
// (pi/4) (n - rem) / n so that both series arguments stay in [0, pi/4]
constexpr CosSin cis(long long e, long long n)
{
    e &= n - 1;
    const long long q = 8 * e;
    const int oct = (int)(q / n);
    const long long rem = q - (long long)oct * n;
    const double x = kQuarterPi * (double)rem / (double)n;
    const double y = kQuarterPi * (double)(n - rem) / (double)n;
    switch (oct) {
    case 0: return {cos_oct(x), sin_oct(x)};
    case 1: return {sin_oct(y), cos_oct(y)};
    case 2: return {-sin_oct(x), cos_oct(x)};
    case 3: return {-cos_oct(y), sin_oct(y)};
    case 4: return {-cos_oct(x), -sin_oct(x)};
    case 5: return {-sin_oct(y), -cos_oct(y)};
    case 6: return {sin_oct(x), -cos_oct(x)};
    default: return {cos_oct(y), -sin_oct(y)};
    }
}

// Two-level table of a 2^LOG2N-point transform: lo[i] = W_N^i for i < 2^LS and
// hi[k] = W_N^(k 2^LS) for k < 2^(LOG2N - LS), LS = ceil(LOG2N / 2), stored
// lo then hi as interleaved (cos, -sin) floats.
template <int LOG2N>
struct TwTable {
    static constexpr int LS = (LOG2N + 1) / 2;
    static constexpr int LO = 1 << LS;
    static constexpr int HI = 1 << (LOG2N - LS);
    alignas(8) float v[2 * (LO + HI)];
};

template <int LOG2N>
constexpr TwTable<LOG2N> make_twtable()
{
    TwTable<LOG2N> t{};
    constexpr long long N = 1ll << LOG2N;
    for (int i = 0; i < TwTable<LOG2N>::LO + TwTable<LOG2N>::HI; i++) {
        const long long e = i < TwTable<LOG2N>::LO ? i : (long long)(i - TwTable<LOG2N>::LO) * TwTable<LOG2N>::LO;
        const CosSin w = cis(e, N);
        t.v[2 * i] = (float)w.c;
        t.v[2 * i + 1] = (float)(-w.s);
    }
    return t;
}

// Row table of the middle pass of the 2^14-point transform with 32 points per
// thread: row k holds W_1024^(k r) for r < 32, rows padded to 33 entries.
constexpr int kTwRowN = 32;      // rows (NS) and columns (R)
constexpr int kTwRowStride = 33; // padded row length
constexpr int kTwRowElems = kTwRowN * kTwRowStride;

struct TwRows {
    alignas(8) float v[2 * kTwRowElems];
};

constexpr TwRows make_twrows()
{
    TwRows t{};
    for (int i = 0; i < kTwRowElems; i++) {
        const int k = i / kTwRowStride, r = i - k * kTwRowStride;
        const CosSin w = cis((long long)k * r, (long long)kTwRowN * kTwRowN);
        t.v[2 * i] = (float)w.c;
        t.v[2 * i + 1] = (float)(-w.s);
    }
    return t;
}

} // namespace twc
} // namespace msosm
