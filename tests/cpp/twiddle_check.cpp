// Checks the compile-time twiddle tables (msosm_gpu_amd/csrc/twiddle_tables.h)
// against the C library's long-double cosl/sinl: every entry must be the
// float nearest to exp(-2 pi i e / N); a one-ulp difference is accepted only
// where the exact value lies within 2^-30 ulp of a rounding tie.  Prints one
// line per table and exits non-zero on any other difference.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>

#include "twiddle_tables.h"

using namespace msosm::twc;

static long double kPiL = 3.141592653589793238462643383279502884L;

// 0 = exact match, 1 = accepted near-tie, 2 = failure
static int check_value(float got, long double exact)
{
    const float want = (float)exact;
    if (std::memcmp(&got, &want, sizeof(float)) == 0 || (got == 0.0f && want == 0.0f))
        return 0;
    const float lo = std::fmin(got, want), hi = std::fmax(got, want);
    if (std::nextafterf(lo, 1e30f) != hi)
        return 2;
    const long double mid = ((long double)lo + (long double)hi) / 2;
    const long double ulp = (long double)hi - (long double)lo;
    return std::fabs((double)((exact - mid) / ulp)) < std::ldexp(1.0, -30) ? 1 : 2;
}

static int check_pairs(const char *name, const float *v, int count, long long n, long long (*expo)(int, const void *),
                       const void *ctx)
{
    int ties = 0, bad = 0;
    for (int i = 0; i < count; i++) {
        const long long e = expo(i, ctx) & (n - 1);
        // reference: quadrant by exact integer reduction (so multiples of
        // pi/2 give exact zeros and ones), long-double cosl/sinl inside it
        const long long q = 4 * e / n, r = 4 * e - q * n;
        const long double a = kPiL / 2 * (long double)r / (long double)n;
        const long double ca = cosl(a), sa = sinl(a);
        const long double c = q == 0 ? ca : q == 1 ? -sa : q == 2 ? -ca : sa;
        const long double s = q == 0 ? sa : q == 1 ? ca : q == 2 ? -sa : -ca;
        const int rc = check_value(v[2 * i], c);
        const int rs = check_value(v[2 * i + 1], -s);
        ties += (rc == 1) + (rs == 1);
        if (rc == 2 || rs == 2) {
            if (bad < 5)
                std::printf("  %s entry %d (e=%lld): got (%.9g, %.9g)\n", name, i, e, v[2 * i], v[2 * i + 1]);
            bad++;
        }
    }
    std::printf("%s: %d entries, %d near-tie, %d wrong\n", name, count, ties, bad);
    return bad;
}

struct TableCtx {
    int lo;
};

static long long two_level(int i, const void *ctx)
{
    const int lo = static_cast<const TableCtx *>(ctx)->lo;
    return i < lo ? i : (long long)(i - lo) * lo;
}

static long long rows(int i, const void *)
{
    const int k = i / kTwRowStride, r = i - k * kTwRowStride;
    return (long long)k * r;
}

template <int L>
static int check_table()
{
    static constexpr TwTable<L> t = make_twtable<L>(); // evaluated at compile time
    char name[32];
    std::snprintf(name, sizeof name, "table 2^%d", L);
    const TableCtx ctx{TwTable<L>::LO};
    return check_pairs(name, t.v, TwTable<L>::LO + TwTable<L>::HI, 1ll << L, two_level, &ctx);
}

template <int... L>
static int check_all(std::integer_sequence<int, L...>)
{
    return (check_table<L>() + ...);
}

int main()
{
    int bad = check_all(std::integer_sequence<int, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14>{});
    static constexpr TwRows r = make_twrows();
    bad += check_pairs("row table", r.v, kTwRowElems, (long long)kTwRowN * kTwRowN, rows, nullptr);
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
