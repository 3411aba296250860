// fold.hip -- pulsar folding behind the C ABI (folding/fold_gpu.cpp).
//
// The reference folds on the CPU, one sample at a time (fold_gpu.cpp:48-173).
// Here the host replays only its per-period bookkeeping with the same float
// state (the fractional drift that skips one sample whenever it passes 1.0,
// the discarded first ceil(Nd/P) periods, the staging of a period that
// straddles two calls) and hands the GPU the start offset of every period to
// add.  The GPU kernel gives each raw bin one thread that adds the periods in
// order, so every bin receives the same float additions in the same order as
// the sequential fold: folded_data_raw is reproduced bit for bit.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "msosm_c.h"
#include "trace.h"

namespace {

#define ferr msosm::set_error

#define FHIP(expr)                                                                                               \
    do {                                                                                                         \
        hipError_t e__ = (expr);                                                                                 \
        if (e__ != hipSuccess)                                                                                   \
            return ferr(MSOSM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e__));                          \
    } while (0)

// tmp[dst + i] = in[src + i], i < len
__global__ void k_fold_stage(float *__restrict__ tmp, const float *__restrict__ in, unsigned long dst,
                             unsigned long src, unsigned long len)
{
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < len;
         i += (unsigned long)gridDim.x * blockDim.x)
        tmp[dst + i] = in[src + i];
}

// raw[i] += tmp[i] (if add_tmp), then raw[i] += in[start_k + i] for every
// listed period k in order -- the reference's summation order per bin.
__global__ void k_fold_add(float *__restrict__ raw, const float *__restrict__ tmp, const float *__restrict__ in,
                           const unsigned long *__restrict__ starts, int nper, unsigned long P, int add_tmp)
{
#pragma clang fp contract(off)
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < P;
         i += (unsigned long)gridDim.x * blockDim.x) {
        float acc = raw[i];
        if (add_tmp)
            acc = acc + tmp[i];
        for (int k = 0; k < nper; k++)
            acc = acc + in[starts[k] + i];
        raw[i] = acc;
    }
}

// folded[i] = (sum_j raw[i*bpt + j]) / bpt, j summed in order (fold_gpu.cpp:185-193)
__global__ void k_fold_bins(float *__restrict__ folded, const float *__restrict__ raw, unsigned long time_bin,
                            int bpt)
{
#pragma clang fp contract(off)
    for (unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x; i < time_bin;
         i += (unsigned long)gridDim.x * blockDim.x) {
        float sum = 0.0f;
        for (int j = 0; j < bpt; j++)
            sum = sum + raw[i * (unsigned long)bpt + j];
        folded[i] = sum / (float)bpt; // IEEE division
    }
}

// Phase fold: sample i of DM trial dm (global index n0 + i) lands in bin
// floor(frac((n0 + i) / P) * nbins), the phase evaluated in FP64.  Each
// workgroup builds a histogram in LDS (float atomics) and adds it to the
// global profile once.  value = sqrt(|a|^2+|b|^2) (the reference's
// intensity, gpu_kernel.cu:352) or |a|^2+|b|^2 (power).
//
// Each workgroup folds one contiguous chunk of its trial's samples, split
// into one contiguous piece per wave, so it touches only the bins its chunk
// spans and adds only those (non-zero) bins to the global profile.  A wave
// reads kFoldUnroll groups of 64*V consecutive samples per step (V = 2: one
// 16-byte load of two samples per lane, when the rows are 16-byte aligned).
//
// The bin of a sample only changes at bin edges, and within one period the
// FP64 phase, hence the bin, is non-decreasing in the sample index (every
// step of the phase evaluation is monotonic; ph - floor(ph) is exact).  So a
// group whose first and last samples fall into the same bin of the same
// period lies wholly in that bin: per step, lanes 2u and 2u+1 evaluate the
// phase of group u's first and last sample, and only groups that straddle a
// bin edge evaluate it per sample.  A wave adds the samples of whole-bin
// groups to a register while the bin stays the same and sums the registers
// with lane shuffles when it changes; a straddling group is split by bin
// with ballots (a few partial sums), and lanes left after that add to the
// LDS histogram one by one.
__device__ __forceinline__ float wave_sum(float x)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
        x += __shfl_xor(x, m);
    return x;
}

constexpr int kFoldThreads = 256;
constexpr int kFoldUnroll = 8;
constexpr int kFoldPeel = 3; // partial sums per straddling group before per-lane atomics

__device__ __forceinline__ float pw(float2 v) { return v.x * v.x + v.y * v.y; }

// V consecutive complex64 samples of one lane: one 8- or 16-byte load
template <int V>
struct Lanes;
template <>
struct Lanes<1> {
    float2 v;
    __device__ __forceinline__ static Lanes load(const float2 *p) { return {*p}; }
    __device__ __forceinline__ float power(int) const { return pw(v); }
};
template <>
struct Lanes<2> {
    float4 v;
    __device__ __forceinline__ static Lanes load(const float2 *p) { return {*reinterpret_cast<const float4 *>(p)}; }
    __device__ __forceinline__ float power(int s) const
    {
        return s ? pw(make_float2(v.z, v.w)) : pw(make_float2(v.x, v.y));
    }
};

// One polarisation: registers for 8 waves per SIMD (no spills); two: as needed.
template <int V, bool TWO>
__global__ __launch_bounds__(kFoldThreads, TWO ? 1 : 8) void k_fold_phase(const float2 *__restrict__ a,
                                                             const float2 *__restrict__ b, long long n,
                                                             long long stride, long long chunk, long long n0,
                                                             double inv_period, int nbins, int power,
                                                             float *__restrict__ prof, float *__restrict__ hits)
{
    static_assert(V == 1 || V == 2, "one or two samples per lane and load");
    static_assert(2 * kFoldUnroll <= 64, "two lanes per group evaluate its end phases");
    extern __shared__ float hist[]; // nbins values (+ nbins hit counts for trial 0)
    const int dm = blockIdx.y;
    const bool count_hits = (hits != nullptr) && dm == 0;
    const int nh = count_hits ? 2 * nbins : nbins;
    for (int i = threadIdx.x; i < nh; i += blockDim.x)
        hist[i] = 0.0f;
    __syncthreads();
    // this wave's parking area for the values of straddling groups
    float *stage = hist + nbins * (hits != nullptr ? 2 : 1) + (threadIdx.x >> 6) * (kFoldUnroll * 64 * V);
    const float2 *ad = a + (long long)dm * stride;
    const float2 *bd = b ? b + (long long)dm * stride : nullptr;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int waves = kFoldThreads / 64;
    constexpr int span = 64 * V; // samples per wave and load
    constexpr int group_step = span * kFoldUnroll;
    // chunk is a multiple of waves * group_step: equal contiguous pieces
    const long long piece = chunk / waves;
    const long long w0 = (long long)blockIdx.x * chunk + wave * piece;
    const long long w1 = min(w0 + piece, n);
    // bin of global sample index i; *whole = floor of the phase (period number)
    auto phase_bin = [&](long long i, double *whole) {
        double ph = (double)i * inv_period;
        const double f = floor(ph);
        ph -= f;
        if (whole)
            *whole = f;
        const int bin = (int)(ph * (double)nbins);
        return bin < nbins ? bin : nbins - 1;
    };
    float acc = 0.f;  // per-lane sum of the samples of bin cur
    float ucnt = 0.f; // samples added to acc (wave-uniform)
    int cur = -1;     // wave-uniform bin being accumulated in registers
    auto flush = [&]() {
        if (cur >= 0) {
            const float sum = wave_sum(acc);
            if (lane == 0) {
                atomicAdd(&hist[cur], sum);
                if (count_hits)
                    atomicAdd(&hist[nbins + cur], ucnt);
            }
        }
        acc = 0.f;
        ucnt = 0.f;
        cur = -1;
    };
    // fold one step of samples base .. base+min(rem, group_step) whose values
    // are in p (uniform base and rem)
    auto process = [&](float (&p)[kFoldUnroll][V], long long base, long long rem) {
        if (!power) {
#pragma unroll
            for (int u = 0; u < kFoldUnroll; u++)
#pragma unroll
                for (int s = 0; s < V; s++)
                    p[u][s] = sqrtf(p[u][s]);
        }
        // end phases of the groups: lane 2u -> first sample of group u, lane
        // 2u+1 -> its last sample (clamped to the piece); other lanes unused
        int ebin, eper;
        {
            const int gu = lane >> 1;
            long long idx = (lane & 1) ? min((long long)(gu + 1) * span, rem) - 1 : (long long)gu * span;
            idx = (lane < 2 * kFoldUnroll && idx >= 0 && idx < rem) ? idx : 0;
            double whole;
            ebin = phase_bin(n0 + base + idx, &whole);
            eper = (int)(long long)whole; // period number, low 32 bits
        }
        unsigned smask = 0; // groups that straddle a bin edge (uniform)
#pragma unroll
        for (int u = 0; u < kFoldUnroll; u++) {
            if (u * span >= rem) // uniform: nothing left in this group or after
                break;
            const int bf = __builtin_amdgcn_readlane(ebin, 2 * u), bl = __builtin_amdgcn_readlane(ebin, 2 * u + 1);
            const int pf = __builtin_amdgcn_readlane(eper, 2 * u), pl = __builtin_amdgcn_readlane(eper, 2 * u + 1);
            if (bf == bl && (pf == pl || nbins == 1)) { // uniform: the whole group is in bin bf
                if (bf != cur) {
                    flush();
                    cur = bf;
                }
#pragma unroll
                for (int s = 0; s < V; s++)
                    acc += p[u][s];
                ucnt += (float)min((long long)span, rem - u * span);
                continue;
            }
            // straddles a bin edge: park the values for the per-sample pass
#pragma unroll
            for (int s = 0; s < V; s++)
                stage[u * span + lane * V + s] = p[u][s];
            smask |= 1u << u;
        }
        // per-sample pass over the straddling groups (a runtime loop: only the
        // values of one group are live)
        while (smask) { // uniform
            const int u = __builtin_ctz(smask);
            smask &= smask - 1;
#pragma unroll
            for (int s = 0; s < V; s++) {
                const int off = u * span + lane * V + s;
                bool left = off < rem;
                const int bin = left ? phase_bin(n0 + base + off, nullptr) : 0;
                const float v = stage[off]; // written by this lane
                for (int it = 0; it < kFoldPeel; it++) {
                    const unsigned long long todo = __ballot(left);
                    if (todo == 0) // uniform
                        break;
                    const int bsel = __shfl(bin, __ffsll((long long)todo) - 1);
                    const bool mine = left && bin == bsel;
                    const float sum = wave_sum(mine ? v : 0.f);
                    const unsigned long long mm = __ballot(mine);
                    if (lane == 0) {
                        atomicAdd(&hist[bsel], sum);
                        if (count_hits)
                            atomicAdd(&hist[nbins + bsel], (float)__popcll(mm));
                    }
                    left = left && !mine;
                }
                if (left) {
                    atomicAdd(&hist[bin], v);
                    if (count_hits)
                        atomicAdd(&hist[nbins + bin], 1.0f);
                }
            }
        }
    };
    // whole steps: every load of a step issued before any is used
    const long long nfull = w1 > w0 ? (w1 - w0) / group_step : 0;
    for (long long k = 0; k < nfull; k++) { // wave-uniform bound
        const long long base = w0 + k * group_step;
        Lanes<V> ra[kFoldUnroll], rb[TWO ? kFoldUnroll : 1];
#pragma unroll
        for (int u = 0; u < kFoldUnroll; u++) {
            ra[u] = Lanes<V>::load(ad + base + u * span + lane * V);
            if constexpr (TWO)
                rb[u] = Lanes<V>::load(bd + base + u * span + lane * V);
        }
        float p[kFoldUnroll][V];
#pragma unroll
        for (int u = 0; u < kFoldUnroll; u++)
#pragma unroll
            for (int s = 0; s < V; s++) {
                p[u][s] = ra[u].power(s);
                if constexpr (TWO)
                    p[u][s] += rb[u].power(s);
            }
        process(p, base, group_step);
    }
    // the last, partial step of a row: one sample at a time
    const long long tb = w0 + nfull * group_step;
    if (tb < w1) {
        const long long rem = w1 - tb;
        float p[kFoldUnroll][V];
#pragma unroll
        for (int u = 0; u < kFoldUnroll; u++)
#pragma unroll
            for (int s = 0; s < V; s++) {
                const int off = u * span + lane * V + s;
                p[u][s] = 0.f;
                if (off < rem) {
                    p[u][s] = pw(ad[tb + off]);
                    if constexpr (TWO)
                        p[u][s] += pw(bd[tb + off]);
                }
            }
        process(p, tb, rem);
    }
    flush();
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        const float h = hist[i];
        if (h != 0.0f)
            atomicAdd(&prof[(long long)dm * nbins + i], h);
        if (count_hits) {
            const float c = hist[nbins + i];
            if (c != 0.0f)
                atomicAdd(&hits[i], c);
        }
    }
}

int grid1(unsigned long n)
{
    unsigned long g = (n + 255) / 256;
    if (g < 1)
        g = 1;
    if (g > 65536)
        g = 65536;
    return (int)g;
}

} // namespace

struct msosm_fold {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // reference state (fold_gpu.h:28-47), same types
    unsigned long size = 0;
    float period_samples_float = 0;
    unsigned long period_samples = 0;
    float diff_samples = 0;
    unsigned long time_bin = 0;
    bool ready = false;
    int discard_count = 0;
    int current_discard = 0;
    float current_diff = 0.0f;
    unsigned long current_index = 0;
    // device buffers
    float *intensity = nullptr; // [size]
    float *raw = nullptr;       // [period_samples]
    float *tmp = nullptr;       // [period_samples]
    float *folded = nullptr;    // [time_bin]
    // period-start upload: two pinned host buffers alternating, guarded by events
    unsigned long *starts_d = nullptr;
    unsigned long *starts_h[2] = {nullptr, nullptr};
    hipEvent_t starts_ev[2] = {nullptr, nullptr};
    int starts_cur = 0;
    size_t max_periods = 0;
};

namespace {

void fold_free(msosm_fold *f)
{
    if (!f)
        return;
    if (f->stream)
        (void)hipStreamSynchronize(f->stream);
    void *bufs[] = {f->intensity, f->raw, f->tmp, f->folded, f->starts_d};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    for (int i = 0; i < 2; i++) {
        if (f->starts_h[i])
            (void)hipHostFree(f->starts_h[i]);
        if (f->starts_ev[i])
            (void)hipEventDestroy(f->starts_ev[i]);
    }
    if (f->own_stream && f->stream)
        (void)hipStreamDestroy(f->stream);
    delete f;
}

void period_done(msosm_fold *f)
{
    if (!f->ready) { // fold_gpu.cpp:111-119
        f->current_discard += 1;
        if (f->current_discard == f->discard_count) {
            f->ready = true;
            f->current_discard = 0;
        }
    }
}

} // namespace

extern "C" {

int msosm_fold_create(float period, float fs, unsigned long size, unsigned long time_bin, void *hip_stream,
                      msosm_fold **out)
{
    if (!out)
        return ferr(MSOSM_ERR_ARG, "fold_create: NULL out");
    *out = nullptr;
    msosm_fold *f = new msosm_fold();
    f->size = size;
    f->period_samples_float = period * fs;                                        // fold_gpu.cpp:6
    f->period_samples = (unsigned long)f->period_samples_float;                   // :7
    f->diff_samples = f->period_samples_float - (float)f->period_samples;         // :8
    f->time_bin = (time_bin == (unsigned long)-1) ? f->period_samples : time_bin; // :9-16
    if (size == 0 || f->period_samples == 0 || f->time_bin == 0 || f->time_bin > f->period_samples) {
        int rc = ferr(MSOSM_ERR_ARG, "fold_create: need size > 0, period*fs >= 1 and 0 < time_bin <= period "
                                     "samples (size=%lu period_samples=%lu time_bin=%lu)",
                      size, f->period_samples, f->time_bin);
        delete f;
        return rc;
    }
    if (hip_stream) {
        f->stream = (hipStream_t)hip_stream;
    } else {
        if (hipStreamCreate(&f->stream) != hipSuccess) {
            delete f;
            return ferr(MSOSM_ERR_HIP, "fold_create: hipStreamCreate failed");
        }
        f->own_stream = true;
    }
    f->max_periods = size / f->period_samples + 2;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc((void **)&f->intensity, sizeof(float) * size);
    if (e == hipSuccess) e = hipMalloc((void **)&f->raw, sizeof(float) * f->period_samples);
    if (e == hipSuccess) e = hipMalloc((void **)&f->tmp, sizeof(float) * f->period_samples);
    if (e == hipSuccess) e = hipMalloc((void **)&f->folded, sizeof(float) * f->time_bin);
    if (e == hipSuccess) e = hipMalloc((void **)&f->starts_d, sizeof(unsigned long) * f->max_periods * 2);
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        e = hipHostMalloc((void **)&f->starts_h[i], sizeof(unsigned long) * f->max_periods, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&f->starts_ev[i], hipEventDisableTiming);
    }
    // The reference leaves these uninitialised (fold_gpu.cpp:17-20); zeroed.
    if (e == hipSuccess) e = hipMemsetAsync(f->intensity, 0, sizeof(float) * size, f->stream);
    if (e == hipSuccess) e = hipMemsetAsync(f->raw, 0, sizeof(float) * f->period_samples, f->stream);
    if (e == hipSuccess) e = hipMemsetAsync(f->tmp, 0, sizeof(float) * f->period_samples, f->stream);
    if (e == hipSuccess) e = hipMemsetAsync(f->folded, 0, sizeof(float) * f->time_bin, f->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
    if (e != hipSuccess) {
        fold_free(f);
        return ferr(MSOSM_ERR_ALLOC, "fold_create: %s", hipGetErrorString(e));
    }
    *out = f;
    return MSOSM_OK;
}

int msosm_fold_destroy(msosm_fold *f)
{
    fold_free(f);
    return MSOSM_OK;
}

// fold_gpu.cpp:29-33.  A count of 0 would leave the reference never "ready"
// (its check at :114 never matches); here it means ready from the start.
int msosm_fold_discard_samples(msosm_fold *f, unsigned long nd)
{
    if (!f)
        return ferr(MSOSM_ERR_ARG, "fold_discard_samples: NULL fold");
    f->discard_count = (int)std::ceil((float)nd / f->period_samples_float);
    f->current_discard = 0;
    f->ready = (f->discard_count == 0);
    return MSOSM_OK;
}

int msosm_fold_intensity(msosm_fold *f, const void *pol1, const void *pol2)
{
    if (!f || !pol1)
        return ferr(MSOSM_ERR_ARG, "fold_intensity: NULL argument");
    msosm::TraceRange range("msosm_fold_intensity");
    FHIP(msosm::launch_intensity((const float2 *)pol1, (const float2 *)pol2, f->intensity, f->size, f->stream));
    return MSOSM_OK;
}

int msosm_fold_intensity_ptr(msosm_fold *f, void **p)
{
    if (!f || !p)
        return ferr(MSOSM_ERR_ARG, "fold_intensity_ptr: NULL argument");
    *p = f->intensity;
    return MSOSM_OK;
}

int msosm_fold_load_intensity(msosm_fold *f, const float *src, int where)
{
    if (!f || !src || (where != 0 && where != 1))
        return ferr(MSOSM_ERR_ARG, "fold_load_intensity: bad argument");
    FHIP(hipMemcpyAsync(f->intensity, src, sizeof(float) * f->size,
                        where == 1 ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, f->stream));
    if (where == 0)
        FHIP(hipStreamSynchronize(f->stream));
    return MSOSM_OK;
}

// fold_gpu.cpp:48-173 -- bookkeeping on the host, additions on the GPU.
int msosm_fold_fold_data(msosm_fold *f)
{
    if (!f)
        return ferr(MSOSM_ERR_ARG, "fold_fold_data: NULL fold");
    msosm::TraceRange range("msosm_fold_fold_data");
    const unsigned long P = f->period_samples;
    const unsigned long NEG1 = (unsigned long)-1;
    if (f->discard_count == 0 && !f->ready)
        f->ready = true;
    unsigned long current_size = f->size;
    unsigned long ci = 0;
    // operations collected in reference order
    bool stage_head = false, add_tmp = false, stage_tail = false;
    unsigned long head_dst = 0, head_src = 0, head_len = 0, tail_src = 0, tail_len = 0;
    int k = f->starts_cur;
    FHIP(hipEventSynchronize(f->starts_ev[k])); // host buffer k free again
    unsigned long *starts = f->starts_h[k];
    int nper = 0;

    bool done = false;
    if (f->current_diff > 1.0f) { // :54-61
        f->current_diff -= 1.0f;
        ci += 1;
        current_size -= 1;
        if (current_size == 0)
            done = true;
    }
    if (!done && f->current_index != 0) { // :62-121
        const unsigned long ci0 = f->current_index;
        if (current_size < P - ci0) {
            if (f->ready) {
                stage_head = true;
                head_dst = ci0;
                head_src = ci;
                head_len = current_size;
            }
            f->current_index += current_size;
            done = true;
        } else {
            if (f->ready) {
                stage_head = true;
                head_dst = ci0;
                head_src = ci;
                head_len = P - ci0;
                add_tmp = true;
            }
            current_size -= P - ci0;
            ci += P - ci0;
            f->current_index = 0;
            f->current_diff += f->diff_samples;
            bool ret = false;
            if (f->current_diff > 1.0f) {
                f->current_diff -= 1.0f;
                ci += 1;
                current_size -= 1;
                if (current_size == NEG1) {
                    f->current_diff += 1.0f;
                    ret = true;
                }
            }
            if (ret)
                done = true;
            else
                period_done(f);
        }
    }
    if (!done) {
        while (current_size >= P) { // :124-152
            if (f->ready)
                starts[nper++] = ci;
            current_size -= P;
            ci += P;
            f->current_diff += f->diff_samples;
            bool ret = false;
            if (f->current_diff > 1.0f) {
                f->current_diff -= 1.0f;
                ci += 1;
                current_size -= 1;
                // the reference wraps current_size here (period ending exactly
                // at the buffer end with a skip due) and reads past the buffer;
                // defined like :105-109: defer the skip to the next call.
                if (current_size == NEG1) {
                    f->current_diff += 1.0f;
                    ret = true;
                }
            }
            if (ret) {
                done = true;
                break;
            }
            period_done(f);
        }
    }
    if (!done && current_size > 0) { // :160-172
        if (f->ready) {
            stage_tail = true;
            tail_src = ci;
            tail_len = current_size;
        }
        f->current_index = current_size;
    }

    hipStream_t s = f->stream;
    if (stage_head && head_len > 0) {
        hipLaunchKernelGGL(k_fold_stage, dim3(grid1(head_len)), dim3(256), 0, s, f->tmp, f->intensity, head_dst,
                           head_src, head_len);
        FHIP(hipGetLastError());
    }
    if (add_tmp || nper > 0) {
        unsigned long *sd = f->starts_d + (size_t)k * f->max_periods;
        if (nper > 0)
            FHIP(hipMemcpyAsync(sd, starts, sizeof(unsigned long) * nper, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_fold_add, dim3(grid1(P)), dim3(256), 0, s, f->raw, f->tmp, f->intensity, sd, nper, P,
                           add_tmp ? 1 : 0);
        FHIP(hipGetLastError());
    }
    FHIP(hipEventRecord(f->starts_ev[k], s));
    f->starts_cur ^= 1;
    if (stage_tail) {
        hipLaunchKernelGGL(k_fold_stage, dim3(grid1(tail_len)), dim3(256), 0, s, f->tmp, f->intensity, 0ul,
                           tail_src, tail_len);
        FHIP(hipGetLastError());
    }
    return MSOSM_OK;
}

int msosm_fold_bins(msosm_fold *f)
{
    if (!f)
        return ferr(MSOSM_ERR_ARG, "fold_bins: NULL fold");
    if (f->time_bin == f->period_samples) { // :182-183, folded_data aliases raw
        FHIP(hipMemcpyAsync(f->folded, f->raw, sizeof(float) * f->period_samples, hipMemcpyDeviceToDevice,
                            f->stream));
        return MSOSM_OK;
    }
    int bpt = (int)(f->period_samples / f->time_bin);
    hipLaunchKernelGGL(k_fold_bins, dim3(grid1(f->time_bin)), dim3(256), 0, f->stream, f->folded, f->raw,
                       f->time_bin, bpt);
    FHIP(hipGetLastError());
    return MSOSM_OK;
}

int msosm_fold_get(msosm_fold *f, float *folded, float *raw, float *intensity)
{
    if (!f)
        return ferr(MSOSM_ERR_ARG, "fold_get: NULL fold");
    if (folded)
        FHIP(hipMemcpyAsync(folded, f->folded, sizeof(float) * f->time_bin, hipMemcpyDeviceToHost, f->stream));
    if (raw)
        FHIP(hipMemcpyAsync(raw, f->raw, sizeof(float) * f->period_samples, hipMemcpyDeviceToHost, f->stream));
    if (intensity)
        FHIP(hipMemcpyAsync(intensity, f->intensity, sizeof(float) * f->size, hipMemcpyDeviceToHost, f->stream));
    FHIP(hipStreamSynchronize(f->stream));
    return MSOSM_OK;
}

int msosm_fold_info(msosm_fold *f, unsigned long *period_samples, unsigned long *time_bin, int *ready)
{
    if (!f)
        return ferr(MSOSM_ERR_ARG, "fold_info: NULL fold");
    if (period_samples)
        *period_samples = f->period_samples;
    if (time_bin)
        *time_bin = f->time_bin;
    if (ready)
        *ready = f->ready ? 1 : 0;
    return MSOSM_OK;
}

int msosm_fold_phase(const void *a, const void *b, size_t n, int ndm, size_t stride, long long n0,
                     double period_samples, int nbins, int power, float *prof, float *hits, void *stream)
{
    if (!a || !prof || n == 0 || ndm < 1 || nbins < 1 || nbins > 8192 || !(period_samples > 0) ||
        (ndm > 1 && stride < n))
        return ferr(MSOSM_ERR_ARG, "fold_phase: bad argument (need n > 0, 1 <= nbins <= 8192, period > 0, "
                                   "stride >= n)");
    if (ndm > 65535)
        return ferr(MSOSM_ERR_ARG, "fold_phase: at most 65535 DM trials per launch");
    msosm::TraceRange range("msosm_fold_phase");
    // two samples per lane and load when every row starts 16-byte aligned
    const bool wide = ((uintptr_t)a % 16 == 0) && (!b || (uintptr_t)b % 16 == 0) && (ndm == 1 || stride % 2 == 0);
    // contiguous chunks of at least 8192 samples (a multiple of one step of
    // the 4 waves), at most 4096 chunks per trial
    constexpr long long min_chunk = 8192;
    const long long step = (long long)(kFoldThreads / 64) * 64 * (wide ? 2 : 1) * kFoldUnroll;
    long long chunk = min_chunk;
    while ((long long)(n + chunk - 1) / chunk > 4096)
        chunk *= 2;
    chunk = (chunk + step - 1) / step * step;
    const long long gx = ((long long)n + chunk - 1) / chunk;
    // histogram (+ hit counts) and one parking area per wave (a step's values)
    const size_t lds = sizeof(float) * ((size_t)nbins * (hits ? 2 : 1) +
                                        (size_t)(kFoldThreads / 64) * kFoldUnroll * 64 * (wide ? 2 : 1));
    auto launch = [&](auto kern) {
        if (lds > 65536) {
            const hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds);
            if (e != hipSuccess)
                return e;
        }
        hipLaunchKernelGGL(kern, dim3((unsigned)gx, ndm), dim3(kFoldThreads), lds, (hipStream_t)stream,
                           (const float2 *)a, (const float2 *)b, (long long)n, (long long)stride, chunk, n0,
                           1.0 / period_samples, nbins, power, prof, hits);
        return hipSuccess;
    };
    const hipError_t le = wide ? (b ? launch(k_fold_phase<2, true>) : launch(k_fold_phase<2, false>))
                               : (b ? launch(k_fold_phase<1, true>) : launch(k_fold_phase<1, false>));
    FHIP(le);
    FHIP(hipGetLastError());
    return MSOSM_OK;
}

} // extern "C"
