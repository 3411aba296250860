// engine.hip -- the dedispersion engine behind the C ABI (include/msosm_c.h).
//
// One engine = one input stream of packed uint16 complex samples, num_dms DM
// trials sharing one FFT size 2M, count blocks of M samples per call.
//
// PLAIN_F32 keeps the reference's structure (msosm_gpu_dm_concurrent.cpp:
// 119-151): unpack into an (count+1)*M buffer whose head holds the previous
// call's tail, forward FFT of count overlapped windows, multiply by H and
// scatter into a per-DM product ring [R/count][dm][count][2M], inverse FFT of
// the ring group at the cursor, discard.  FUSED_F32 / FUSED_F16 store each
// forward spectrum once in a ring shared by all DMs and fold the per-bin delay
// gather, the H multiply, the discard and the scale into the inverse-FFT
// kernel.  All forms give output [dm][count][M] complex64 for block g equal to
// the last M points of IDFT(X_{g-d(n)}[n] H[n]) / (2M)^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "engine_internal.h"
#include "kernels.h"
#include "mfma_fft.h"
#include "msosm_c.h"
#include "plan.h"
#include "trace.h"

namespace {
thread_local std::string g_last_error;
} // namespace

int msosm::set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

namespace {
using msosm::set_error;

#define HIP_TRY(expr)                                                                                            \
    do {                                                                                                         \
        hipError_t e__ = (expr);                                                                                 \
        if (e__ != hipSuccess)                                                                                   \
            return set_error(MSOSM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e__), __FILE__,     \
                             __LINE__);                                                                          \
    } while (0)

int ilog2_exact(unsigned long v)
{
    int l = 0;
    while ((1ul << l) < v)
        l++;
    return l;
}

template <class T>
hipError_t dev_alloc(T **p, size_t elems)
{
    *p = nullptr;
    if (elems == 0)
        elems = 1;
    return hipMalloc((void **)p, elems * sizeof(T));
}

} // namespace

struct msosm_plan {
    msosm::Plan p;
};

struct StageTimer {
    int stage;
    hipEvent_t a, b;
};

struct msosm_engine {
    msosm::Plan plan;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int log2n = 0;
    int form = 0;
    // shared buffers
    uint32_t *in_stage = nullptr;   // 2 x count*M packed input (host input path, double-buffered)
    float2 *H32 = nullptr;          // [ndm][2M]
    __half2 *H16 = nullptr;         // [ndm][2M] = H/(2M) (FP16 forms)
    __half2 *Hp = nullptr;          // H16 in the matrix-core gather order (FUSED_F16_MFMA)
    uint32_t *dpk_mfma = nullptr;   // delays in that order, packed in pairs (FUSED_F16_MFMA)
    uint8_t *mfma_tab = nullptr;    // DFT tiles and twiddles of the matrix-core inverse
    int32_t *delay32 = nullptr;     // [ndm][2M] (PLAIN)
    uint16_t *delay16 = nullptr;    // [ndm][2M] (FUSED)
    uint32_t *delay_pk = nullptr;   // [ndm][E/2][NT] packed pairs (FUSED, one transform per workgroup)
    float2 *out = nullptr;          // [ndm][count][M] + 1 sentinel
    uint32_t *out_u16 = nullptr;    // packed output + 1 sentinel
    // PLAIN_F32
    float2 *inbuf = nullptr;        // (count+1)*M
    float2 *fftblk = nullptr;       // count*2M
    float2 *ring = nullptr;         // ndm*R*2M
    float2 *ifftblk = nullptr;      // ndm*count*2M
    float2 *fft_scratch = nullptr;  // two-pass FFT scratch
    int current_index = 0;
    // FUSED
    uint32_t *tail = nullptr;       // 2 x M: previous call's last M input samples / this call's
    void *xring = nullptr;          // rx*2M spectra (float2 or half2)
    int rx = 0;
    // FUSED at 2^15..2^21 points (four-step): ring, H and delays in position
    // order, FP32 scratch of the forward and of one inverse chunk
    bool big = false;
    float2 *Hbig = nullptr;         // [ndm][2M] H in position order (FUSED_F32)
    float2 *big_fwd = nullptr;      // big_fwd_max*2M
    float2 *big_inv = nullptr;      // big_inv_max*2M
    int big_fwd_max = 0, big_inv_max = 0; // windows / units the scratch holds
    // HANKEL (delay-line accumulate as per-bin GEMMs on MFMA)
    void *ybuf = nullptr;           // [ndm][count][2M] products (half2 or float2)
    int hankel_mode = -1;
    int kmax = 0;                   // largest delay + 1 over all trials
    int hankel_tpg = 0;             // DM tiles per Hankel workgroup (grid z groups)
    msosm_tuning tuning{};          // launch-shape requests (0 = the built-in rules)
    int hankel_kspan = 0;           // widest delay band one Hankel workgroup stages
    std::vector<int> hankel_lo_h;   // lowest delay per (trial group, bin group)
    int *hankel_lo = nullptr;       // its device copy
    long long g0 = 0;               // global index of the next block
    bool have_output = false;
    uint64_t plan_hash = 0;         // identity of the plan (state import check)
    // host input: copies run on copy_stream into alternate halves of in_stage;
    // stage_free[i] is recorded on the engine stream after the kernels that
    // read half i, and the next copy into that half waits for it
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_done = nullptr;
    hipEvent_t stage_free[2] = {nullptr, nullptr};
    bool stage_used[2] = {false, false};
    int stage_slot = 0;
    // timing
    bool timing = false;
    std::vector<StageTimer> pending;
    std::vector<hipEvent_t> pool;
    double stage_ms[MSOSM_NUM_STAGES] = {0};
    long long stage_calls[MSOSM_NUM_STAGES] = {0};
};

namespace {

hipEvent_t pool_get(msosm_engine *e)
{
    if (!e->pool.empty()) {
        hipEvent_t ev = e->pool.back();
        e->pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess)
        return nullptr;
    return ev;
}

void harvest(msosm_engine *e)
{
    for (auto &t : e->pending) {
        float ms = 0;
        if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            e->stage_ms[t.stage] += ms;
            e->stage_calls[t.stage] += 1;
        }
        e->pool.push_back(t.a);
        e->pool.push_back(t.b);
    }
    e->pending.clear();
}

struct StageScope {
    msosm_engine *e;
    int stage;
    hipStream_t s;
    hipEvent_t a = nullptr;
    StageScope(msosm_engine *eng, int st, hipStream_t on = nullptr) : e(eng), stage(st), s(on ? on : eng->stream)
    {
        if (e->timing) {
            if (e->pending.size() > 4096)
                harvest(e);
            a = pool_get(e);
            if (a)
                (void)hipEventRecord(a, s);
        }
    }
    ~StageScope()
    {
        if (e->timing && a) {
            hipEvent_t b = pool_get(e);
            if (b) {
                (void)hipEventRecord(b, s);
                e->pending.push_back({stage, a, b});
            } else {
                e->pool.push_back(a);
            }
        }
    }
};

void free_engine(msosm_engine *e)
{
    if (!e)
        return;
    if (e->stream)
        (void)hipStreamSynchronize(e->stream);
    harvest(e);
    for (auto ev : e->pool)
        (void)hipEventDestroy(ev);
    if (e->copy_stream) {
        (void)hipStreamSynchronize(e->copy_stream);
        (void)hipStreamDestroy(e->copy_stream);
    }
    if (e->copy_done)
        (void)hipEventDestroy(e->copy_done);
    for (hipEvent_t ev : e->stage_free)
        if (ev)
            (void)hipEventDestroy(ev);
    void *bufs[] = {e->in_stage, e->H32,  e->H16,     e->delay32,  e->delay16, e->delay_pk, e->Hp,
                    e->dpk_mfma, e->mfma_tab, e->out, e->out_u16, e->inbuf,
                    e->fftblk, e->ring, e->ifftblk, e->fft_scratch, e->tail, e->xring, e->ybuf, e->hankel_lo,
                    e->Hbig, e->big_fwd, e->big_inv};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    if (e->own_stream && e->stream)
        (void)hipStreamDestroy(e->stream);
    delete e;
}

// Filter tables H = DFT_2M(zero-pad(IDFT_M(C))) for every DM, on the GPU
// (msosm_prepare.cpp:37-84; the /M of :68-72 is overwritten by the memcpy at
// :74, so the IDFT is unnormalised).
int build_filter_tables(msosm_engine *e)
{
    const msosm::Plan &p = e->plan;
    const unsigned long M = p.M, N = p.N;
    const int ndm = p.num_dms;
    std::vector<msosm::ChirpParams> prm(ndm);
    for (int i = 0; i < ndm; i++) {
        const msosm::DmModel &m = p.dms[i];
        prm[i].fs = m.fs;
        prm[i].step = (float)(2 * 3.14159265358979323846 * (double)m.fs / (double)M); // :42
        prm[i].w0 = m.w0;
        prm[i].kdm = m.kdm;
        prm[i].dm = m.dm;
        prm[i].f0 = m.f0;
    }
    msosm::ChirpParams *prm_d = nullptr;
    float2 *c = nullptr, *h = nullptr, *pad = nullptr, *scratch = nullptr;
    int rc = MSOSM_OK;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(e->stream);
        if (prm_d)
            (void)hipFree(prm_d);
        if (c)
            (void)hipFree(c);
        if (h)
            (void)hipFree(h);
        if (pad)
            (void)hipFree(pad);
        if (scratch)
            (void)hipFree(scratch);
    };
    auto fail = [&](hipError_t err, const char *what) {
        cleanup();
        return set_error(MSOSM_ERR_HIP, "filter tables: %s: %s", what, hipGetErrorString(err));
    };
    hipError_t err;
    if ((err = dev_alloc(&prm_d, ndm)) != hipSuccess)
        return fail(err, "alloc params");
    if ((err = hipMemcpyAsync(prm_d, prm.data(), sizeof(msosm::ChirpParams) * ndm, hipMemcpyHostToDevice,
                              e->stream)) != hipSuccess)
        return fail(err, "copy params");
    if ((err = dev_alloc(&c, (size_t)ndm * M)) != hipSuccess || (err = dev_alloc(&h, (size_t)ndm * M)) != hipSuccess ||
        (err = dev_alloc(&pad, (size_t)ndm * N)) != hipSuccess)
        return fail(err, "alloc tables");
    size_t sc = std::max(msosm::fft_c2c_scratch_elems(ilog2_exact(M), ndm),
                         msosm::fft_c2c_scratch_elems(ilog2_exact(N), ndm));
    if (sc && (err = dev_alloc(&scratch, sc)) != hipSuccess)
        return fail(err, "alloc scratch");
    if ((err = msosm::launch_chirp(c, M, p.desc.chirp, prm_d, ndm, e->stream)) != hipSuccess)
        return fail(err, "chirp");
    if ((err = msosm::launch_fft_c2c(c, h, ilog2_exact(M), ndm, M, M, true, 1.0f, scratch, e->stream)) != hipSuccess)
        return fail(err, "IDFT_M");
    if ((err = msosm::launch_zero_pad(pad, h, M, ndm, e->stream)) != hipSuccess)
        return fail(err, "pad");
    if ((err = msosm::launch_fft_c2c(pad, e->H32, ilog2_exact(N), ndm, N, N, false, 1.0f, scratch, e->stream)) !=
        hipSuccess)
        return fail(err, "DFT_2M");
    if (e->H16) {
        // H/(2M) = H/N: with the ring's X/N the product is X H / N^2, whose
        // unnormalised inverse transform is the output (the FP32 forms'
        // IDFT(X H) / N^2) with no scale at the store
        if ((err = msosm::launch_to_half(e->H16, e->H32, (size_t)ndm * N, 0.5f / (float)M, e->stream)) != hipSuccess)
            return fail(err, "H to FP16");
    }
    cleanup();
    return rc;
}

bool form_f16_ring(int form)
{
    return form == MSOSM_FORM_FUSED_F16 || form == MSOSM_FORM_FUSED_F16_MFMA || form == MSOSM_FORM_HANKEL_F16;
}
bool form_hankel(int form)
{
    return form == MSOSM_FORM_HANKEL_F16 || form == MSOSM_FORM_HANKEL_BF16 || form == MSOSM_FORM_HANKEL_BF16X3;
}
int hankel_mode_of(int form)
{
    return form == MSOSM_FORM_HANKEL_F16 ? msosm::HANKEL_F16
           : form == MSOSM_FORM_HANKEL_BF16 ? msosm::HANKEL_BF16 : msosm::HANKEL_BF16X3;
}

// FNV-1a over what determines an engine's tables and buffer shapes: the
// trials' (bw, dm, f0), FFT size, blocks per call, method, chirp mode, form.
uint64_t plan_identity(const msosm::Plan &p)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *d, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(d);
        for (size_t i = 0; i < n; i++) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    const int ints[5] = {p.num_dms, p.count, p.desc.method, p.desc.chirp, p.desc.form};
    mix(ints, sizeof ints);
    mix(&p.N, sizeof p.N);
    for (int i = 0; i < p.num_dms; i++) {
        const float v[3] = {p.dms[i].bw, p.dms[i].dm, p.dms[i].f0};
        mix(v, sizeof v);
    }
    return h;
}

int create_engine(const msosm::Plan &plan, void *stream, msosm_engine **out)
{
    msosm_engine *e = new msosm_engine();
    e->plan = plan;
    e->form = plan.desc.form;
    e->log2n = ilog2_exact(plan.N);
    const msosm::Plan &p = e->plan;
    const unsigned long M = p.M, N = p.N;
    const int ndm = p.num_dms, count = p.count;
    auto bail = [&](int rc) {
        free_engine(e);
        return rc;
    };
    if (e->form == MSOSM_FORM_HANKEL_BF16)
        return bail(set_error(MSOSM_ERR_UNSUPPORTED,
                              "MSOSM_FORM_HANKEL_BF16 (single BF16 operands, 2^-8 relative rounding) misses the 1e-3 "
                              "mixed-precision gate; use MSOSM_FORM_HANKEL_BF16X3 (hi+lo split) or "
                              "MSOSM_FORM_HANKEL_F16"));
    e->plan_hash = plan_identity(plan);
    // the launchers size their runs from the device's CU count, cached per
    // device on first use: fill the cache here, outside any graph capture
    (void)msosm::device_cu_count();
    e->big = (e->form == MSOSM_FORM_FUSED_F32 || e->form == MSOSM_FORM_FUSED_F16) && msosm::big_supported(e->log2n);
    if (e->form != MSOSM_FORM_PLAIN_F32 && !msosm::fused_supported(e->log2n) && !e->big)
        return bail(set_error(MSOSM_ERR_UNSUPPORTED, "the fused forms support FFT sizes 4..2097152, the matrix-core "
                                                     "and Hankel forms 4..16384 (got %lu); use MSOSM_FORM_PLAIN_F32",
                              N));
    if (e->form == MSOSM_FORM_FUSED_F16_MFMA && e->log2n != msosm::mfft::kLog2N)
        return bail(set_error(MSOSM_ERR_UNSUPPORTED, "MSOSM_FORM_FUSED_F16_MFMA needs an FFT size of %d (plan has %lu)",
                              msosm::mfft::kN, N));
    if (e->form != MSOSM_FORM_PLAIN_F32) {
        // the shared-ring forms keep delays in 16 bits and address the spectra
        // ring with 32-bit byte offsets
        if (p.K_max > 65536)
            return bail(set_error(MSOSM_ERR_UNSUPPORTED,
                                  "delays of up to %d blocks do not fit the 16-bit delay tables of the shared-ring "
                                  "forms; use a larger FFT size or MSOSM_FORM_PLAIN_F32",
                                  p.K_max - 1));
        const size_t ring_bytes = (size_t)(p.K_max - 1 + p.count) * N *
                                  (form_f16_ring(e->form) ? sizeof(__half2) : sizeof(float2));
        if (ring_bytes >= ((size_t)1 << 32))
            return bail(set_error(MSOSM_ERR_UNSUPPORTED,
                                  "the spectra ring (%zu bytes for %d blocks of history and %d per call) must stay "
                                  "below 4 GiB: use fewer blocks per call or MSOSM_FORM_PLAIN_F32",
                                  ring_bytes, p.K_max - 1, p.count));
    }
    if (form_hankel(e->form)) {
        e->hankel_mode = hankel_mode_of(e->form);
        int km = 0;
        for (int32_t d : p.delay)
            km = std::max(km, (int)d + 1);
        e->kmax = km;
        if (N % (unsigned long)msosm::hankel_bins_per_wg(e->hankel_mode) != 0)
            return bail(set_error(MSOSM_ERR_UNSUPPORTED, "Hankel forms need an FFT size that is a multiple of %d",
                                  msosm::hankel_bins_per_wg(e->hankel_mode)));
        if (km > msosm::hankel_max_kmax(e->hankel_mode))
            return bail(set_error(MSOSM_ERR_UNSUPPORTED,
                                  "Hankel forms hold up to %d delay segments per bin in LDS (plan has %d); use a "
                                  "larger FFT size or a fused form",
                                  msosm::hankel_max_kmax(e->hankel_mode), km));
        msosm::hankel_groups(e->hankel_mode, p.delay.data(), (int)p.num_dms, (int)N, km, 0, &e->hankel_tpg,
                             &e->hankel_kspan, &e->hankel_lo_h);
    }
    int ndev = 0;
    hipError_t herr = hipGetDeviceCount(&ndev);
    if (herr != hipSuccess || ndev < 1)
        return bail(set_error(MSOSM_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(herr)));
    if (stream) {
        e->stream = (hipStream_t)stream;
    } else {
        herr = hipStreamCreate(&e->stream);
        if (herr != hipSuccess)
            return bail(set_error(MSOSM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(herr)));
        e->own_stream = true;
    }
#define ALLOC(ptr, n)                                                                                            \
    do {                                                                                                         \
        hipError_t a__ = dev_alloc(&(ptr), (n));                                                                 \
        if (a__ != hipSuccess)                                                                                   \
            return bail(set_error(MSOSM_ERR_ALLOC, "hipMalloc %s (%zu bytes): %s", #ptr,                         \
                                  (size_t)(n) * sizeof(*(ptr)), hipGetErrorString(a__)));                       \
    } while (0)
    const size_t out_elems = (size_t)ndm * count * M + 1;
    ALLOC(e->in_stage, 2 * (size_t)count * M);
    ALLOC(e->H32, (size_t)ndm * N);
    ALLOC(e->out, out_elems);
    ALLOC(e->out_u16, out_elems);
    if (e->form == MSOSM_FORM_PLAIN_F32) {
        ALLOC(e->delay32, (size_t)ndm * N);
        ALLOC(e->inbuf, (size_t)(count + 1) * M);
        ALLOC(e->fftblk, (size_t)count * N);
        ALLOC(e->ring, (size_t)ndm * p.R * N);
        ALLOC(e->ifftblk, (size_t)ndm * count * N);
        size_t sc = std::max(msosm::fft_c2c_scratch_elems(e->log2n, count),
                             msosm::fft_c2c_scratch_elems(e->log2n, (long long)ndm * count));
        if (sc)
            ALLOC(e->fft_scratch, sc);
        HIP_TRY(hipMemsetAsync(e->inbuf, 0, sizeof(float2) * (count + 1) * M, e->stream));
        HIP_TRY(hipMemsetAsync(e->ring, 0, sizeof(float2) * (size_t)ndm * p.R * N, e->stream));
        HIP_TRY(hipMemcpyAsync(e->delay32, p.delay.data(), sizeof(int32_t) * ndm * N, hipMemcpyHostToDevice,
                               e->stream));
    } else {
        const bool f16 = form_f16_ring(e->form);
        e->rx = p.K_max - 1 + count;
        if (form_hankel(e->form)) {
            if (f16) {
                __half2 *y = nullptr;
                ALLOC(y, (size_t)ndm * count * N);
                e->ybuf = y;
            } else {
                float2 *y = nullptr;
                ALLOC(y, (size_t)ndm * count * N);
                e->ybuf = y;
            }
            ALLOC(e->hankel_lo, e->hankel_lo_h.size());
            HIP_TRY(hipMemcpy(e->hankel_lo, e->hankel_lo_h.data(), sizeof(int) * e->hankel_lo_h.size(),
                              hipMemcpyHostToDevice));
        }
        ALLOC(e->delay16, (size_t)ndm * N);
        ALLOC(e->tail, 2 * (size_t)M); // two halves, alternating by call
        if (e->big) {
            e->big_fwd_max = msosm::big_chunk(e->log2n, count, 0);
            e->big_inv_max = msosm::big_chunk(e->log2n, (long long)ndm * count, 0);
            ALLOC(e->big_fwd, (size_t)e->big_fwd_max * N);
            ALLOC(e->big_inv, (size_t)e->big_inv_max * N);
            if (!f16)
                ALLOC(e->Hbig, (size_t)ndm * N);
        }
        if (f16) {
            ALLOC(e->H16, (size_t)ndm * N);
            __half2 *r = nullptr;
            ALLOC(r, (size_t)e->rx * N);
            e->xring = r;
            HIP_TRY(hipMemsetAsync(e->xring, 0, sizeof(__half2) * (size_t)e->rx * N, e->stream));
        } else {
            float2 *r = nullptr;
            ALLOC(r, (size_t)e->rx * N);
            e->xring = r;
            HIP_TRY(hipMemsetAsync(e->xring, 0, sizeof(float2) * (size_t)e->rx * N, e->stream));
        }
        HIP_TRY(hipMemsetAsync(e->tail, 0, sizeof(uint32_t) * 2 * M, e->stream));
        int32_t *tmp = nullptr;
        ALLOC(tmp, (size_t)ndm * N);
        std::vector<int32_t> perm;
        const int32_t *dsrc = p.delay.data();
        if (e->big) { // delays in the ring's position order
            perm.resize((size_t)ndm * N);
            for (int d = 0; d < ndm; d++)
                for (size_t q = 0; q < N; q++)
                    perm[(size_t)d * N + q] = p.delay[(size_t)d * N + (size_t)msosm::big_position_bin(e->log2n, (long long)q)];
            dsrc = perm.data();
        }
        herr = hipMemcpyAsync(tmp, dsrc, sizeof(int32_t) * ndm * N, hipMemcpyHostToDevice, e->stream);
        if (herr == hipSuccess)
            herr = msosm::launch_i32_to_u16(e->delay16, tmp, (size_t)ndm * N, e->stream);
        (void)hipStreamSynchronize(e->stream);
        (void)hipFree(tmp);
        if (herr != hipSuccess)
            return bail(set_error(MSOSM_ERR_HIP, "delay table upload: %s", hipGetErrorString(herr)));
        if (e->form == MSOSM_FORM_FUSED_F16_MFMA) {
            // delays in the matrix-core gather order, pairs of positions p0
            // and p0 + 64 (elements i and i + 1 of a lane's K-tile)
            std::vector<uint32_t> pk((size_t)ndm * N / 2);
            for (int d = 0; d < ndm; d++) {
                const int32_t *row = p.delay.data() + (size_t)d * N;
                uint32_t *dst = pk.data() + (size_t)d * N / 2;
                for (int pp = 0; pp < (int)(N / 2); pp++) {
                    const int p0 = ((pp >> 7) * 4 + 2 * ((pp >> 6) & 1)) * 64 + (pp & 63);
                    dst[pp] = (uint32_t)row[msosm::mfft::ring_bin(p0)] |
                              ((uint32_t)row[msosm::mfft::ring_bin(p0 + 64)] << 16);
                }
            }
            ALLOC(e->dpk_mfma, pk.size());
            HIP_TRY(hipMemcpy(e->dpk_mfma, pk.data(), sizeof(uint32_t) * pk.size(), hipMemcpyHostToDevice));
            ALLOC(e->Hp, (size_t)ndm * N);
            ALLOC(e->mfma_tab, msosm::inv_mfma_tables_bytes());
        } else if (!form_hankel(e->form) && msosm::fused_uses_packed_delay(e->log2n)) {
            const int NT = msosm::fused_delay_threads(e->log2n);
            const int E = (int)(N / NT);
            std::vector<uint32_t> pk((size_t)ndm * N / 2);
            for (int d = 0; d < ndm; d++) {
                const int32_t *row = p.delay.data() + (size_t)d * N;
                uint32_t *dst = pk.data() + (size_t)d * N / 2;
                for (int j = 0; j < E / 2; j++)
                    for (int t = 0; t < NT; t++)
                        dst[(size_t)j * NT + t] = (uint32_t)row[t + 2 * j * NT] |
                                                  ((uint32_t)row[t + (2 * j + 1) * NT] << 16);
            }
            ALLOC(e->delay_pk, pk.size());
            HIP_TRY(hipMemcpy(e->delay_pk, pk.data(), sizeof(uint32_t) * pk.size(), hipMemcpyHostToDevice));
        }
    }
#undef ALLOC
    HIP_TRY(hipMemsetAsync(e->out, 0, sizeof(float2) * out_elems, e->stream));
    HIP_TRY(hipMemsetAsync(e->out_u16, 0, sizeof(uint32_t) * out_elems, e->stream));
    int rc = build_filter_tables(e);
    if (rc != MSOSM_OK)
        return bail(rc);
    if (e->big) // H/(2M) (FP16) or H (FP32) in the ring's position order
        HIP_TRY(msosm::launch_big_permute(e->H32, e->H16 ? (void *)e->H16 : (void *)e->Hbig, e->H16 != nullptr,
                                          e->H16 ? 0.5f / (float)M : 1.0f, e->log2n, ndm, e->stream));
    if (e->form == MSOSM_FORM_FUSED_F16_MFMA) {
        HIP_TRY(msosm::launch_inv_mfma_filter(e->Hp, e->H16, ndm, e->stream));
        HIP_TRY(msosm::launch_inv_mfma_tables(e->mfma_tab, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    *out = e;
    return MSOSM_OK;
}

int filter_plain(msosm_engine *e, const uint32_t *cur)
{
    const msosm::Plan &p = e->plan;
    const unsigned long M = p.M, N = p.N;
    const int ndm = p.num_dms, count = p.count;
    hipStream_t s = e->stream;
    {
        StageScope st(e, MSOSM_STAGE_COPY); // overlap-save history (msosm_gpu_batch.cpp:80)
        HIP_TRY(msosm::launch_copy_u32(e->inbuf, e->inbuf + (size_t)count * M, 2 * M, s));
    }
    {
        StageScope st(e, MSOSM_STAGE_UNPACK);
        HIP_TRY(msosm::launch_unpack(e->inbuf + M, cur, (size_t)count * M, s));
    }
    {
        StageScope st(e, MSOSM_STAGE_FFT_FWD);
        HIP_TRY(msosm::launch_fft_c2c(e->inbuf, e->fftblk, e->log2n, count, M, N, false, 1.0f, e->fft_scratch, s));
    }
    {
        StageScope st(e, MSOSM_STAGE_MULTIPLY);
        HIP_TRY(msosm::launch_multiply_delay(e->fftblk, e->H32, (int)N, e->ring, e->delay32, p.R, e->current_index,
                                             count, ndm, s));
    }
    {
        StageScope st(e, MSOSM_STAGE_FFT_INV);
        const float2 *grp = e->ring + (size_t)e->current_index * N * ndm;
        HIP_TRY(msosm::launch_fft_c2c(grp, e->ifftblk, e->log2n, (long long)ndm * count, N, N, true, 1.0f,
                                      e->fft_scratch, s));
    }
    e->current_index += count; // msosm_gpu_dm_concurrent.cpp:142-146
    if (e->current_index >= p.R)
        e->current_index -= p.R;
    {
        StageScope st(e, MSOSM_STAGE_DISCARD);
        HIP_TRY(msosm::launch_discard(e->ifftblk, e->out, (int)M, count, ndm, s));
    }
    return MSOSM_OK;
}

// Shared-ring forms keep the previous call's last M input samples in one
// half of e->tail and write this call's into the other (the forward kernel
// reads one while its last window writes the other), alternating by call.
uint32_t *tail_in(msosm_engine *e) { return e->tail + (size_t)((e->g0 / e->plan.count) & 1) * e->plan.M; }
uint32_t *tail_out(msosm_engine *e) { return e->tail + (size_t)(((e->g0 / e->plan.count) & 1) ^ 1) * e->plan.M; }

// Hankel forms: shared forward spectra as in the fused forms, the
// delay-line accumulate as per-bin GEMMs on MFMA into dense products, then
// the inverse transform with the discard fused.
int filter_hankel(msosm_engine *e, const uint32_t *cur)
{
    const msosm::Plan &p = e->plan;
    const unsigned long M = p.M, N = p.N;
    const int ndm = p.num_dms, count = p.count;
    const bool f16 = form_f16_ring(e->form);
    hipStream_t s = e->stream;
    {
        StageScope st(e, MSOSM_STAGE_FFT_FWD);
        HIP_TRY(msosm::launch_fwd_u16(cur, tail_in(e), tail_out(e), e->xring, f16, e->log2n, count, e->g0, e->rx,
                                      false, s));
    }
    {
        StageScope st(e, MSOSM_STAGE_MULTIPLY);
        HIP_TRY(msosm::launch_hankel(e->xring, f16 ? (const void *)e->H16 : (const void *)e->H32, e->delay16, e->ybuf,
                                     e->hankel_mode, e->log2n, count, ndm, e->g0, e->rx, e->kmax, e->hankel_tpg,
                                     e->hankel_kspan, e->hankel_lo, s));
    }
    {
        StageScope st(e, MSOSM_STAGE_FFT_INV);
        // FP16 products are (X/N)(H/(2M)), FP32 products X H: both give
        // IDFT(X H) / N^2 after these scales
        const float scale = f16 ? 1.0f : 1.0f / ((float)N * (float)N);
        HIP_TRY(msosm::launch_inv_dense(e->ybuf, e->out, f16, e->log2n, (long long)ndm * count, scale,
                                        e->tuning.inv_blocks_per_wg, s));
    }
    e->g0 += count;
    return MSOSM_OK;
}

int filter_fused(msosm_engine *e, const uint32_t *cur)
{
    const msosm::Plan &p = e->plan;
    const unsigned long M = p.M;
    const int ndm = p.num_dms, count = p.count;
    const bool mfma = e->form == MSOSM_FORM_FUSED_F16_MFMA;
    const bool f16 = e->form == MSOSM_FORM_FUSED_F16 || mfma;
    hipStream_t s = e->stream;
    if (e->big) {
        // scratch chunks: the requested size, within what was allocated
        const size_t req = (size_t)e->tuning.scratch_chunk_kib << 10;
        const int fw = std::min(e->big_fwd_max, msosm::big_chunk(e->log2n, count, req));
        const int iu = std::min(e->big_inv_max, msosm::big_chunk(e->log2n, (long long)ndm * count, req));
        {
            StageScope st(e, MSOSM_STAGE_FFT_FWD);
            HIP_TRY(msosm::launch_big_forward(cur, tail_in(e), tail_out(e), e->xring, e->big_fwd, fw, f16, e->log2n,
                                              count, e->g0, e->rx, s));
        }
        {
            StageScope st(e, MSOSM_STAGE_FFT_INV);
            HIP_TRY(msosm::launch_big_inverse(e->xring, f16 ? (const void *)e->H16 : (const void *)e->Hbig,
                                              e->delay16, e->big_inv, iu, e->out, f16, e->log2n, count, ndm, e->g0,
                                              e->rx, s));
        }
        e->g0 += count;
        return MSOSM_OK;
    }
    {
        StageScope st(e, MSOSM_STAGE_FFT_FWD);
        HIP_TRY(msosm::launch_fwd_u16(cur, tail_in(e), tail_out(e), e->xring, f16, e->log2n, count, e->g0, e->rx,
                                      mfma, s));
    }
    {
        StageScope st(e, MSOSM_STAGE_FFT_INV);
        if (mfma)
            HIP_TRY(msosm::launch_inv_mfma(static_cast<const __half2 *>(e->xring), e->Hp, e->dpk_mfma, e->mfma_tab,
                                           e->out, count, ndm, e->g0, e->rx, e->tuning.inv_blocks_per_wg, s));
        else
            HIP_TRY(msosm::launch_inv_fused(e->xring, f16 ? (const void *)e->H16 : (const void *)e->H32, e->delay16,
                                            e->delay_pk, e->out, f16, e->log2n, count, ndm, e->g0, e->rx,
                                            e->tuning.inv_blocks_per_wg, s));
    }
    e->g0 += count;
    return MSOSM_OK;
}

} // namespace

// Zero the stream history with fill kernels only (safe inside graph capture).
int msosm::engine_reset_async(msosm_engine *e)
{
    const msosm::Plan &p = e->plan;
    if (e->form == MSOSM_FORM_PLAIN_F32) {
        HIP_TRY(msosm::launch_fill_u32(e->inbuf, 2 * (size_t)(p.count + 1) * p.M, 0, e->stream));
        HIP_TRY(msosm::launch_fill_u32(e->ring, 2 * (size_t)p.num_dms * p.R * p.N, 0, e->stream));
        e->current_index = 0;
    } else {
        size_t words = (form_f16_ring(e->form) ? 1 : 2) * (size_t)e->rx * p.N;
        HIP_TRY(msosm::launch_fill_u32(e->xring, words, 0, e->stream));
        HIP_TRY(msosm::launch_fill_u32(e->tail, 2 * (size_t)p.M, 0, e->stream));
        e->g0 = 0;
    }
    return MSOSM_OK;
}

hipStream_t msosm::engine_stream(msosm_engine *e) { return e->stream; }
float2 *msosm::engine_output(msosm_engine *e) { return e->out; }
const msosm::Plan &msosm::engine_plan(msosm_engine *e) { return e->plan; }
msosm::EngineCursor msosm::engine_cursor(const msosm_engine *e)
{
    msosm::EngineCursor c;
    c.g0 = e->g0;
    c.current_index = e->current_index;
    c.have_output = e->have_output;
    return c;
}
void msosm::engine_set_cursor(msosm_engine *e, const msosm::EngineCursor &c)
{
    e->g0 = c.g0;
    e->current_index = c.current_index;
    e->have_output = c.have_output;
}
bool msosm::engine_timing(const msosm_engine *e) { return e->timing; }

namespace {
// Stream state of an engine as one byte image: this header, then the
// sections of state_sections() in order.
struct StateHeader {
    char magic[8];          // "MSOSMST1"
    uint64_t plan_hash;
    int32_t form, num_dms, count, ring_slots;
    uint64_t M;
    int64_t g0;
    int32_t current_index, have_output;
    uint64_t section_bytes[2];
};
constexpr char kStateMagic[8] = {'M', 'S', 'O', 'S', 'M', 'S', 'T', '1'};

// The history a filter call reads: reference-structured form -- the last M
// input samples (head of the next call's overlapped input) and the product
// ring; shared-ring forms -- both halves of the packed-input tail and the
// spectra ring.
void state_sections(msosm_engine *e, void *ptr[2], size_t bytes[2])
{
    const msosm::Plan &p = e->plan;
    if (e->form == MSOSM_FORM_PLAIN_F32) {
        ptr[0] = e->inbuf + (size_t)p.count * p.M;
        bytes[0] = sizeof(float2) * p.M;
        ptr[1] = e->ring;
        bytes[1] = sizeof(float2) * (size_t)p.num_dms * p.R * p.N;
    } else {
        ptr[0] = e->tail;
        bytes[0] = sizeof(uint32_t) * 2 * p.M;
        ptr[1] = e->xring;
        bytes[1] = (form_f16_ring(e->form) ? sizeof(__half2) : sizeof(float2)) * (size_t)e->rx * p.N;
    }
}
} // namespace

// ===========================================================================
// C ABI

extern "C" {

int msosm_host_alloc(size_t bytes, void **out)
{
    if (!out || bytes == 0)
        return set_error(MSOSM_ERR_ARG, "host_alloc: bad argument");
    *out = nullptr;
    HIP_TRY(hipHostMalloc(out, bytes, hipHostMallocDefault));
    return MSOSM_OK;
}

int msosm_host_free(void *p)
{
    if (p)
        HIP_TRY(hipHostFree(p));
    return MSOSM_OK;
}

const char *msosm_last_error(void) { return g_last_error.c_str(); }
const char *msosm_version(void) { return "msosm-mi355x 0.1 (gfx950)"; }

int msosm_plan_create(const msosm_plan_desc *desc, msosm_plan **out)
{
    if (!desc || !out)
        return set_error(MSOSM_ERR_ARG, "plan_create: NULL argument");
    *out = nullptr;
    msosm_plan *p = new msosm_plan();
    std::string msg;
    int rc = msosm::build_plan(*desc, p->p, msg);
    if (rc != MSOSM_OK) {
        delete p;
        return set_error(rc, "%s", msg.c_str());
    }
    *out = p;
    return MSOSM_OK;
}

void msosm_plan_destroy(msosm_plan *plan) { delete plan; }

int msosm_dm_model(float bw, float dm, float f0, int method, unsigned long *nd, unsigned long *m_min)
{
    if (!(bw > 0) || !(f0 > 0) || !(dm >= 0) || (method != MSOSM_METHOD_MSOSM && method != MSOSM_METHOD_OSM))
        return set_error(MSOSM_ERR_ARG, "dm_model: need bw > 0, f0 > 0, dm >= 0 and a known method");
    msosm::DmModel m = msosm::make_dm_model(bw, dm, f0);
    if (nd)
        *nd = m.nd;
    if (m_min)
        *m_min = method == MSOSM_METHOD_MSOSM ? msosm::msosm_min_order(m.nd) : msosm::osm_min_order(m.nd);
    return MSOSM_OK;
}

int msosm_plan_get_info(const msosm_plan *plan, msosm_plan_info *info)
{
    if (!plan || !info)
        return set_error(MSOSM_ERR_ARG, "plan_get_info: NULL argument");
    const msosm::Plan &p = plan->p;
    info->M = p.M;
    info->M_min = p.M_min;
    info->fftpoint = p.N;
    info->count = p.count;
    info->num_dms = p.num_dms;
    info->K_max = p.K_max;
    info->R = p.R;
    info->clamped = p.clamped;
    return MSOSM_OK;
}

int msosm_plan_dm_info(const msosm_plan *plan, int i, unsigned long *nd, unsigned long *m_min, int *delaycount,
                       int *delay_block_size)
{
    if (!plan || i < 0 || i >= plan->p.num_dms)
        return set_error(MSOSM_ERR_ARG, "plan_dm_info: bad plan or DM index %d", i);
    const msosm::Plan &p = plan->p;
    if (nd)
        *nd = p.dms[i].nd;
    if (m_min)
        *m_min = p.m_min[i];
    if (delaycount)
        *delaycount = p.K[i];
    if (delay_block_size)
        *delay_block_size = p.delay_block_size[i];
    return MSOSM_OK;
}

int msosm_plan_delay_points(const msosm_plan *plan, int i, int32_t *out)
{
    if (!plan || !out || i < 0 || i >= plan->p.num_dms)
        return set_error(MSOSM_ERR_ARG, "plan_delay_points: bad argument");
    const msosm::Plan &p = plan->p;
    std::memcpy(out, p.delay.data() + (size_t)i * p.N, sizeof(int32_t) * p.N);
    return MSOSM_OK;
}

int msosm_engine_create(const msosm_plan *plan, void *hip_stream, msosm_engine **out)
{
    if (!plan || !out)
        return set_error(MSOSM_ERR_ARG, "engine_create: NULL argument");
    *out = nullptr;
    return create_engine(plan->p, hip_stream, out);
}

int msosm_engine_destroy(msosm_engine *eng)
{
    free_engine(eng);
    return MSOSM_OK;
}

int msosm_engine_filter_u16(msosm_engine *e, const void *input, int where)
{
    if (!e || !input)
        return set_error(MSOSM_ERR_ARG, "filter_u16: NULL argument");
    msosm::TraceRange range("msosm_engine_filter_u16");
    const msosm::Plan &p = e->plan;
    const uint32_t *cur = (const uint32_t *)input;
    const int slot = e->stage_slot;
    if (where == 0) {
        // The caller may refill its host buffer as soon as this returns (the
        // reference's blocking cudaMemcpy), so wait for the copy.  It runs on
        // the copy stream into the staging half that the call before the
        // previous one read, so it overlaps the previous call's kernels
        // instead of queueing behind them.
        if (!e->copy_stream) {
            HIP_TRY(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&e->copy_done, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&e->stage_free[0], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&e->stage_free[1], hipEventDisableTiming));
        }
        uint32_t *dst = e->in_stage + (size_t)slot * p.count * p.M;
        if (e->stage_used[slot])
            HIP_TRY(hipStreamWaitEvent(e->copy_stream, e->stage_free[slot], 0));
        {
            StageScope st(e, MSOSM_STAGE_H2D, e->copy_stream);
            HIP_TRY(hipMemcpyAsync(dst, input, sizeof(uint32_t) * (size_t)p.count * p.M, hipMemcpyHostToDevice,
                                   e->copy_stream));
        }
        HIP_TRY(hipEventRecord(e->copy_done, e->copy_stream));
        HIP_TRY(hipEventSynchronize(e->copy_done));
        HIP_TRY(hipStreamWaitEvent(e->stream, e->copy_done, 0));
        cur = dst;
    } else if (where != 1) {
        return set_error(MSOSM_ERR_ARG, "filter_u16: where must be 0 (host) or 1 (device)");
    }
    int rc = (e->form == MSOSM_FORM_PLAIN_F32) ? filter_plain(e, cur)
             : form_hankel(e->form)             ? filter_hankel(e, cur)
                                                : filter_fused(e, cur);
    if (rc == MSOSM_OK)
        e->have_output = true;
    if (where == 0) {
        HIP_TRY(hipEventRecord(e->stage_free[slot], e->stream));
        e->stage_used[slot] = true;
        e->stage_slot = slot ^ 1;
    }
    return rc;
}

int msosm_engine_output_c64(msosm_engine *e, void *dst, int where)
{
    if (!e || !dst)
        return set_error(MSOSM_ERR_ARG, "output_c64: NULL argument");
    msosm::TraceRange range("msosm_engine_output_c64");
    const msosm::Plan &p = e->plan;
    size_t bytes = sizeof(float2) * (size_t)p.num_dms * p.count * p.M;
    HIP_TRY(hipMemcpyAsync(dst, e->out, bytes, where == 1 ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           e->stream));
    if (where != 1)
        HIP_TRY(hipStreamSynchronize(e->stream));
    return MSOSM_OK;
}

int msosm_engine_output_u16(msosm_engine *e, void *dst, int where)
{
    if (!e || !dst)
        return set_error(MSOSM_ERR_ARG, "output_u16: NULL argument");
    msosm::TraceRange range("msosm_engine_output_u16");
    const msosm::Plan &p = e->plan;
    size_t n = (size_t)p.num_dms * p.count * p.M;
    HIP_TRY(msosm::launch_pack(e->out_u16, e->out, n, e->stream));
    HIP_TRY(hipMemcpyAsync(dst, e->out_u16, sizeof(uint32_t) * n,
                           where == 1 ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, e->stream));
    if (where != 1)
        HIP_TRY(hipStreamSynchronize(e->stream));
    return MSOSM_OK;
}

int msosm_engine_output_ptr(msosm_engine *e, void **dev_ptr)
{
    if (!e || !dev_ptr)
        return set_error(MSOSM_ERR_ARG, "output_ptr: NULL argument");
    *dev_ptr = e->out;
    return MSOSM_OK;
}

int msosm_engine_dedisp_params(msosm_engine *e, int i, float *host_out)
{
    if (!e || !host_out || i < 0 || i >= e->plan.num_dms)
        return set_error(MSOSM_ERR_ARG, "dedisp_params: bad argument");
    HIP_TRY(hipMemcpyAsync(host_out, e->H32 + (size_t)i * e->plan.N, sizeof(float2) * e->plan.N,
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return MSOSM_OK;
}

int msosm_engine_synchronize(msosm_engine *e)
{
    if (!e)
        return set_error(MSOSM_ERR_ARG, "synchronize: NULL engine");
    HIP_TRY(hipStreamSynchronize(e->stream));
    return MSOSM_OK;
}

int msosm_engine_reset(msosm_engine *e)
{
    if (!e)
        return set_error(MSOSM_ERR_ARG, "reset: NULL engine");
    int rc = msosm::engine_reset_async(e);
    if (rc != MSOSM_OK)
        return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return MSOSM_OK;
}

int msosm_engine_state_size(msosm_engine *e, size_t *bytes)
{
    if (!e || !bytes)
        return set_error(MSOSM_ERR_ARG, "state_size: NULL argument");
    void *ptr[2];
    size_t sec[2];
    state_sections(e, ptr, sec);
    *bytes = sizeof(StateHeader) + sec[0] + sec[1];
    return MSOSM_OK;
}

int msosm_engine_state_export(msosm_engine *e, void *dst, size_t bytes, int where)
{
    if (!e || !dst || (where != 0 && where != 1))
        return set_error(MSOSM_ERR_ARG, "state_export: bad argument");
    msosm::TraceRange range("msosm_engine_state_export");
    void *ptr[2];
    size_t sec[2];
    state_sections(e, ptr, sec);
    if (bytes < sizeof(StateHeader) + sec[0] + sec[1])
        return set_error(MSOSM_ERR_ARG, "state_export: buffer of %zu bytes, need %zu", bytes,
                         sizeof(StateHeader) + sec[0] + sec[1]);
    StateHeader h{};
    std::memcpy(h.magic, kStateMagic, sizeof h.magic);
    h.plan_hash = e->plan_hash;
    h.form = e->form;
    h.num_dms = e->plan.num_dms;
    h.count = e->plan.count;
    h.ring_slots = e->form == MSOSM_FORM_PLAIN_F32 ? e->plan.R : e->rx;
    h.M = e->plan.M;
    h.g0 = e->g0;
    h.current_index = e->current_index;
    h.have_output = e->have_output ? 1 : 0;
    h.section_bytes[0] = sec[0];
    h.section_bytes[1] = sec[1];
    char *d = static_cast<char *>(dst);
    const hipMemcpyKind kind = where == 1 ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (where == 1)
        HIP_TRY(hipMemcpyAsync(d, &h, sizeof h, hipMemcpyHostToDevice, e->stream));
    else
        std::memcpy(d, &h, sizeof h);
    HIP_TRY(hipMemcpyAsync(d + sizeof h, ptr[0], sec[0], kind, e->stream));
    HIP_TRY(hipMemcpyAsync(d + sizeof h + sec[0], ptr[1], sec[1], kind, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return MSOSM_OK;
}

int msosm_engine_state_import(msosm_engine *e, const void *src, size_t bytes, int where)
{
    if (!e || !src || (where != 0 && where != 1))
        return set_error(MSOSM_ERR_ARG, "state_import: bad argument");
    msosm::TraceRange range("msosm_engine_state_import");
    void *ptr[2];
    size_t sec[2];
    state_sections(e, ptr, sec);
    if (bytes < sizeof(StateHeader) + sec[0] + sec[1])
        return set_error(MSOSM_ERR_ARG, "state_import: image of %zu bytes, this engine's state is %zu", bytes,
                         sizeof(StateHeader) + sec[0] + sec[1]);
    const char *s = static_cast<const char *>(src);
    StateHeader h{};
    if (where == 1)
        HIP_TRY(hipMemcpy(&h, s, sizeof h, hipMemcpyDeviceToHost));
    else
        std::memcpy(&h, s, sizeof h);
    if (std::memcmp(h.magic, kStateMagic, sizeof h.magic) != 0)
        return set_error(MSOSM_ERR_ARG, "state_import: not an engine state image");
    const int slots = e->form == MSOSM_FORM_PLAIN_F32 ? e->plan.R : e->rx;
    if (h.plan_hash != e->plan_hash || h.form != e->form || h.num_dms != e->plan.num_dms ||
        h.count != e->plan.count || h.ring_slots != slots || h.M != e->plan.M || h.section_bytes[0] != sec[0] ||
        h.section_bytes[1] != sec[1])
        return set_error(MSOSM_ERR_ARG, "state_import: the image belongs to a different plan (trials, FFT size, "
                                        "blocks per call, method, chirp mode or form differ)");
    const hipMemcpyKind kind = where == 1 ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpyAsync(ptr[0], s + sizeof h, sec[0], kind, e->stream));
    HIP_TRY(hipMemcpyAsync(ptr[1], s + sizeof h + sec[0], sec[1], kind, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->g0 = h.g0;
    e->current_index = h.current_index;
    e->have_output = h.have_output != 0;
    return MSOSM_OK;
}

int msosm_engine_set_tuning(msosm_engine *e, const msosm_tuning *t)
{
    if (!e || !t || t->inv_blocks_per_wg < 0 || t->inv_blocks_per_wg > 64 || t->hankel_tiles_per_group < 0 ||
        t->scratch_chunk_kib < 0)
        return set_error(MSOSM_ERR_ARG, "set_tuning: bad argument (0 <= inv_blocks_per_wg <= 64, "
                                        "hankel_tiles_per_group >= 0, scratch_chunk_kib >= 0)");
    if (form_hankel(e->form) && t->hankel_tiles_per_group != e->tuning.hankel_tiles_per_group) {
        int tpg = 0, kspan = 0;
        std::vector<int> lo;
        msosm::hankel_groups(e->hankel_mode, e->plan.delay.data(), e->plan.num_dms, (int)e->plan.N, e->kmax,
                             t->hankel_tiles_per_group, &tpg, &kspan, &lo);
        int *lo_d = nullptr;
        HIP_TRY(dev_alloc(&lo_d, lo.size()));
        hipError_t herr = hipMemcpy(lo_d, lo.data(), sizeof(int) * lo.size(), hipMemcpyHostToDevice);
        if (herr != hipSuccess) {
            (void)hipFree(lo_d);
            return set_error(MSOSM_ERR_HIP, "set_tuning: band table upload: %s", hipGetErrorString(herr));
        }
        HIP_TRY(hipStreamSynchronize(e->stream)); // the old table may still be read
        (void)hipFree(e->hankel_lo);
        e->hankel_lo = lo_d;
        e->hankel_lo_h.swap(lo);
        e->hankel_tpg = tpg;
        e->hankel_kspan = kspan;
    }
    e->tuning = *t;
    return MSOSM_OK;
}

int msosm_engine_get_tuning(msosm_engine *e, msosm_tuning *t)
{
    if (!e || !t)
        return set_error(MSOSM_ERR_ARG, "get_tuning: NULL argument");
    *t = e->tuning;
    if (form_hankel(e->form))
        t->hankel_tiles_per_group = e->hankel_tpg;
    return MSOSM_OK;
}

int msosm_range_push(const char *name)
{
    roctxRangePushA(name ? name : "");
    return MSOSM_OK;
}

int msosm_range_pop(void)
{
    roctxRangePop();
    return MSOSM_OK;
}

int msosm_engine_stream(msosm_engine *e, void **s)
{
    if (!e || !s)
        return set_error(MSOSM_ERR_ARG, "stream: NULL argument");
    *s = (void *)e->stream;
    return MSOSM_OK;
}

int msosm_engine_set_timing(msosm_engine *e, int on)
{
    if (!e)
        return set_error(MSOSM_ERR_ARG, "set_timing: NULL engine");
    harvest(e);
    e->timing = on != 0;
    for (int i = 0; i < MSOSM_NUM_STAGES; i++) {
        e->stage_ms[i] = 0;
        e->stage_calls[i] = 0;
    }
    return MSOSM_OK;
}

int msosm_engine_get_timing(msosm_engine *e, double *times_ms, long long *calls, int n)
{
    if (!e)
        return set_error(MSOSM_ERR_ARG, "get_timing: NULL engine");
    harvest(e);
    for (int i = 0; i < n && i < MSOSM_NUM_STAGES; i++) {
        if (times_ms)
            times_ms[i] = e->stage_ms[i];
        if (calls)
            calls[i] = e->stage_calls[i];
    }
    return MSOSM_OK;
}

// ---- element kernels -------------------------------------------------------
int msosm_k_uint16_to_complex(float *dst, const void *src, size_t n, void *stream)
{
    HIP_TRY(msosm::launch_unpack((float2 *)dst, (const uint32_t *)src, n, (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_k_complex_to_uint16(void *dst, const float *src, size_t n, void *stream)
{
    HIP_TRY(msosm::launch_pack((uint32_t *)dst, (const float2 *)src, n, (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_k_intensity(const float *a, const float *b, float *out, size_t n, void *stream)
{
    HIP_TRY(msosm::launch_intensity((const float2 *)a, (const float2 *)b, out, n, (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_k_multiply_batch(const float *a, const float *b, int N, float *out, int count, int ndm, void *stream)
{
    if (N < 1 || count < 1 || ndm < 1)
        return set_error(MSOSM_ERR_ARG, "multiply_batch: bad sizes");
    HIP_TRY(msosm::launch_multiply((const float2 *)a, (const float2 *)b, N, (float2 *)out, count, ndm,
                                   (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_k_multiply_delay_batch(const float *a, const float *b, int N, float *ring, const int32_t *delay, int dbs,
                                 int index, int count, int ndm, void *stream)
{
    if (N < 1 || count < 1 || ndm < 1 || dbs < 1 || dbs % count != 0 || index < 0)
        return set_error(MSOSM_ERR_ARG, "multiply_delay_batch: bad sizes (delay_block_size must be a positive "
                                        "multiple of count)");
    HIP_TRY(msosm::launch_multiply_delay((const float2 *)a, (const float2 *)b, N, (float2 *)ring, delay, dbs, index,
                                         count, ndm, (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_k_discard_batch(const float *a, float *b, int M, int count, int ndm, void *stream)
{
    if (M < 1 || count < 1 || ndm < 1)
        return set_error(MSOSM_ERR_ARG, "discard_batch: bad sizes");
    HIP_TRY(msosm::launch_discard((const float2 *)a, (float2 *)b, M, count, ndm, (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_k_hankel_delay_gemm(const void *ring, const void *H, const uint16_t *delay, void *Y, int form, int log2n,
                              int count, int num_dms, long long g0, int rx, int kmax, void *stream)
{
    if (!ring || !H || !delay || !Y || !form_hankel(form) || log2n < 5 || log2n > 14 || count < 1 ||
        num_dms < 1 || rx < count || kmax < 1 || kmax > rx - count + 1 || g0 < 0)
        return set_error(MSOSM_ERR_ARG, "hankel_delay_gemm: bad argument (Hankel form, 5 <= log2n <= 14, "
                                        "rx >= count, 1 <= kmax <= rx - count + 1)");
    const int mode = hankel_mode_of(form);
    if (kmax > msosm::hankel_max_kmax(mode))
        return set_error(MSOSM_ERR_UNSUPPORTED, "hankel_delay_gemm: kmax %d above %d", kmax,
                         msosm::hankel_max_kmax(mode));
    HIP_TRY(msosm::launch_hankel(ring, H, delay, Y, mode, log2n, count, num_dms, g0, rx, kmax, 0, kmax, nullptr,
                                 (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_k_inverse_discard(const void *Y, float *out, int y_f16, int log2n, size_t units, float scale, void *stream)
{
    if (!Y || !out || log2n < 2 || log2n > 14 || units < 1)
        return set_error(MSOSM_ERR_ARG, "inverse_discard: bad argument (2 <= log2n <= 14, units >= 1)");
    HIP_TRY(msosm::launch_inv_dense(Y, (float2 *)out, y_f16 != 0, log2n, (long long)units, scale, 0,
                                    (hipStream_t)stream));
    return MSOSM_OK;
}

struct msosm_fft_plan {
    int log2n = 0;
    size_t batch = 0, idist = 0, odist = 0;
    float2 *scratch = nullptr;
};

int msosm_fft_plan_create(size_t n, size_t batch, size_t idist, size_t odist, msosm_fft_plan **out)
{
    if (!out || n < 2 || (n & (n - 1)) || n > (1ul << 28) || batch < 1 || idist < 1 || odist < 1)
        return set_error(MSOSM_ERR_ARG, "fft_plan_create: n must be a power of two in [2, 2^28], batch >= 1");
    *out = nullptr;
    msosm_fft_plan *p = new msosm_fft_plan();
    p->log2n = ilog2_exact(n);
    p->batch = batch;
    p->idist = idist;
    p->odist = odist;
    const size_t sc = msosm::fft_c2c_scratch_elems(p->log2n, (long long)batch);
    if (sc) {
        hipError_t e = dev_alloc(&p->scratch, sc);
        if (e != hipSuccess) {
            delete p;
            return set_error(MSOSM_ERR_ALLOC, "fft_plan_create: scratch of %zu bytes: %s", sc * sizeof(float2),
                             hipGetErrorString(e));
        }
    }
    *out = p;
    return MSOSM_OK;
}

int msosm_fft_plan_exec(msosm_fft_plan *p, const float *in, float *out, int direction, void *stream)
{
    if (!p || !in || !out || (direction != -1 && direction != 1))
        return set_error(MSOSM_ERR_ARG, "fft_plan_exec: bad argument (direction -1 or +1)");
    HIP_TRY(msosm::launch_fft_c2c((const float2 *)in, (float2 *)out, p->log2n, (long long)p->batch,
                                  (long long)p->idist, (long long)p->odist, direction == 1, 1.0f, p->scratch,
                                  (hipStream_t)stream));
    return MSOSM_OK;
}

int msosm_fft_plan_destroy(msosm_fft_plan *p)
{
    if (!p)
        return MSOSM_OK;
    if (p->scratch) {
        // execs may still be in flight on any stream
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(p->scratch));
    }
    delete p;
    return MSOSM_OK;
}

int msosm_k_fft_c2c(const float *in, float *out, size_t n, size_t batch, size_t idist, size_t odist, int direction,
                    void *stream)
{
    if (n < 2 || (n & (n - 1)) || n > (1ul << 28) || batch < 1 || (direction != -1 && direction != 1))
        return set_error(MSOSM_ERR_ARG, "fft_c2c: n must be a power of two in [2, 2^28], direction -1 or +1");
    HIP_TRY(msosm::launch_fft_c2c((const float2 *)in, (float2 *)out, ilog2_exact(n), (long long)batch,
                                  (long long)idist, (long long)odist, direction == 1, 1.0f, nullptr,
                                  (hipStream_t)stream));
    return MSOSM_OK;
}

} // extern "C"
