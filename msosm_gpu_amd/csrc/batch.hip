// batch.hip -- many independent input streams in one graph launch, and the
// FFT-size autotuner.
//
// A batch owns nothing but its graph and its fold profiles: the engines
// (one per input stream: its own bandwidth, DM trials, FFT size, count) are
// created by the caller.  At creation the whole pass -- reset of every
// engine, every filter call over every stream's resident input, and a phase
// fold of every call's output -- is captured once into a hipGraph with one
// branch per stream (streams fork from and join back to the origin stream),
// so one hipGraphLaunch replays a full pass over all streams.  Resets use
// fill kernels, never memset nodes.
//
// The reference runs independent streams as one CUDA stream per engine with
// host-side launches per block (dedispersion/msosm_stream/
// msosm_gpu_batch_stream.cpp:105-149); it has no graph and no autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "engine_internal.h"
#include "kernels.h"
#include "msosm_c.h"
#include "trace.h"

namespace {

#define BHIP(expr)                                                                                               \
    do {                                                                                                         \
        hipError_t e__ = (expr);                                                                                 \
        if (e__ != hipSuccess)                                                                                   \
            return msosm::set_error(MSOSM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e__));              \
    } while (0)

} // namespace

struct msosm_batch {
    std::vector<msosm_engine *> engines;
    std::vector<const uint32_t *> inputs;
    std::vector<unsigned long> samples;
    std::vector<double> period;
    int nbins = 0;
    std::vector<float *> prof, hits;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    hipStream_t origin = nullptr;
    bool own_origin = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
    // host-side engine positions after one replayed pass: capture walks each
    // engine's cursor through the pass without running it, so the cursors
    // are put back after capture and moved to these after each launch
    std::vector<msosm::EngineCursor> after_pass;
};

namespace {

void batch_free(msosm_batch *b)
{
    if (!b)
        return;
    if (b->origin)
        (void)hipStreamSynchronize(b->origin);
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    for (float *p : b->prof)
        if (p)
            (void)hipFree(p);
    for (float *p : b->hits)
        if (p)
            (void)hipFree(p);
    for (hipEvent_t ev : b->join)
        if (ev)
            (void)hipEventDestroy(ev);
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    if (b->own_origin && b->origin)
        (void)hipStreamDestroy(b->origin);
    delete b;
}

// Enqueue one full pass of stream i on its engine's stream (during capture).
int enqueue_stream(msosm_batch *b, int i)
{
    msosm_engine *e = b->engines[i];
    const msosm::Plan &p = msosm::engine_plan(e);
    hipStream_t s = msosm::engine_stream(e);
    const size_t per = (size_t)p.count * p.M;
    const size_t calls = b->samples[i] / per;
    int rc = msosm::engine_reset_async(e);
    if (rc != MSOSM_OK)
        return rc;
    if (b->prof[i]) {
        BHIP(msosm::launch_fill_u32(b->prof[i], (size_t)p.num_dms * b->nbins, 0, s));
        BHIP(msosm::launch_fill_u32(b->hits[i], (size_t)b->nbins, 0, s));
    }
    for (size_t c = 0; c < calls; c++) {
        rc = msosm_engine_filter_u16(e, b->inputs[i] + c * per, 1);
        if (rc != MSOSM_OK)
            return rc;
        if (b->prof[i]) {
            rc = msosm_fold_phase(msosm::engine_output(e), nullptr, per, p.num_dms, per, (long long)(c * per),
                                  b->period[i], b->nbins, 1, b->prof[i], b->hits[i], (void *)s);
            if (rc != MSOSM_OK)
                return rc;
        }
    }
    return MSOSM_OK;
}

} // namespace

extern "C" {

int msosm_batch_create(msosm_engine *const *engines, int n, const void *const *inputs,
                       const unsigned long *num_samples, const double *period_samples, int nbins, void *stream,
                       msosm_batch **out)
{
    if (!engines || !inputs || !num_samples || !out || n < 1 || (period_samples && (nbins < 1 || nbins > 8192)))
        return msosm::set_error(MSOSM_ERR_ARG, "batch_create: bad argument");
    *out = nullptr;
    msosm_batch *b = new msosm_batch();
    auto fail = [&](int rc) {
        batch_free(b);
        return rc;
    };
    b->nbins = period_samples ? nbins : 0;
    for (int i = 0; i < n; i++) {
        if (!engines[i] || !inputs[i])
            return fail(msosm::set_error(MSOSM_ERR_ARG, "batch_create: stream %d has no engine or input", i));
        const msosm::Plan &p = msosm::engine_plan(engines[i]);
        const unsigned long per = (unsigned long)p.count * p.M;
        if (num_samples[i] < per)
            return fail(msosm::set_error(MSOSM_ERR_ARG, "batch_create: stream %d has %lu samples, fewer than one "
                                                        "call (%lu)", i, num_samples[i], per));
        for (int j = 0; j < i; j++)
            if (engines[j] == engines[i])
                return fail(msosm::set_error(MSOSM_ERR_ARG, "batch_create: engine %d listed twice", i));
        if (msosm::engine_timing(engines[i]))
            return fail(msosm::set_error(MSOSM_ERR_ARG, "batch_create: engine %d has per-stage timing on (its "
                                                        "event records cannot be captured); turn it off first",
                                         i));
        b->engines.push_back(engines[i]);
        b->inputs.push_back((const uint32_t *)inputs[i]);
        b->samples.push_back(num_samples[i]);
        double per_s = period_samples ? period_samples[i] : 0.0;
        b->period.push_back(per_s);
        float *pr = nullptr, *ht = nullptr;
        if (period_samples && per_s > 0) {
            hipError_t e1 = hipMalloc((void **)&pr, sizeof(float) * (size_t)p.num_dms * nbins);
            hipError_t e2 = e1 == hipSuccess ? hipMalloc((void **)&ht, sizeof(float) * nbins) : e1;
            if (e1 != hipSuccess || e2 != hipSuccess) {
                b->prof.push_back(pr);
                b->hits.push_back(ht);
                return fail(msosm::set_error(MSOSM_ERR_ALLOC, "batch_create: profile allocation failed"));
            }
        }
        b->prof.push_back(pr);
        b->hits.push_back(ht);
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return fail(msosm::set_error(MSOSM_ERR_HIP, "batch_create: event"));
        b->join.push_back(ev);
    }
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        return fail(msosm::set_error(MSOSM_ERR_HIP, "batch_create: event"));
    if (stream) {
        b->origin = (hipStream_t)stream;
    } else {
        if (hipStreamCreate(&b->origin) != hipSuccess)
            return fail(msosm::set_error(MSOSM_ERR_HIP, "batch_create: stream"));
        b->own_origin = true;
    }
    // make sure no engine has work in flight, then capture one pass
    for (auto *e : b->engines)
        if (hipStreamSynchronize(msosm::engine_stream(e)) != hipSuccess)
            return fail(msosm::set_error(MSOSM_ERR_HIP, "batch_create: engine stream not idle"));
    std::vector<msosm::EngineCursor> before;
    for (auto *e : b->engines)
        before.push_back(msosm::engine_cursor(e));
    auto restore = [&]() {
        for (size_t i = 0; i < b->engines.size(); i++)
            msosm::engine_set_cursor(b->engines[i], before[i]);
    };
    if (hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(msosm::set_error(MSOSM_ERR_HIP, "batch_create: begin capture"));
    int rc = MSOSM_OK;
    hipError_t herr = hipEventRecord(b->fork, b->origin);
    for (int i = 0; i < n && rc == MSOSM_OK && herr == hipSuccess; i++) {
        hipStream_t s = msosm::engine_stream(b->engines[i]);
        if (s != b->origin)
            herr = hipStreamWaitEvent(s, b->fork, 0);
        if (herr != hipSuccess)
            break;
        rc = enqueue_stream(b, i);
        if (rc == MSOSM_OK && s != b->origin) {
            herr = hipEventRecord(b->join[i], s);
            if (herr == hipSuccess)
                herr = hipStreamWaitEvent(b->origin, b->join[i], 0);
        }
    }
    hipGraph_t g = nullptr;
    hipError_t cerr = hipStreamEndCapture(b->origin, &g);
    b->graph = g;
    for (auto *e : b->engines)
        b->after_pass.push_back(msosm::engine_cursor(e));
    restore();
    if (rc != MSOSM_OK)
        return fail(rc);
    if (herr != hipSuccess || cerr != hipSuccess || !g)
        return fail(msosm::set_error(MSOSM_ERR_HIP, "batch_create: capture failed (%s / %s)", hipGetErrorString(herr),
                                     hipGetErrorString(cerr)));
    if (hipGraphGetNodes(g, nullptr, &b->nodes) != hipSuccess)
        b->nodes = 0;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess)
        return fail(msosm::set_error(MSOSM_ERR_HIP, "batch_create: graph instantiate failed"));
    *out = b;
    return MSOSM_OK;
}

int msosm_batch_launch(msosm_batch *b)
{
    if (!b || !b->exec)
        return msosm::set_error(MSOSM_ERR_ARG, "batch_launch: bad batch");
    msosm::TraceRange range("msosm_batch_launch");
    BHIP(hipGraphLaunch(b->exec, b->origin));
    for (size_t i = 0; i < b->engines.size(); i++)
        msosm::engine_set_cursor(b->engines[i], b->after_pass[i]);
    return MSOSM_OK;
}

int msosm_batch_synchronize(msosm_batch *b)
{
    if (!b)
        return msosm::set_error(MSOSM_ERR_ARG, "batch_synchronize: bad batch");
    BHIP(hipStreamSynchronize(b->origin));
    return MSOSM_OK;
}

int msosm_batch_info(msosm_batch *b, int *num_streams, size_t *graph_nodes)
{
    if (!b)
        return msosm::set_error(MSOSM_ERR_ARG, "batch_info: bad batch");
    if (num_streams)
        *num_streams = (int)b->engines.size();
    if (graph_nodes)
        *graph_nodes = b->nodes;
    return MSOSM_OK;
}

int msosm_batch_profiles(msosm_batch *b, int i, float *prof, float *hits)
{
    if (!b || i < 0 || i >= (int)b->engines.size())
        return msosm::set_error(MSOSM_ERR_ARG, "batch_profiles: bad stream index");
    if (!b->prof[i])
        return msosm::set_error(MSOSM_ERR_STATE, "batch_profiles: stream %d is not folded", i);
    const msosm::Plan &p = msosm::engine_plan(b->engines[i]);
    BHIP(hipStreamSynchronize(b->origin));
    if (prof)
        BHIP(hipMemcpy(prof, b->prof[i], sizeof(float) * (size_t)p.num_dms * b->nbins, hipMemcpyDeviceToHost));
    if (hits)
        BHIP(hipMemcpy(hits, b->hits[i], sizeof(float) * (size_t)b->nbins, hipMemcpyDeviceToHost));
    return MSOSM_OK;
}

int msosm_batch_destroy(msosm_batch *b)
{
    batch_free(b);
    return MSOSM_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: measured choice of the FFT size, the form it runs with and the
// launch shapes of that form.

static int ilog2_fp(unsigned long v)
{
    int l = 0;
    while ((1ul << (l + 1)) <= v)
        l++;
    return l;
}

// The form a descriptor's form runs with at FFT size 2^log2n: the fused
// forms go up to 2^21, the matrix-core and Hankel forms up to 2^14 (above,
// the fused FP16-storage form, the same precision class); beyond 2^21 the
// reference-structured form.
static int form_at(int form, int log2n)
{
    if (form == MSOSM_FORM_PLAIN_F32)
        return form;
    const bool single = msosm::fused_supported(log2n);
    const bool big = msosm::big_supported(log2n);
    if (form == MSOSM_FORM_FUSED_F32 || form == MSOSM_FORM_FUSED_F16)
        return (single || big) ? form : MSOSM_FORM_PLAIN_F32;
    if (single && (form != MSOSM_FORM_FUSED_F16_MFMA || log2n == 14))
        return form;
    return big || single ? MSOSM_FORM_FUSED_F16 : MSOSM_FORM_PLAIN_F32;
}

// Launch-shape candidates of a form at a size (the rule first).
static std::vector<msosm_tuning> tuning_candidates(int form, int log2n, int count, int ndm)
{
    std::vector<msosm_tuning> out(1, msosm_tuning{0, 0, 0});
    const bool hankel =
        form == MSOSM_FORM_HANKEL_F16 || form == MSOSM_FORM_HANKEL_BF16 || form == MSOSM_FORM_HANKEL_BF16X3;
    if (form != MSOSM_FORM_PLAIN_F32 && log2n >= 12 && log2n <= 14) {
        for (int nb : {4, 8, 16, 32})
            if (nb <= count)
                out.push_back(msosm_tuning{nb, 0, 0});
    }
    if (hankel) {
        const int ntiles = (ndm + 15) / 16;
        for (int t : {1, 2, 4, 8})
            if (t < ntiles)
                out.push_back(msosm_tuning{0, t, 0});
    }
    if ((form == MSOSM_FORM_FUSED_F32 || form == MSOSM_FORM_FUSED_F16) && msosm::big_supported(log2n))
        out.push_back(msosm_tuning{0, 0, 64 * 1024});
    return out;
}

int msosm_autotune(const msosm_plan_desc *desc, unsigned long min_fftpoint, unsigned long max_fftpoint,
                   unsigned long target_samples, int calls, void *stream, msosm_autotune_result *result)
{
    static std::map<std::string, msosm_autotune_result> cache;
    static std::mutex cache_mu; // held for lookups and inserts, not while timing
    if (!desc || !result || calls < 1 || target_samples < 1 || desc->num_dms < 1)
        return msosm::set_error(MSOSM_ERR_ARG, "autotune: bad argument");
    std::string key;
    {
        char buf[128];
        snprintf(buf, sizeof buf, "%d|%lu|%lu|%lu|%d|%d|%d|%d|", desc->num_dms, min_fftpoint, max_fftpoint,
                 target_samples, calls, desc->method, desc->chirp, desc->form);
        key = buf;
        for (int i = 0; i < desc->num_dms; i++) {
            snprintf(buf, sizeof buf, "%a,%a,%a;", desc->bw[i], desc->dm[i], desc->f0[i]);
            key += buf;
        }
    }
    {
        std::lock_guard<std::mutex> lk(cache_mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            *result = it->second;
            return MSOSM_OK;
        }
    }
    msosm_autotune_result best{};
    best.ns_per_dm_sample = 1e300;
    uint32_t *input = nullptr;
    size_t input_words = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BHIP(hipEventCreate(&e0));
    BHIP(hipEventCreate(&e1));
    // smallest size the trials allow (a plan with fftpoint 0 takes it without
    // the below-minimum warning)
    unsigned long fp_min = 4;
    {
        msosm_plan_desc d = *desc;
        d.fftpoint = 0;
        d.input_size = 0;
        d.count = 1;
        msosm_plan *plan = nullptr;
        if (msosm_plan_create(&d, &plan) == MSOSM_OK) {
            msosm_plan_info info{};
            msosm_plan_get_info(plan, &info);
            fp_min = info.fftpoint;
            msosm_plan_destroy(plan);
        }
    }
    int timed = 0;
    for (unsigned long fp = 4; fp <= max_fftpoint && fp <= (1ul << 28); fp <<= 1) {
        if (fp < min_fftpoint || fp < fp_min)
            continue;
        msosm_plan_desc d = *desc;
        d.fftpoint = fp;
        d.input_size = 0;
        d.form = form_at(desc->form, ilog2_fp(fp));
        d.count = (int)std::max(1ul, target_samples / (fp / 2));
        msosm_plan *plan = nullptr;
        if (msosm_plan_create(&d, &plan) != MSOSM_OK)
            continue;
        msosm_plan_info info{};
        msosm_plan_get_info(plan, &info);
        if (info.clamped || info.fftpoint != fp) { // below the minimum for these trials
            msosm_plan_destroy(plan);
            continue;
        }
        msosm_engine *eng = nullptr;
        if (msosm_engine_create(plan, stream, &eng) != MSOSM_OK) {
            msosm_plan_destroy(plan);
            continue;
        }
        size_t words = (size_t)info.count * info.M;
        if (words > input_words) {
            if (input)
                (void)hipFree(input);
            input = nullptr;
            if (hipMalloc((void **)&input, words * sizeof(uint32_t)) != hipSuccess) {
                msosm_engine_destroy(eng);
                msosm_plan_destroy(plan);
                return msosm::set_error(MSOSM_ERR_ALLOC, "autotune: input allocation");
            }
            input_words = words;
            (void)msosm::launch_fill_random_u16(input, words, 0x5EED, msosm::engine_stream(eng));
        }
        hipStream_t es = msosm::engine_stream(eng);
        for (const msosm_tuning &tn : tuning_candidates(d.form, ilog2_fp(fp), info.count, info.num_dms)) {
            int rc = msosm_engine_set_tuning(eng, &tn);
            if (rc == MSOSM_OK)
                rc = msosm_engine_filter_u16(eng, input, 1); // warm-up
            (void)hipEventRecord(e0, es);
            for (int c = 0; c < calls && rc == MSOSM_OK; c++)
                rc = msosm_engine_filter_u16(eng, input, 1);
            (void)hipEventRecord(e1, es);
            float ms = 0;
            if (rc == MSOSM_OK && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
                timed++;
                const double ns = ms * 1e6 / ((double)calls * words * info.num_dms);
                if (ns < best.ns_per_dm_sample) {
                    best.ns_per_dm_sample = ns;
                    best.fftpoint = fp;
                    best.form = d.form;
                    best.tuning = tn;
                }
            }
        }
        msosm_engine_destroy(eng);
        msosm_plan_destroy(plan);
    }
    if (input)
        (void)hipFree(input);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!best.fftpoint)
        return msosm::set_error(MSOSM_ERR_UNSUPPORTED, "autotune: no FFT size in [%lu, %lu] fits these DM trials",
                                min_fftpoint, max_fftpoint);
    best.candidates = timed;
    {
        std::lock_guard<std::mutex> lk(cache_mu);
        cache[key] = best;
    }
    *result = best;
    return MSOSM_OK;
}

int msosm_autotune_fftpoint(const msosm_plan_desc *desc, unsigned long min_fftpoint, unsigned long max_fftpoint,
                            unsigned long target_samples, int calls, void *stream, unsigned long *best_fftpoint,
                            double *best_ns_per_dm_sample, int *best_form)
{
    if (!best_fftpoint)
        return msosm::set_error(MSOSM_ERR_ARG, "autotune: bad argument");
    msosm_autotune_result r{};
    int rc = msosm_autotune(desc, min_fftpoint, max_fftpoint, target_samples, calls, stream, &r);
    if (rc != MSOSM_OK)
        return rc;
    *best_fftpoint = r.fftpoint;
    if (best_ns_per_dm_sample)
        *best_ns_per_dm_sample = r.ns_per_dm_sample;
    if (best_form)
        *best_form = r.form;
    return MSOSM_OK;
}

} // extern "C"
