// outring.hip -- output streaming from the GPU engines to a host consumer.
//
// The reference hands each call's output to a writer thread through one host
// buffer, a sentinel element after the data that the consumer sets to 1 and
// the device-to-host copy overwrites with 0, and a plain bool the producer
// spins on (msosm_gpu_dm_concurrent.cpp:153-195,
// multiple_dm_concurrent_msosm.cpp:84-146).  That relies on the copy landing
// its last element last and on unsynchronised flags.  Here: a ring of slots,
// each a device staging buffer plus a page-locked host buffer.
//
//   push (producer thread): on the engine's stream, optionally pack to uint16,
//     then copy the output into the slot's device staging buffer (so the
//     engine may overwrite its output with the next call at once); on the
//     ring's copy stream, after that, copy staging -> host and record the
//     slot's completion event.  Blocks while every slot is queued or held.
//   pop (consumer thread): wait for the oldest queued slot's event; return
//     its host data and sequence number.
//   release (consumer): the popped slot is free again.
//
// One producer and one consumer thread; slots are handed over in push order,
// each exactly once.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <vector>

#include "common.h"
#include "engine_internal.h"
#include "kernels.h"
#include "msosm_c.h"
#include "trace.h"

namespace {

struct Slot {
    void *dev = nullptr;
    void *host = nullptr;
    size_t bytes = 0; // valid bytes of the current content
    long long seq = -1;
    hipEvent_t staged = nullptr; // staging copy done (engine stream)
    hipEvent_t ready = nullptr;  // host copy done (copy stream)
};

} // namespace

struct msosm_outring {
    size_t slot_bytes = 0;
    std::vector<Slot> slots;
    hipStream_t copy_stream = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> free_slots; // available to the producer
    std::deque<int> queued;     // pushed, not yet popped (push order)
    int held = -1;              // popped, not yet released
    long long next_seq = 0;
};

namespace {

void ring_free(msosm_outring *r)
{
    if (!r)
        return;
    if (r->copy_stream)
        (void)hipStreamSynchronize(r->copy_stream);
    for (Slot &s : r->slots) {
        if (s.dev)
            (void)hipFree(s.dev);
        if (s.host)
            (void)hipHostFree(s.host);
        if (s.staged)
            (void)hipEventDestroy(s.staged);
        if (s.ready)
            (void)hipEventDestroy(s.ready);
    }
    if (r->copy_stream)
        (void)hipStreamDestroy(r->copy_stream);
    delete r;
}

// wait on cv until pred or timeout (timeout_ms < 0: forever); lock held
template <typename Pred>
bool wait_for(msosm_outring *r, std::unique_lock<std::mutex> &lk, int timeout_ms, Pred pred)
{
    if (timeout_ms < 0) {
        r->cv.wait(lk, pred);
        return true;
    }
    return r->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
}

} // namespace

extern "C" {

int msosm_outring_create(size_t slot_bytes, int slots, msosm_outring **out)
{
    if (!out || slot_bytes == 0 || slots < 1 || slots > 1024)
        return msosm::set_error(MSOSM_ERR_ARG, "outring_create: need slot_bytes > 0 and 1 <= slots <= 1024");
    *out = nullptr;
    msosm_outring *r = new msosm_outring();
    r->slot_bytes = slot_bytes;
    r->slots.resize(slots);
    auto fail = [&](hipError_t e, const char *what) {
        ring_free(r);
        return msosm::set_error(MSOSM_ERR_ALLOC, "outring_create: %s: %s", what, hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&r->copy_stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "stream");
    for (int i = 0; i < slots; i++) {
        Slot &s = r->slots[i];
        if ((e = hipMalloc(&s.dev, slot_bytes)) != hipSuccess)
            return fail(e, "device staging");
        if ((e = hipHostMalloc(&s.host, slot_bytes, hipHostMallocDefault)) != hipSuccess)
            return fail(e, "host buffer");
        if ((e = hipEventCreateWithFlags(&s.staged, hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&s.ready, hipEventDisableTiming)) != hipSuccess)
            return fail(e, "events");
        r->free_slots.push_back(i);
    }
    *out = r;
    return MSOSM_OK;
}

int msosm_outring_destroy(msosm_outring *r)
{
    ring_free(r);
    return MSOSM_OK;
}

int msosm_outring_push(msosm_outring *r, msosm_engine *e, int format, int timeout_ms, long long *seq)
{
    if (!r || !e || (format != 0 && format != 1))
        return msosm::set_error(MSOSM_ERR_ARG, "outring_push: bad argument (format 0 complex64, 1 uint16)");
    msosm::TraceRange range("msosm_outring_push");
    const msosm::Plan &p = msosm::engine_plan(e);
    const size_t n = (size_t)p.num_dms * p.count * p.M;
    const size_t bytes = n * (format == 0 ? sizeof(float2) : sizeof(uint32_t));
    if (bytes > r->slot_bytes)
        return msosm::set_error(MSOSM_ERR_ARG, "outring_push: output of %zu bytes exceeds the slot size %zu", bytes,
                                r->slot_bytes);
    int idx;
    {
        std::unique_lock<std::mutex> lk(r->mu);
        if (!wait_for(r, lk, timeout_ms, [&] { return !r->free_slots.empty(); }))
            return msosm::set_error(MSOSM_ERR_STATE, "outring_push: no free slot within %d ms", timeout_ms);
        idx = r->free_slots.front();
        r->free_slots.pop_front();
    }
    Slot &s = r->slots[idx];
    hipStream_t es = msosm::engine_stream(e);
    const float2 *out = msosm::engine_output(e);
    // a failure after the slot was taken hands it back (at the front, so the
    // next push reuses it) instead of shrinking the ring for good
    hipError_t herr = hipSuccess;
    const char *what = nullptr;
    auto step = [&](hipError_t x, const char *name) {
        if (herr == hipSuccess && x != hipSuccess) {
            herr = x;
            what = name;
        }
        return herr == hipSuccess;
    };
    if (format == 0)
        step(hipMemcpyAsync(s.dev, out, bytes, hipMemcpyDeviceToDevice, es), "staging copy");
    else
        step(msosm::launch_pack(static_cast<uint32_t *>(s.dev), out, n, es), "uint16 pack");
    if (herr == hipSuccess)
        step(hipEventRecord(s.staged, es), "staged event");
    if (herr == hipSuccess)
        step(hipStreamWaitEvent(r->copy_stream, s.staged, 0), "copy-stream wait");
    if (herr == hipSuccess)
        step(hipMemcpyAsync(s.host, s.dev, bytes, hipMemcpyDeviceToHost, r->copy_stream), "host copy");
    if (herr == hipSuccess)
        step(hipEventRecord(s.ready, r->copy_stream), "ready event");
    if (herr != hipSuccess) {
        {
            std::lock_guard<std::mutex> lk(r->mu);
            r->free_slots.push_front(idx);
        }
        r->cv.notify_all();
        return msosm::set_error(MSOSM_ERR_HIP, "outring_push: %s failed: %s", what, hipGetErrorString(herr));
    }
    {
        std::lock_guard<std::mutex> lk(r->mu);
        s.bytes = bytes;
        s.seq = r->next_seq++;
        if (seq)
            *seq = s.seq;
        r->queued.push_back(idx);
    }
    r->cv.notify_all();
    return MSOSM_OK;
}

int msosm_outring_pop(msosm_outring *r, int timeout_ms, const void **data, size_t *bytes, long long *seq)
{
    if (!r || !data)
        return msosm::set_error(MSOSM_ERR_ARG, "outring_pop: bad argument");
    int idx;
    {
        std::unique_lock<std::mutex> lk(r->mu);
        if (r->held >= 0)
            return msosm::set_error(MSOSM_ERR_STATE, "outring_pop: release the previous slot first");
        if (!wait_for(r, lk, timeout_ms, [&] { return !r->queued.empty(); }))
            return msosm::set_error(MSOSM_ERR_STATE, "outring_pop: nothing queued within %d ms", timeout_ms);
        idx = r->queued.front();
        r->queued.pop_front();
        r->held = idx;
    }
    Slot &s = r->slots[idx];
    hipError_t herr = hipEventSynchronize(s.ready);
    if (herr != hipSuccess) {
        // the slot's content is not valid: free it rather than hand it out
        {
            std::lock_guard<std::mutex> lk(r->mu);
            r->held = -1;
            r->free_slots.push_back(idx);
        }
        r->cv.notify_all();
        return msosm::set_error(MSOSM_ERR_HIP, "outring_pop: waiting for slot %d: %s", idx, hipGetErrorString(herr));
    }
    *data = s.host;
    if (bytes)
        *bytes = s.bytes;
    if (seq)
        *seq = s.seq;
    return MSOSM_OK;
}

int msosm_outring_release(msosm_outring *r)
{
    if (!r)
        return msosm::set_error(MSOSM_ERR_ARG, "outring_release: NULL ring");
    {
        std::lock_guard<std::mutex> lk(r->mu);
        if (r->held < 0)
            return msosm::set_error(MSOSM_ERR_STATE, "outring_release: no slot held");
        r->free_slots.push_back(r->held);
        r->held = -1;
    }
    r->cv.notify_all();
    return MSOSM_OK;
}

} // extern "C"
