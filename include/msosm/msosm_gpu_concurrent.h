// msosm_gpu_concurrent.h -- drop-in facade for MSOSM_GPU_DM_concurrent
// (dedispersion/include/msosm_gpu_concurrent.h:15-64) and the two
// multi-stream engines (dedispersion/include/msosm_gpu_batch_stream.h:14-93).
//
// Output handshake: the reference copies count*M*numDMs + 1 elements, the
// last being a sentinel a consumer thread spins on, and guards its host
// buffer with plain bool flags (msosm_gpu_dm_concurrent.cpp:153-195).  The
// facade keeps the same members and the sentinel and makes the flags atomic.
// get_output returns once the copy is complete.  get_output_thread (one
// writer thread per buffer, multiple_dm_concurrent_msosm.cpp:117-146) hands
// the output over asynchronously through the library's output ring
// (msosm_outring_*): the call pushes this call's output into a ring slot (a
// copy on the engine's stream, then to page-locked memory on the ring's
// stream) and returns, so the engine's next call can start at once; a
// delivery thread takes the slots in push order, copies each into the
// buffer given with it and only then clears that buffer's sentinel (a
// release store), so a writer never sees the sentinel before the data.
#pragma once
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "globals.h"

class MSOSM_GPU_DM_concurrent {
public:
    MSOSM_GPU_DM_concurrent(float *bw, float *dm, float *f0, int numDMs)
        : numDMs(numDMs), bw_(bw, bw + numDMs), dm_(dm, dm + numDMs), f0_(f0, f0 + numDMs)
    {
        Nd_.resize(numDMs);
        M_min_.resize(numDMs);
        K_.resize(numDMs);
        for (int i = 0; i < numDMs; i++) {
            msosm_facade::check(msosm_dm_model(bw[i], dm[i], f0[i], MSOSM_METHOD_MSOSM, &Nd_[i], &M_min_[i]),
                                "MSOSM_GPU_DM_concurrent");
        }
        Nd = Nd_.data();
        M_min = M_min_.data();
        M = M_min_.data();
        delaycount = K_.data();
        wait_for_cpu = false;
    }
    ~MSOSM_GPU_DM_concurrent()
    {
        stop_delivery();
        core_.destroy();
    }
    MSOSM_GPU_DM_concurrent(const MSOSM_GPU_DM_concurrent &) = delete;
    MSOSM_GPU_DM_concurrent &operator=(const MSOSM_GPU_DM_concurrent &) = delete;

    void get_device_info() {}
    // msosm_gpu_dm_concurrent.cpp:22-108
    void initialize_uint16(unsigned long fftpoint, int count = 1, unsigned long input_size = 0,
                           bool compute_only = false)
    {
        stop_delivery();
        core_.destroy();
        core_.create(numDMs, bw_.data(), dm_.data(), f0_.data(), fftpoint, count, input_size, MSOSM_METHOD_MSOSM);
        M_common = core_.info.M;
        M_common_min = core_.info.M_min;
        this->count = core_.info.count;
        this->compute_only = compute_only;
        for (int i = 0; i < numDMs; i++)
            msosm_facade::check(msosm_plan_dm_info(core_.plan, i, nullptr, nullptr, &K_[i], nullptr), "dm_info");
    }
    // msosm_gpu_dm_concurrent.cpp:110-117
    void initialize_threads_flag(int thread_num)
    {
        nthreads_ = thread_num;
        wait_for_threads_.reset(new std::atomic<bool>[thread_num]);
        for (int i = 0; i < thread_num; i++)
            wait_for_threads_[i] = false;
        wait_for_threads = wait_for_threads_.get();
    }
    // msosm_gpu_dm_concurrent.cpp:119-151 (host input)
    void filter_block_uint16(uint16_pair *input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, input, 0), "filter_block_uint16");
    }
    void filter_block_uint16_device(const uint16_pair *d_input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, d_input, 1), "filter_block_uint16_device");
    }
    // msosm_gpu_dm_concurrent.cpp:153-165
    void get_output(Complex *output)
    {
        acquire(wait_for_cpu);
        msosm_facade::check(msosm_engine_output_c64(core_.eng, output, 0), "get_output");
        output[total()] = Complex{0.0f, 0.0f};
    }
    // msosm_gpu_dm_concurrent.cpp:167-180
    void get_output(uint16_pair *output)
    {
        acquire(wait_for_cpu);
        msosm_facade::check(msosm_engine_output_u16(core_.eng, output, 0), "get_output");
        output[total()] = uint16_pair{0, 0};
    }
    // msosm_gpu_dm_concurrent.cpp:182-195: asynchronous hand-over through
    // the output ring (see the header comment); the sentinel output[total]
    // reads 0 once the data is in `output`
    void get_output_thread(uint16_pair *output, int thread_id)
    {
        acquire(wait_for_threads[thread_id]);
        start_delivery();
        {
            std::lock_guard<std::mutex> lk(dmu_);
            dest_.push_back(output); // before the push: the delivery thread pops in push order
            pending_++;
        }
        long long seq = 0;
        const int rc = msosm_outring_push(ring_, core_.eng, 1, -1, &seq);
        if (rc != MSOSM_OK) {
            // nothing was queued: take the destination back (this thread is
            // the only producer, so it is the last entry)
            std::lock_guard<std::mutex> lk(dmu_);
            dest_.pop_back();
            pending_--;
            wait_for_threads[thread_id] = false;
        }
        msosm_facade::check(rc, "get_output_thread");
    }
    // waits for the engine and for every output handed to writer threads to
    // be in its buffer (the reference's cudaDeviceSynchronize covers both)
    void synchronize()
    {
        msosm_facade::check(msosm_engine_synchronize(core_.eng), "synchronize");
        std::unique_lock<std::mutex> lk(dmu_);
        dcv_.wait(lk, [&] { return pending_ == 0 || delivery_error_; });
        if (delivery_error_)
            msosm_facade::check(MSOSM_ERR_HIP, "output delivery");
    }
    // The reference calls cudaDeviceReset (msosm_gpu_dm_concurrent.cpp:202-205),
    // which tears down every allocation of the process; here only this
    // engine's stream history is cleared.
    void reset_device() { msosm_facade::check(msosm_engine_reset(core_.eng), "reset_device"); }
    Complex *output_device() const { return core_.output_device(); }
    msosm_engine *engine() const { return core_.eng; }

public:
    int count = 0;
    std::atomic<bool> wait_for_cpu;
    std::atomic<bool> *wait_for_threads = nullptr;
    unsigned long M_common = 0, M_common_min = 0;
    unsigned long *M = nullptr, *M_min = nullptr, *Nd = nullptr;
    int *delaycount = nullptr;
    int numDMs;

private:
    size_t total() const { return (size_t)count * M_common * numDMs; }
    void acquire(std::atomic<bool> &flag)
    {
        if (!compute_only) {
            bool expect = false;
            while (!flag.compare_exchange_weak(expect, true))
                expect = false;
        }
    }
    // ring of (writer threads + 1) slots and the thread that empties it
    void start_delivery()
    {
        if (ring_)
            return;
        int threads = 1;
        if (wait_for_threads_)
            threads = nthreads_;
        msosm_facade::check(msosm_outring_create(sizeof(uint16_pair) * total(), threads + 1, &ring_),
                            "output ring");
        stop_ = false;
        delivery_ = std::thread([this] { deliver(); });
    }
    void deliver()
    {
        const size_t n = total();
        for (;;) {
            const void *data = nullptr;
            size_t bytes = 0;
            long long seq = 0;
            const int rc = msosm_outring_pop(ring_, 50, &data, &bytes, &seq);
            if (rc != MSOSM_OK) {
                std::lock_guard<std::mutex> lk(dmu_);
                if (rc != MSOSM_ERR_STATE) { // not a time-out
                    delivery_error_ = true;
                    dcv_.notify_all();
                    return;
                }
                if (stop_ && pending_ == 0)
                    return;
                continue;
            }
            uint16_pair *dst = nullptr;
            {
                std::lock_guard<std::mutex> lk(dmu_);
                dst = dest_.front();
                dest_.pop_front();
            }
            std::memcpy(dst, data, bytes);
            __atomic_store_n(&dst[n].second, (uint16_t)0, __ATOMIC_RELAXED);
            __atomic_store_n(&dst[n].first, (uint16_t)0, __ATOMIC_RELEASE); // after the data
            msosm_outring_release(ring_);
            std::lock_guard<std::mutex> lk(dmu_);
            pending_--;
            dcv_.notify_all();
        }
    }
    void stop_delivery()
    {
        if (!ring_)
            return;
        {
            std::unique_lock<std::mutex> lk(dmu_);
            dcv_.wait(lk, [&] { return pending_ == 0 || delivery_error_; });
            stop_ = true;
        }
        delivery_.join();
        msosm_outring_destroy(ring_);
        ring_ = nullptr;
        dest_.clear();
        pending_ = 0;
        delivery_error_ = false;
    }
    bool compute_only = false;
    std::vector<float> bw_, dm_, f0_;
    std::vector<unsigned long> Nd_, M_min_;
    std::vector<int> K_;
    std::unique_ptr<std::atomic<bool>[]> wait_for_threads_;
    int nthreads_ = 0;
    msosm_facade::Core core_;
    msosm_outring *ring_ = nullptr;
    std::thread delivery_;
    std::mutex dmu_;
    std::condition_variable dcv_;
    std::deque<uint16_pair *> dest_;
    long long pending_ = 0;
    bool stop_ = false, delivery_error_ = false;
};

// dedispersion/include/msosm_gpu_batch_stream.h:52-93: many DM trials of one
// input, outputs handed out per trial with a per-trial flag.
class MSOSM_GPU_DM_BATCH_stream {
public:
    MSOSM_GPU_DM_BATCH_stream(float *bw, float *dm, float *f0, int numStreams)
        : numStreams(numStreams), bw_(bw, bw + numStreams), dm_(dm, dm + numStreams), f0_(f0, f0 + numStreams),
          flags_(new std::atomic<bool>[numStreams])
    {
        for (int i = 0; i < numStreams; i++)
            flags_[i] = false;
        wait_for_cpu = flags_.get();
    }
    ~MSOSM_GPU_DM_BATCH_stream() { core_.destroy(); }
    MSOSM_GPU_DM_BATCH_stream(const MSOSM_GPU_DM_BATCH_stream &) = delete;
    MSOSM_GPU_DM_BATCH_stream &operator=(const MSOSM_GPU_DM_BATCH_stream &) = delete;

    void get_device_info() {}
    void initialize_uint16(unsigned long fftpoint, int count = 1)
    {
        core_.destroy();
        core_.create(numStreams, bw_.data(), dm_.data(), f0_.data(), fftpoint, count, 0, MSOSM_METHOD_MSOSM);
        M_stream = core_.info.M;
        M_stream_min = core_.info.M_min;
        this->count = core_.info.count;
    }
    void filter_block_uint16(uint16_pair *input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, input, 0), "filter_block_uint16");
        have_c_ = have_u_ = false;
    }
    // msosm_gpu_dm_stream.cpp:160-181: output[i] gets count*M_stream + 1
    void get_output(Complex **output)
    {
        const size_t per = (size_t)count * M_stream;
        if (!have_c_) {
            host_c_.resize(per * numStreams);
            msosm_facade::check(msosm_engine_output_c64(core_.eng, host_c_.data(), 0), "get_output");
            have_c_ = true;
        }
        for (int i = 0; i < numStreams; i++) {
            bool expect = false;
            while (!wait_for_cpu[i].compare_exchange_weak(expect, true))
                expect = false;
            std::memcpy(output[i], host_c_.data() + per * i, sizeof(Complex) * per);
            output[i][per] = Complex{0.0f, 0.0f};
        }
    }
    void get_output(uint16_pair **output)
    {
        const size_t per = (size_t)count * M_stream;
        if (!have_u_) {
            host_u_.resize(per * numStreams);
            msosm_facade::check(msosm_engine_output_u16(core_.eng, host_u_.data(), 0), "get_output");
            have_u_ = true;
        }
        for (int i = 0; i < numStreams; i++) {
            bool expect = false;
            while (!wait_for_cpu[i].compare_exchange_weak(expect, true))
                expect = false;
            std::memcpy(output[i], host_u_.data() + per * i, sizeof(uint16_pair) * per);
            output[i][per] = uint16_pair{0, 0};
        }
    }
    void synchronize() { msosm_facade::check(msosm_engine_synchronize(core_.eng), "synchronize"); }
    void reset_device() { msosm_facade::check(msosm_engine_reset(core_.eng), "reset_device"); }

public:
    int count = 0;
    std::atomic<bool> *wait_for_cpu = nullptr;
    unsigned long M_stream = 0, M_stream_min = 0;
    int numStreams;

private:
    std::vector<float> bw_, dm_, f0_;
    std::unique_ptr<std::atomic<bool>[]> flags_;
    std::vector<Complex> host_c_;
    std::vector<uint16_pair> host_u_;
    bool have_c_ = false, have_u_ = false;
    msosm_facade::Core core_;
};

// dedispersion/include/msosm_gpu_batch_stream.h:14-49: independent input
// streams, each with its own FFT size; one engine per stream.
class MSOSM_GPU_BATCH_stream {
public:
    MSOSM_GPU_BATCH_stream(float *bw, float *dm, float *f0, int numStreams)
        : numStreams(numStreams), bw_(bw, bw + numStreams), dm_(dm, dm + numStreams), f0_(f0, f0 + numStreams),
          cores_(numStreams), outs_(numStreams, nullptr), M_(numStreams, 0)
    {
        M = M_.data();
    }
    ~MSOSM_GPU_BATCH_stream()
    {
        for (auto &c : cores_)
            c.destroy();
    }
    MSOSM_GPU_BATCH_stream(const MSOSM_GPU_BATCH_stream &) = delete;
    MSOSM_GPU_BATCH_stream &operator=(const MSOSM_GPU_BATCH_stream &) = delete;

    void get_device_info() {}
    // msosm_gpu_batch_stream.cpp:33-103; fftpoint[i] is updated to 2*M[i]
    void initialize_uint16(unsigned long *fftpoint, int count = 1)
    {
        this->count = count;
        for (int i = 0; i < numStreams; i++) {
            cores_[i].destroy();
            cores_[i].create(1, &bw_[i], &dm_[i], &f0_[i], fftpoint[i], count, 0, MSOSM_METHOD_MSOSM);
            M_[i] = cores_[i].info.M;
            fftpoint[i] = 2 * M_[i];
            outs_[i] = cores_[i].output_device();
        }
        output_buffer_d = outs_.data();
    }
    void filter_block_uint16(uint16_pair **input)
    {
        for (int i = 0; i < numStreams; i++)
            msosm_facade::check(msosm_engine_filter_u16(cores_[i].eng, input[i], 0), "filter_block_uint16");
    }
    void get_output(Complex **output)
    {
        for (int i = 0; i < numStreams; i++)
            msosm_facade::check(msosm_engine_output_c64(cores_[i].eng, output[i], 0), "get_output");
    }
    void synchronize()
    {
        for (auto &c : cores_)
            msosm_facade::check(msosm_engine_synchronize(c.eng), "synchronize");
    }

public:
    int count = 0;
    Complex **output_buffer_d = nullptr;
    unsigned long *M = nullptr;
    int numStreams;

private:
    std::vector<float> bw_, dm_, f0_;
    std::vector<msosm_facade::Core> cores_;
    std::vector<Complex *> outs_;
    std::vector<unsigned long> M_;
};
