// osm_gpu.h -- drop-in facade for the single-segment overlap-save engines
// (dedispersion/include/osm_gpu_batch.h:7-37, osm_gpu_concurrent.h:7-42).
#pragma once
#include <memory>
#include <vector>

#include "globals.h"

class OSM_GPU_BATCH {
public:
    OSM_GPU_BATCH(float bw = 0, float dm = 0, float f0 = 0) : bw_(bw), dm_(dm), f0_(f0)
    {
        if (bw > 0 && f0 > 0)
            msosm_facade::check(msosm_dm_model(bw, dm, f0, MSOSM_METHOD_OSM, &Nd, &M_min), "OSM_GPU_BATCH");
    }
    ~OSM_GPU_BATCH() { core_.destroy(); }
    OSM_GPU_BATCH(const OSM_GPU_BATCH &) = delete;
    OSM_GPU_BATCH &operator=(const OSM_GPU_BATCH &) = delete;

    void get_device_info() {}
    // osm_gpu_batch.cpp:19-92 (streaming only chose a CUDA stream there; the
    // engine always runs on its own stream)
    void initialize_uint16(unsigned long fftpoint = 0, int count = 1, bool streaming = false)
    {
        (void)streaming;
        core_.destroy();
        core_.create(1, &bw_, &dm_, &f0_, fftpoint, count, 0, MSOSM_METHOD_OSM);
        M = core_.info.M;
        M_min = core_.info.M_min;
        this->count = core_.info.count;
        output_buffer_d = core_.output_device();
        const unsigned long N = 2 * M;
        // device bytes held by the engine (the reference sums its buffers and
        // cuFFT work areas, osm_gpu_batch.cpp:46-91)
        total_memory = sizeof(Complex) * ((count + 1) * M + 3 * count * N + count * M + N) +
                       sizeof(uint16_pair) * (2 * count * M);
    }
    void filter_block_uint16(uint16_pair *input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, input, 0), "filter_block_uint16");
    }
    void get_output(Complex *output)
    {
        msosm_facade::check(msosm_engine_output_c64(core_.eng, output, 0), "get_output");
    }
    void synchronize() { msosm_facade::check(msosm_engine_synchronize(core_.eng), "synchronize"); }

public:
    unsigned long M = 0, M_min = 0, Nd = 0;
    int count = 0;
    unsigned long total_memory = 0;
    Complex *output_buffer_d = nullptr;

private:
    float bw_, dm_, f0_;
    msosm_facade::Core core_;
};

class OSM_GPU_DM_concurrent {
public:
    OSM_GPU_DM_concurrent(float *bw, float *dm, float *f0, int numDMs)
        : numDMs(numDMs), bw_(bw, bw + numDMs), dm_(dm, dm + numDMs), f0_(f0, f0 + numDMs)
    {
        wait_for_cpu = false;
    }
    ~OSM_GPU_DM_concurrent() { core_.destroy(); }
    OSM_GPU_DM_concurrent(const OSM_GPU_DM_concurrent &) = delete;
    OSM_GPU_DM_concurrent &operator=(const OSM_GPU_DM_concurrent &) = delete;

    void get_device_info() {}
    // osm_gpu_dm_concurrent.cpp:20-79
    void initialize_uint16(unsigned long fftpoint = 0, int count = 1, bool compute_only = false)
    {
        core_.destroy();
        core_.create(numDMs, bw_.data(), dm_.data(), f0_.data(), fftpoint, count, 0, MSOSM_METHOD_OSM);
        M_common = core_.info.M;
        M_common_min = core_.info.M_min;
        this->count = core_.info.count;
        this->compute_only = compute_only;
    }
    void filter_block_uint16(uint16_pair *input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, input, 0), "filter_block_uint16");
    }
    void get_output(Complex *output)
    {
        acquire();
        msosm_facade::check(msosm_engine_output_c64(core_.eng, output, 0), "get_output");
        output[(size_t)count * M_common * numDMs] = Complex{0.0f, 0.0f};
    }
    void get_output(uint16_pair *output)
    {
        acquire();
        msosm_facade::check(msosm_engine_output_u16(core_.eng, output, 0), "get_output");
        output[(size_t)count * M_common * numDMs] = uint16_pair{0, 0};
    }
    void synchronize() { msosm_facade::check(msosm_engine_synchronize(core_.eng), "synchronize"); }
    void reset_device() { msosm_facade::check(msosm_engine_reset(core_.eng), "reset_device"); }

public:
    int count = 0;
    std::atomic<bool> wait_for_cpu;
    unsigned long M_common = 0, M_common_min = 0;
    int numDMs;

private:
    void acquire()
    {
        if (!compute_only) {
            bool expect = false;
            while (!wait_for_cpu.compare_exchange_weak(expect, true))
                expect = false;
        }
    }
    bool compute_only = false;
    std::vector<float> bw_, dm_, f0_;
    msosm_facade::Core core_;
};
