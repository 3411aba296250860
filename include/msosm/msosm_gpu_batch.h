// msosm_gpu_batch.h -- drop-in facade for the reference's single-DM and
// DM-loop MS-OSM engines (dedispersion/include/msosm_gpu_batch.h:7-78) over
// the C ABI of libmsosm.so.  Same class names, constructor arguments, method
// names, public members and units (bw, f0 in Hz, f0 = lowest frequency).
#pragma once
#include <cstring>
#include <vector>

#include "globals.h"

// dedispersion/include/msosm_gpu_batch.h:7-41 (+ Prepare_MSOSM members,
// dedispersion/include/prepare.h:7-49)
class MSOSM_GPU_BATCH {
public:
    MSOSM_GPU_BATCH(float bw = 0, float dm = 0, float f0 = 0) : bw_(bw), dm_(dm), f0_(f0)
    {
        if (bw > 0 && f0 > 0)
            msosm_facade::check(msosm_dm_model(bw, dm, f0, MSOSM_METHOD_MSOSM, &Nd, &M_min), "MSOSM_GPU_BATCH");
    }
    ~MSOSM_GPU_BATCH() { core_.destroy(); }
    MSOSM_GPU_BATCH(const MSOSM_GPU_BATCH &) = delete; // the reference copied raw device pointers
    MSOSM_GPU_BATCH &operator=(const MSOSM_GPU_BATCH &) = delete;

    void get_device_info() {}
    // msosm_gpu_batch.cpp:10-73
    void initialize_uint16(unsigned long fftpoint = 0, int count = 1)
    {
        core_.destroy();
        core_.create(1, &bw_, &dm_, &f0_, fftpoint, count, 0, MSOSM_METHOD_MSOSM);
        M = core_.info.M;
        M_min = core_.info.M_min;
        this->count = core_.info.count;
        int dbs = 0;
        msosm_facade::check(msosm_plan_dm_info(core_.plan, 0, &Nd, nullptr, &delaycount, &dbs), "dm_info");
        delay_points_.resize(2 * M);
        msosm_facade::check(msosm_plan_delay_points(core_.plan, 0, delay_points_.data()), "delay_points");
        delay_points = reinterpret_cast<int *>(delay_points_.data());
        params_.resize(2 * M);
        msosm_facade::check(msosm_engine_dedisp_params(core_.eng, 0, &params_[0].x), "dedisp_params");
        dedisp_params = params_.data();
        output_buffer_d = core_.output_device();
    }
    // msosm_gpu_batch.cpp:75-94 (input in host memory, as the reference)
    void filter_block_uint16(uint16_pair *input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, input, 0), "filter_block_uint16");
    }
    // Extension: input already in device memory.
    void filter_block_uint16_device(const uint16_pair *d_input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, d_input, 1), "filter_block_uint16_device");
    }
    // msosm_gpu_batch.cpp:96-99
    void get_output(Complex *output)
    {
        msosm_facade::check(msosm_engine_output_c64(core_.eng, output, 0), "get_output");
    }
    void synchronize() { msosm_facade::check(msosm_engine_synchronize(core_.eng), "synchronize"); }
    msosm_engine *engine() const { return core_.eng; }

public:
    unsigned long M = 0, M_min = 0, Nd = 0;
    int delaycount = 0;
    int *delay_points = nullptr;
    Complex *dedisp_params = nullptr; // host copy of the filter table (2M)
    int count = 0;
    Complex *output_buffer_d = nullptr; // device, count*M complex

private:
    float bw_, dm_, f0_;
    msosm_facade::Core core_;
    std::vector<int32_t> delay_points_;
    std::vector<Complex> params_;
};

// dedispersion/include/msosm_gpu_batch.h:43-78: one shared forward FFT, then
// one DM at a time.  Here input_process runs every DM trial in one engine
// call and filter_block_uint16(output, i) hands out trial i.
class MSOSM_GPU_DM_BATCH_loop {
public:
    MSOSM_GPU_DM_BATCH_loop(float *bw, float *dm, float *f0, int numOutput)
        : numOutput(numOutput), bw_(bw, bw + numOutput), dm_(dm, dm + numOutput), f0_(f0, f0 + numOutput)
    {
        Nd_.resize(numOutput);
        M_min_.resize(numOutput);
        M_.resize(numOutput);
        K_.resize(numOutput);
        for (int i = 0; i < numOutput; i++)
            msosm_facade::check(msosm_dm_model(bw[i], dm[i], f0[i], MSOSM_METHOD_MSOSM, &Nd_[i], &M_min_[i]),
                                "MSOSM_GPU_DM_BATCH_loop");
        Nd = Nd_.data();
        M_min = M_min_.data();
        M = M_.data();
        delaycount = K_.data();
    }
    ~MSOSM_GPU_DM_BATCH_loop() { core_.destroy(); }
    MSOSM_GPU_DM_BATCH_loop(const MSOSM_GPU_DM_BATCH_loop &) = delete;
    MSOSM_GPU_DM_BATCH_loop &operator=(const MSOSM_GPU_DM_BATCH_loop &) = delete;

    void get_device_info() {}
    // msosm_gpu_dm_loop.cpp:23-92
    void initialize_uint16(unsigned long fftpoint, int count = 1)
    {
        core_.destroy();
        core_.create(numOutput, bw_.data(), dm_.data(), f0_.data(), fftpoint, count, 0, MSOSM_METHOD_MSOSM);
        M_loop = core_.info.M;
        M_loop_min = core_.info.M_min;
        this->count = core_.info.count;
        for (int i = 0; i < numOutput; i++) {
            M_[i] = M_loop;
            msosm_facade::check(msosm_plan_dm_info(core_.plan, i, nullptr, nullptr, &K_[i], nullptr), "dm_info");
        }
        outs_.assign(numOutput, nullptr);
        Complex *base = core_.output_device();
        for (int i = 0; i < numOutput; i++)
            outs_[i] = base + (size_t)i * this->count * M_loop;
        output_buffer_d = outs_.data();
    }
    // msosm_gpu_dm_loop.cpp:94-104
    void input_process(uint16_pair *input)
    {
        msosm_facade::check(msosm_engine_filter_u16(core_.eng, input, 0), "input_process");
        host_.resize((size_t)numOutput * count * M_loop);
        have_ = false;
    }
    // msosm_gpu_dm_loop.cpp:106-122: output[i] <- trial i of this call
    void filter_block_uint16(Complex **output, int i)
    {
        if (!have_) {
            msosm_facade::check(msosm_engine_output_c64(core_.eng, host_.data(), 0), "filter_block_uint16");
            have_ = true;
        }
        std::memcpy(output[i], host_.data() + (size_t)i * count * M_loop, sizeof(Complex) * count * M_loop);
    }
    // msosm_gpu_dm_loop.cpp:124-142 frees the device buffers; here the engine
    // stays usable and this only waits for it.
    void synchronize() { msosm_facade::check(msosm_engine_synchronize(core_.eng), "synchronize"); }

public:
    unsigned long *M = nullptr, *M_min = nullptr;
    unsigned long M_loop = 0, M_loop_min = 0;
    int *delaycount = nullptr;
    unsigned long *Nd = nullptr;
    int numOutput;
    int count = 0;
    Complex **output_buffer_d = nullptr;

private:
    std::vector<float> bw_, dm_, f0_;
    std::vector<unsigned long> Nd_, M_min_, M_;
    std::vector<int> K_;
    std::vector<Complex *> outs_;
    std::vector<Complex> host_;
    bool have_ = false;
    msosm_facade::Core core_;
};
