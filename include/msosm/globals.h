// globals.h -- shared types of the drop-in C++ facade
// (utils/include/globals.h:21-29 of the reference).
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "../msosm_c.h"

// Packed offset-binary I/Q sample (globals.h:21-25).
struct __attribute__((packed)) uint16_pair {
    uint16_t first;
    uint16_t second;
};

// Complex sample, layout of cufftComplex / float2 (globals.h:28).
struct Complex {
    float x, y;
};
static_assert(sizeof(Complex) == 8, "Complex must be two packed floats");
static_assert(sizeof(uint16_pair) == 4, "uint16_pair must be 4 bytes");

constexpr auto pi = M_PI; // globals.h:29
inline bool verbose = false; // globals.cpp:3

// Named profiler ranges (the reference's NVTX PUSH_RANGE / POP_RANGE,
// gpu/include/gpu_common.cuh:16-39): rocTX ranges on the calling thread,
// shown by rocprofv3 --marker-trace.  The colour id has no rocTX counterpart.
#ifndef PUSH_RANGE
#define PUSH_RANGE(name, cid)                                                                                    \
    {                                                                                                            \
        (void)(cid);                                                                                             \
        (void)msosm_range_push(name);                                                                            \
    }
#define POP_RANGE (void)msosm_range_pop();
#endif

namespace msosm_facade {

// The reference prints and exits on any CUDA/cuFFT failure
// (gpu/include/gpu_common.cuh:41-61); the facade does the same with the C ABI
// status.
inline void check(int rc, const char *what)
{
    if (rc != MSOSM_OK) {
        std::cerr << "msosm error " << rc << " in " << what << ": " << msosm_last_error() << std::endl;
        std::exit(1);
    }
}

// Owning wrapper of a plan + engine pair built from per-DM arrays.
struct Core {
    msosm_plan *plan = nullptr;
    msosm_engine *eng = nullptr;
    msosm_plan_info info{};

    void create(int n, const float *bw, const float *dm, const float *f0, unsigned long fftpoint, int count,
                unsigned long input_size, int method, int form = MSOSM_FORM_FUSED_F32)
    {
        msosm_plan_desc d{};
        d.num_dms = n;
        d.bw = bw;
        d.dm = dm;
        d.f0 = f0;
        d.fftpoint = fftpoint;
        d.count = count;
        d.input_size = input_size;
        d.method = method;
        d.chirp = MSOSM_CHIRP_REF_F32W;
        d.form = form;
        check(msosm_plan_create(&d, &plan), "plan");
        check(msosm_plan_get_info(plan, &info), "plan info");
        int rc = MSOSM_ERR_UNSUPPORTED;
        // the fused kernels take FFT sizes up to 2^21, keep delays in 16 bits
        // and address their spectra ring with 32-bit offsets; plans beyond
        // that use the reference-structured form
        if (d.form == MSOSM_FORM_PLAIN_F32 || info.fftpoint <= 2097152)
            rc = msosm_engine_create(plan, nullptr, &eng);
        if (rc == MSOSM_ERR_UNSUPPORTED && d.form != MSOSM_FORM_PLAIN_F32) {
            msosm_plan_destroy(plan);
            d.form = MSOSM_FORM_PLAIN_F32;
            check(msosm_plan_create(&d, &plan), "plan");
            check(msosm_plan_get_info(plan, &info), "plan info");
            rc = msosm_engine_create(plan, nullptr, &eng);
        }
        check(rc, "engine");
    }
    void destroy()
    {
        if (eng)
            msosm_engine_destroy(eng);
        if (plan)
            msosm_plan_destroy(plan);
        eng = nullptr;
        plan = nullptr;
    }
    Complex *output_device() const
    {
        void *p = nullptr;
        check(msosm_engine_output_ptr(eng, &p), "output_ptr");
        return static_cast<Complex *>(p);
    }
};

} // namespace msosm_facade
