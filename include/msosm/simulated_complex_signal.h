// simulated_complex_signal.h -- drop-in facade for the reference drivers'
// test-signal generator (simulated_pulsar/include/simulated_complex_signal.h,
// simulated_complex_signal.cpp, AWGN.cpp): same class name, constructor,
// generate_pulsar_signal and public members.  The generation itself is
// msosm_refsim_generate (transforms on the GPU, noise from the C++
// standard library's default_random_engine / normal_distribution<float>).
// plot_abs (matplotlib) has no counterpart.
#pragma once
#include <string>

#include "globals.h"

#ifndef FFTW3_H
typedef float fftwf_complex[2]; // the reference's sample type for this class (fftw3.h)
#endif

class SimulatedComplexSignal {
public:
    SimulatedComplexSignal(float bw, float dm, float f0, float period, std::string mode = "complex")
        : data_type(mode), bw_(bw), dm_(dm), f0_(f0), period_(period)
    {
        msosm_facade::check(msosm_refsim_info(bw, dm, f0, period, &Np, &Nd, &range, nullptr),
                            "SimulatedComplexSignal");
        if (verbose)
            std::cout << "Period samples: " << Np << "\nDispersion samples: " << Nd << std::endl;
    }
    ~SimulatedComplexSignal()
    {
        if (pinned_) {
            msosm_host_free(signal);
            msosm_host_free(signal_u16);
        } else {
            std::free(signal);
            std::free(signal_u16);
        }
    }
    SimulatedComplexSignal(const SimulatedComplexSignal &) = delete;
    SimulatedComplexSignal &operator=(const SimulatedComplexSignal &) = delete;

    // simulated_complex_signal.cpp:40-236
    void generate_pulsar_signal(unsigned long repeat = 1, bool add_noise = false, float SNR = 0, bool pinned = true)
    {
        signal_size = repeat * Np;
        pinned_ = pinned;
        signal = static_cast<fftwf_complex *>(alloc(sizeof(fftwf_complex) * signal_size));
        if (data_type == "uint16")
            signal_u16 = static_cast<uint16_pair *>(alloc(sizeof(uint16_pair) * signal_size));
        float power = 0;
        msosm_facade::check(msosm_refsim_generate(bw_, dm_, f0_, period_, repeat, add_noise ? 1 : 0, SNR,
                                                  &signal[0][0], signal_u16, &power),
                            "generate_pulsar_signal");
        if (verbose)
            std::cout << "Signal measured power: " << power << " dB" << std::endl;
    }

public:
    fftwf_complex *signal = nullptr;
    uint16_pair *signal_u16 = nullptr;
    std::string data_type;
    unsigned long signal_size = 0;
    unsigned long Np = 0;
    unsigned long Nd = 0;
    unsigned long M = 0;
    unsigned long M_min = 0;
    unsigned long range = 0;

private:
    void *alloc(size_t bytes)
    {
        void *p = nullptr;
        if (pinned_)
            msosm_facade::check(msosm_host_alloc(bytes, &p), "generate_pulsar_signal");
        else
            p = std::malloc(bytes);
        return p;
    }
    float bw_, dm_, f0_, period_;
    bool pinned_ = false;
};
