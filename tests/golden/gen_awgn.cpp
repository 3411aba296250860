// gen_awgn.cpp -- writes the first complex noise samples that the reference
// drivers' AWGN generator (simulated_pulsar/AWGN.cpp:29-45: one
// std::default_random_engine with its default seed, one
// std::normal_distribution<float>(0, 1) per component, real drawn before
// imaginary, both scaled by sigma / sqrt(2), sigma = sqrt(10^(-SNR/10)))
// produces with this C++ standard library, as text: index re im.
// Build and run: g++ -O2 gen_awgn.cpp -o gen_awgn && ./gen_awgn 10 64 > awgn_snr10.txt
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char **argv)
{
    const float snr = argc > 1 ? std::strtof(argv[1], nullptr) : 10.0f;
    const unsigned long n = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 64;
    const float sigma = (float)std::sqrt(std::pow(10, (-snr / 10)));
    std::default_random_engine engine;
    std::normal_distribution<float> re(0.0f, 1.0f), im(0.0f, 1.0f);
    for (unsigned long i = 0; i < n; i++) {
        const float a = (float)(sigma / std::sqrt(2) * re(engine));
        const float b = (float)(sigma / std::sqrt(2) * im(engine));
        std::printf("%lu %.9g %.9g\n", i, a, b);
    }
    return 0;
}
