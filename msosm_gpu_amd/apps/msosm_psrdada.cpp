// msosm_psrdada -- coherent dedispersion and folding of a dual-polarisation
// PSRDADA recording, end to end on the GPU: the drop-in for the reference's
// msosm_psrdada driver (msosm_psrdada.cpp:92-143), written against the same
// C++ classes (PSRDADA, MSOSM_GPU_BATCH, Fold_GPU).
//
// Per block of count*M samples: read pol1 and pol2 frames, dedisperse each
// polarisation with its own engine, intensity sqrt(|p1|^2 + |p2|^2) on the
// GPU, fold with the reference's sample-exact fold (the first Nd samples,
// the filter transient, are discarded).  Leading blocks whose first pol1
// sample is the "no data" code 0 are skipped, as the reference does.  The
// folded profile (time_bin values) is written to <file stem>_dedisp_folded.txt.
//
// Options: --file F  --dm DM  --period P (s)  [--verbose]
//          [--fftpoint N (default 2097152)] [--count C (default 1)]
//          [--time-bin B (default 1024)] [--out PATH] [--header]
#include <chrono>
#include <cstring>
#include <iostream>
#include <string>

#include "msosm/fold_gpu.h"
#include "msosm/msosm_gpu_batch.h"
#include "msosm/psrdada.h"

namespace {

void usage()
{
    std::cout << "usage: msosm_psrdada --file F --dm DM --period P [--fftpoint N] [--count C] [--time-bin B] "
                 "[--out PATH] [--header] [--verbose]"
              << std::endl;
}

} // namespace

int main(int argc, char **argv)
{
    std::string file, out;
    float dm = 750.0f, period = 0.0f;
    unsigned long fftpoint = 2097152, time_bin = 1024;
    int count = 1;
    bool header_only = false;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--file")
            file = next();
        else if (a == "--dm")
            dm = std::stof(next());
        else if (a == "--period")
            period = std::stof(next());
        else if (a == "--fftpoint")
            fftpoint = std::stoul(next());
        else if (a == "--count")
            count = std::stoi(next());
        else if (a == "--time-bin")
            time_bin = std::stoul(next());
        else if (a == "--out")
            out = next();
        else if (a == "--header")
            header_only = true;
        else if (a == "--verbose")
            verbose = true;
        else {
            usage();
            return 2;
        }
    }
    if (file.empty()) {
        usage();
        return 2;
    }
    PSRDADA dada(file);
    if (header_only) {
        std::cout << "BW " << dada.bw << "\nFREQ " << dada.fc << "\nNBIT " << dada.nbit << "\nNCHAN " << dada.nchan
                  << "\nNDIM " << dada.ndim << "\nNPOL " << dada.npol << "\nfilesize " << dada.filesize << std::endl;
        return 0;
    }
    if (period == 0) {
        std::cout << "Please specify the period." << std::endl;
        return 1;
    }
    // the header gives the band centre and width in MHz; the engines take the
    // lowest frequency and the width in Hz (msosm_psrdada.cpp:100-104)
    const float bw = dada.bw * 1e6f;
    const float f0 = (dada.fc - dada.bw / 2) * 1e6f;
    MSOSM_GPU_BATCH pol1(bw, dm, f0), pol2(bw, dm, f0);
    pol1.initialize_uint16(fftpoint, count);
    pol2.initialize_uint16(fftpoint, count);
    const unsigned long n = (unsigned long)pol1.count * pol1.M;
    dada.initBuffer(n);
    Fold_GPU fold(period, bw, n, out.empty() ? dada.outfileName : out, time_bin);
    fold.discard_samples(pol1.Nd);
    std::cout << "M " << pol1.M << ", count " << pol1.count << ", Nd " << pol1.Nd << ", reads " << dada.readCount
              << std::endl;

    const auto t0 = std::chrono::steady_clock::now();
    bool leading = true;
    unsigned long used = 0;
    for (unsigned long r = 0; r < dada.readCount; r++) {
        dada.readSamples();
        if (leading && dada.pol1_in[0].first == 0)
            continue; // recording not started yet
        leading = false;
        pol1.filter_block_uint16(dada.pol1_in);
        pol2.filter_block_uint16(dada.pol2_in);
        fold.calculate_intensity<MSOSM_GPU_BATCH>(&pol1, &pol2);
        fold.fold_data();
        used++;
        if (dada.readCount >= 100 && r % (dada.readCount / 100) == 0)
            std::cout << "\r" << r * 100 / dada.readCount << "%" << std::flush;
    }
    fold.fold_data_bins();
    fold.write_to_file();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "\rblocks used " << used << " of " << dada.readCount << ", " << s << " s, "
              << (s > 0 ? 2.0 * used * n / s / 1e6 : 0.0) << " Msamples/s (both polarisations)" << std::endl;
    return 0;
}
