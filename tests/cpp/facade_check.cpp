// facade_check.cpp -- drives the C++ drop-in facade the way the reference's
// drivers drive its classes (test_processing_speed_msosm.cpp,
// multiple_dm_concurrent_msosm.cpp, msosm_psrdada.cpp) and writes every
// output to files that tests/test_gpu_facade.py checks against the oracle.
//
// usage: facade_check <input.u16> <num_samples> <out_prefix> [default_plan | writer_threads]
//   default_plan: only the single-DM engine at the reference recording
//   program's plan (fftpoint 2097152, count 1, msosm_psrdada.cpp:94-95) at
//   128 MHz, DM 100, f0 1336 MHz
//   writer_threads: only the concurrent engine with two writer threads fed
//   by get_output_thread, alternating by call, each writer spinning on its
//   buffer's sentinel and appending what it gets to its own file, as
//   multiple_dm_concurrent_msosm.cpp:117-146 does
#include <cstdio>
#include <fstream>
#include <thread>
#include <vector>

#include "msosm/fold_gpu.h"
#include "msosm/gpu_kernel.h"
#include "msosm/msosm_gpu_batch.h"
#include "msosm/msosm_gpu_concurrent.h"
#include "msosm/osm_gpu.h"
#include "msosm/simulated_complex_signal.h"

template <class T>
static void dump(const std::string &path, const T *p, size_t n)
{
    std::ofstream f(path, std::ios::binary | std::ios::app);
    f.write(reinterpret_cast<const char *>(p), sizeof(T) * n);
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s input.u16 num_samples out_prefix\n", argv[0]);
        return 2;
    }
    const std::string pre = argv[3];
    const unsigned long L = std::stoul(argv[2]);
    std::vector<uint16_pair> in(L);
    {
        std::ifstream f(argv[1], std::ios::binary);
        f.read(reinterpret_cast<char *>(in.data()), sizeof(uint16_pair) * L);
        if (!f) {
            std::fprintf(stderr, "cannot read %lu samples from %s\n", L, argv[1]);
            return 2;
        }
    }
    if (argc >= 5 && std::string(argv[4]) == "default_plan") {
        std::remove((pre + ".big.c64").c_str());
        MSOSM_GPU_BATCH big(128e6f, 100.0f, 1336e6f);
        big.initialize_uint16(2097152, 1);
        std::ofstream meta(pre + ".meta");
        meta << "big M " << big.M << " Nd " << big.Nd << " K " << big.delaycount << " count " << big.count << "\n";
        std::vector<Complex> out((size_t)big.count * big.M);
        const unsigned long per = (unsigned long)big.count * big.M;
        for (unsigned long c = 0; c + per <= L; c += per) {
            big.filter_block_uint16(in.data() + c);
            big.get_output(out.data());
            dump(pre + ".big.c64", out.data(), out.size());
        }
        return 0;
    }
    if (argc >= 5 && std::string(argv[4]) == "writer_threads") {
        float bws[3] = {16e6f, 16e6f, 16e6f}, dms[3] = {5.0f, 10.0f, 20.0f}, f0s[3] = {1392e6f, 1392e6f, 1392e6f};
        MSOSM_GPU_DM_concurrent conc(bws, dms, f0s, 3);
        conc.initialize_uint16(1024, 8);
        conc.initialize_threads_flag(2);
        const size_t tot = (size_t)conc.count * conc.M_common * 3;
        const unsigned long per = (unsigned long)conc.count * conc.M_common;
        const int calls = (int)(L / per);
        std::vector<uint16_pair> buf[2] = {std::vector<uint16_pair>(tot + 1), std::vector<uint16_pair>(tot + 1)};
        std::vector<std::thread> writers;
        for (int id = 0; id < 2; id++) {
            std::remove((pre + ".thread" + std::to_string(id) + ".u16").c_str());
            buf[id][tot].first = 1;
            const int mine = (calls - id + 1) / 2;
            writers.emplace_back([&, id, mine] {
                uint16_pair *src = buf[id].data();
                for (int k = 0; k < mine; k++) {
                    while (__atomic_load_n(&src[tot].first, __ATOMIC_ACQUIRE) == 1)
                        ;
                    PUSH_RANGE("Output Copy", id)
                    __atomic_store_n(&src[tot].first, (uint16_t)1, __ATOMIC_RELAXED);
                    dump(pre + ".thread" + std::to_string(id) + ".u16", src, tot);
                    conc.wait_for_threads[id] = false;
                    POP_RANGE
                }
            });
        }
        for (int c = 0; c < calls; c++) {
            conc.filter_block_uint16(in.data() + (size_t)c * per);
            conc.get_output_thread(buf[c % 2].data(), c % 2);
        }
        for (auto &w : writers)
            w.join();
        conc.synchronize();
        std::ofstream meta(pre + ".meta");
        meta << "writer_threads calls " << calls << "\n";
        return 0;
    }
    for (const char *s : {".batch.c64", ".conc.c64", ".conc.u16", ".osm.c64", ".loop.c64", ".fold.f32", ".inten.f32",
                          ".dmstream.c64", ".dmstream.u16", ".bstream0.c64", ".bstream1.c64", ".sim.u16", ".sim.c64"})
        std::remove((pre + s).c_str());
    std::ofstream meta(pre + ".meta");

    // single DM (msosm_gpu_batch.h), 16 MHz, DM 10, f0 1392 MHz, N 1024
    const float bw = 16e6f, f0 = 1392e6f;
    MSOSM_GPU_BATCH one(bw, 10.0f, f0);
    one.initialize_uint16(1024, 8);
    meta << "batch M " << one.M << " M_min " << one.M_min << " Nd " << one.Nd << " K " << one.delaycount
         << " count " << one.count << "\n";
    std::vector<Complex> out((size_t)one.count * one.M);
    const unsigned long per = (unsigned long)one.count * one.M;
    for (unsigned long c = 0; c + per <= L; c += per) {
        one.filter_block_uint16(in.data() + c);
        one.get_output(out.data());
        dump(pre + ".batch.c64", out.data(), out.size());
    }

    // three DM trials in one engine (msosm_gpu_concurrent.h), sentinel layout
    float bws[3] = {bw, bw, bw}, dms[3] = {5.0f, 10.0f, 20.0f}, f0s[3] = {f0, f0, f0};
    MSOSM_GPU_DM_concurrent conc(bws, dms, f0s, 3);
    conc.initialize_uint16(1024, 8, 0, false);
    meta << "conc M " << conc.M_common << " count " << conc.count << " K " << conc.delaycount[0] << " "
         << conc.delaycount[1] << " " << conc.delaycount[2] << "\n";
    const size_t tot = (size_t)conc.count * conc.M_common * 3;
    std::vector<Complex> cbuf(tot + 1);
    std::vector<uint16_pair> ubuf(tot + 1);
    for (unsigned long c = 0; c + per <= L; c += per) {
        conc.filter_block_uint16(in.data() + c);
        cbuf[tot] = Complex{1.0f, 1.0f};
        PUSH_RANGE("Output Copy", 1) // as the reference drivers mark it
        conc.get_output(cbuf.data());
        POP_RANGE
        conc.wait_for_cpu = false; // what the consumer thread does once it has the block
        if (cbuf[tot].x != 0.0f)
            return 3; // sentinel must read 0 once the data is there
        dump(pre + ".conc.c64", cbuf.data(), tot);
        conc.get_output(ubuf.data());
        conc.wait_for_cpu = false;
        dump(pre + ".conc.u16", ubuf.data(), tot);
    }

    // DM loop engine: trial i handed out separately
    MSOSM_GPU_DM_BATCH_loop loop(bws, dms, f0s, 3);
    loop.initialize_uint16(1024, 8);
    std::vector<std::vector<Complex>> lo(3, std::vector<Complex>(per));
    Complex *lp[3] = {lo[0].data(), lo[1].data(), lo[2].data()};
    for (unsigned long c = 0; c + per <= L; c += per) {
        loop.input_process(in.data() + c);
        for (int i = 0; i < 3; i++)
            loop.filter_block_uint16(lp, i);
        for (int i = 0; i < 3; i++)
            dump(pre + ".loop.c64", lp[i], per);
    }

    // DM stream engine (msosm_gpu_batch_stream.h:52-93): trial i handed out
    // with its own flag and a sentinel after its count*M samples
    MSOSM_GPU_DM_BATCH_stream dms_eng(bws, dms, f0s, 3);
    dms_eng.initialize_uint16(1024, 8);
    std::vector<std::vector<Complex>> ds(3, std::vector<Complex>(per + 1));
    std::vector<std::vector<uint16_pair>> du(3, std::vector<uint16_pair>(per + 1));
    Complex *dsp[3] = {ds[0].data(), ds[1].data(), ds[2].data()};
    uint16_pair *dup[3] = {du[0].data(), du[1].data(), du[2].data()};
    for (unsigned long c = 0; c + per <= L; c += per) {
        dms_eng.filter_block_uint16(in.data() + c);
        dms_eng.get_output(dsp);
        for (int i = 0; i < 3; i++)
            dms_eng.wait_for_cpu[i] = false;
        dms_eng.get_output(dup);
        for (int i = 0; i < 3; i++) {
            dms_eng.wait_for_cpu[i] = false;
            dump(pre + ".dmstream.c64", dsp[i], per);
            dump(pre + ".dmstream.u16", dup[i], per);
        }
    }

    // independent streams with their own FFT sizes (msosm_gpu_batch_stream.h:14-49):
    // stream 0 DM 10 at N 1024, stream 1 DM 20 at N 2048; each call consumes
    // count*M[i] samples of stream i's own input (here both read the same
    // sample array, each at its own position)
    float sb[2] = {bw, bw}, sd[2] = {10.0f, 20.0f}, sf[2] = {f0, f0};
    unsigned long sfp[2] = {1024, 2048};
    MSOSM_GPU_BATCH_stream bst(sb, sd, sf, 2);
    bst.initialize_uint16(sfp, 4);
    meta << "bstream M " << bst.M[0] << " " << bst.M[1] << " fftpoint " << sfp[0] << " " << sfp[1] << "\n";
    std::vector<std::vector<Complex>> bo{std::vector<Complex>((size_t)bst.count * bst.M[0]),
                                         std::vector<Complex>((size_t)bst.count * bst.M[1])};
    Complex *bop[2] = {bo[0].data(), bo[1].data()};
    unsigned long pos[2] = {0, 0};
    while (pos[1] + (unsigned long)bst.count * bst.M[1] <= L) {
        uint16_pair *ip[2] = {in.data() + pos[0], in.data() + pos[1]};
        bst.filter_block_uint16(ip);
        bst.get_output(bop);
        dump(pre + ".bstream0.c64", bop[0], bo[0].size());
        dump(pre + ".bstream1.c64", bop[1], bo[1].size());
        pos[0] += (unsigned long)bst.count * bst.M[0];
        pos[1] += (unsigned long)bst.count * bst.M[1];
    }

    // single-segment OSM at DM 2 (M = 2 * nextpow2(Nd))
    OSM_GPU_BATCH osm(bw, 2.0f, f0);
    osm.initialize_uint16(0, 2);
    meta << "osm M " << osm.M << " count " << osm.count << " Nd " << osm.Nd << "\n";
    const unsigned long operr = (unsigned long)osm.count * osm.M;
    std::vector<Complex> oout(operr);
    for (unsigned long c = 0; c + operr <= L; c += operr) {
        osm.filter_block_uint16(in.data() + c);
        osm.get_output(oout.data());
        dump(pre + ".osm.c64", oout.data(), oout.size());
    }

    // fold of two "polarisations" (msosm_psrdada.cpp:105-142): same input
    MSOSM_GPU_BATCH p1(bw, 10.0f, f0), p2(bw, 10.0f, f0);
    p1.initialize_uint16(1024, 8);
    p2.initialize_uint16(1024, 8);
    Fold_GPU fold(0.0005f, bw, per, (pre + ".fold.txt"), 100);
    fold.discard_samples(p1.Nd);
    for (unsigned long c = 0; c + per <= L; c += per) {
        p1.filter_block_uint16(in.data() + c);
        p2.filter_block_uint16(in.data() + c);
        p1.synchronize();
        p2.synchronize();
        fold.calculate_intensity<MSOSM_GPU_BATCH>(&p1, &p2);
        dump(pre + ".inten.f32", fold.total_intensity, per);
        fold.fold_data();
    }
    fold.fold_data_bins();
    fold.write_to_file();
    dump(pre + ".fold.f32", fold.folded_data, fold.time_bin);
    meta << "fold P " << fold.period_samples << " time_bin " << fold.time_bin << "\n";
    // the reference drivers' test signal (check_msosm.cpp:101-105): uint16,
    // noise free, through the single-DM engine
    {
        SimulatedComplexSignal sim(bw, 10.0f, f0, 0.005f, "uint16");
        sim.generate_pulsar_signal(4, false, 0, false);
        meta << "sim Np " << sim.Np << " Nd " << sim.Nd << " range " << sim.range << " size " << sim.signal_size
             << "\n";
        dump(pre + ".sim.u16", sim.signal_u16, sim.signal_size);
        dump(pre + ".sim.c64", &sim.signal[0][0], 2 * sim.signal_size);
    }
    std::printf("facade_check ok\n");
    return 0;
}
