// fold_gpu.h -- drop-in facade for Fold_GPU (folding/include/fold_gpu.h:10-41).
// The fold itself runs on the GPU (libmsosm.so); the public host arrays of
// the reference are kept and filled where the reference fills them.
#pragma once
#include <fstream>
#include <string>
#include <vector>

#include "globals.h"

class Fold_GPU {
public:
    // fold_gpu.cpp:3-27
    Fold_GPU(float period, float fs, unsigned long size, std::string outfileName = "fold.txt",
             unsigned long time_bin = -1)
        : size_(size)
    {
        msosm_facade::check(msosm_fold_create(period, fs, size, time_bin, nullptr, &f_), "Fold_GPU");
        msosm_facade::check(msosm_fold_info(f_, &period_samples, &this->time_bin, nullptr), "Fold_GPU");
        intensity_.assign(size, 0.0f);
        folded_.assign(this->time_bin, 0.0f);
        tmp_.assign(period_samples, 0.0f);
        counts_.assign(period_samples, 0);
        total_intensity = intensity_.data();
        folded_data = folded_.data();
        tmp_fold = tmp_.data();
        fold_count = counts_.data();
        outfile.open(outfileName);
    }
    ~Fold_GPU()
    {
        msosm_fold_destroy(f_);
        outfile.close();
    }
    Fold_GPU(const Fold_GPU &) = delete;
    Fold_GPU &operator=(const Fold_GPU &) = delete;

    // fold_gpu.cpp:29-33
    void discard_samples(unsigned long Nd) { msosm_facade::check(msosm_fold_discard_samples(f_, Nd), "discard"); }
    // fold_gpu.cpp:35-43: intensity of two polarisations' device outputs,
    // copied to total_intensity as the reference does.  The engines run on
    // their own streams, so their last calls are waited for first (the
    // reference ran everything on one stream).
    template <typename T>
    void calculate_intensity(T *pol1, T *pol2)
    {
        pol1->synchronize();
        pol2->synchronize();
        msosm_facade::check(msosm_fold_intensity(f_, pol1->output_buffer_d, pol2->output_buffer_d), "intensity");
        msosm_facade::check(msosm_fold_get(f_, nullptr, nullptr, total_intensity), "intensity");
    }
    // fold_gpu.cpp:48-173
    void fold_data() { msosm_facade::check(msosm_fold_fold_data(f_), "fold_data"); }
    // fold_gpu.cpp:175-194
    void fold_data_bins()
    {
        msosm_facade::check(msosm_fold_bins(f_), "fold_data_bins");
        msosm_facade::check(msosm_fold_get(f_, folded_data, nullptr, nullptr), "fold_data_bins");
    }
    // fold_gpu.cpp:196-202
    void write_to_file()
    {
        for (unsigned long i = 0; i < time_bin; i++)
            outfile << folded_data[i] << std::endl;
    }
    msosm_fold *handle() const { return f_; }

public:
    float *folded_data = nullptr;
    unsigned long time_bin = 0;
    float *total_intensity = nullptr;
    unsigned long period_samples = 0;
    unsigned long *fold_count = nullptr; // kept for source compatibility (the reference never reads it)
    float *tmp_fold = nullptr;

private:
    unsigned long size_;
    msosm_fold *f_ = nullptr;
    std::vector<float> intensity_, folded_, tmp_;
    std::vector<unsigned long> counts_;
    std::ofstream outfile;
};
