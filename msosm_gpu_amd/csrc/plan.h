// plan.h -- host-side plan of the MS-OSM / OSM engine (no GPU calls).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "msosm_c.h"

namespace msosm {

// Scalars of the dispersion model for one DM trial, kept in the reference's
// float types (prepare.h:37-48).
struct DmModel {
    float bw = 0, dm = 0, f0 = 0;
    float fs = 0, kdm = 0, w0 = 0;
    unsigned long nd = 0;
};

DmModel make_dm_model(float bw, float dm, float f0);
unsigned long msosm_min_order(unsigned long nd);
unsigned long osm_min_order(unsigned long nd);
int delaycount_of(unsigned long nd, unsigned long M);
// delay_points[2M]; returns K (>=1); holes = bins left unassigned.
int segment_delays(const DmModel &m, unsigned long M, int32_t *delay_points, int *holes);

struct Plan {
    msosm_plan_desc desc{};
    std::vector<DmModel> dms;
    std::vector<unsigned long> m_min;  // per DM
    std::vector<int> K;                // per DM delaycount
    std::vector<int> delay_block_size; // per DM (msosm_gpu_dm_concurrent.cpp:90-94)
    std::vector<int32_t> delay;        // [num_dms][2M]
    unsigned long M = 0, M_min = 0, N = 0;
    int count = 0, K_max = 1, R = 1;
    int clamped = 0;
    int num_dms = 0;
};

// Builds the plan; returns MSOSM_OK or an error code with msg filled.
int build_plan(const msosm_plan_desc &d, Plan &p, std::string &msg);

} // namespace msosm
