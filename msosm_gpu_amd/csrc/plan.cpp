// plan.cpp -- host planner: dispersion smearing Nd, filter order M, segment
// count K and the per-bin delay table, with the reference's float/double
// mixing reproduced operation by operation.  Built with -ffp-contract=off so
// no multiply-add is fused (the reference's x86-64 g++ build has no FMA).
//
// Restates dedispersion/msosm/msosm_prepare.cpp:3-30,86-152,
// dedispersion/msosm_dm_concurrent/msosm_prepare_dm_concurrent.cpp:3-49,
// dedispersion/overlap_save/osm_prepare.cpp:17-30,227-240 and the
// fftpoint/count/ring sizing of msosm_gpu_dm_concurrent.cpp:28-103.
#include "plan.h"

#include <cmath>
#include <cstdio>

namespace msosm {

static constexpr double kPi = 3.14159265358979323846; // globals.h:29 (M_PI)

DmModel make_dm_model(float bw, float dm, float f0)
{
    DmModel m;
    m.bw = bw;
    m.dm = dm;
    m.f0 = f0;
    m.fs = bw;                              // msosm_prepare.cpp:5
    m.kdm = (float)(4.15 * 1e15);           // :6, stored in a float member
    m.w0 = (float)(2 * kPi * (double)f0);   // :7
    // :8  floor(kdm*dm*(1/f0^2 - 1/(f0+bw)^2) / (1/fs)); kdm*dm and f0+bw are
    // float operations, the rest double, 1/fs float.
    float kd = m.kdm * dm;
    double inv0 = 1 / std::pow((double)f0, 2);
    float f1 = f0 + bw;
    double inv1 = 1 / std::pow((double)f1, 2);
    double t = (double)kd * (inv0 - inv1);
    float inv_fs = 1 / m.fs;
    t = t / (double)inv_fs;
    m.nd = (t >= 0) ? (unsigned long)std::floor(t) : 0;
    return m;
}

// msosm_prepare.cpp:17-30
unsigned long msosm_min_order(unsigned long nd)
{
    double order = 2 * std::sqrt((double)nd);
    return (unsigned long)std::pow(2, std::ceil(std::log2(order)));
}

// osm_prepare.cpp:17-30 with next_power_of_2(int) at :227-240
unsigned long osm_min_order(unsigned long nd)
{
    int n = (int)nd;
    int order = 1;
    if (n > 0) {
        --n;
        n |= n >> 1;
        n |= n >> 2;
        n |= n >> 4;
        n |= n >> 8;
        n |= n >> 16;
        order = n + 1;
    }
    return 2 * (unsigned long)order;
}

// msosm_prepare.cpp:88 -- ceil in float of Nd/(float)M
int delaycount_of(unsigned long nd, unsigned long M) { return (int)std::ceil((float)nd / (float)M); }

// msosm_prepare.cpp:86-152
int segment_delays(const DmModel &m, unsigned long M, int32_t *dp, int *holes)
{
    const unsigned long N = 2 * M;
    int K = delaycount_of(m.nd, M);
    if (K <= 0) {
        // Nd = 0 (DM = 0): the reference reads an uninitialised table here;
        // defined as one segment with zero delay.
        for (unsigned long k = 0; k < N; k++)
            dp[k] = 0;
        if (holes)
            *holes = 0;
        return 1;
    }
    std::vector<float> w(K + 1);
    w[0] = -1;
    w[K] = 1;
    int pos_index = 1;
    for (int i = 1; i < K; i++) {
        // :102, term by term: 1/w0^2 (double), i*M/fs (float), the constant
        // 4*pi*pi*kdm*dm (double), sqrt, then the normalisation to [-1, 1].
        double a = 1 / std::pow((double)m.w0, 2);
        float b = (float)((unsigned long)i * M) / m.fs;
        double c = (double)b / (4 * kPi * kPi * (double)m.kdm * (double)m.dm);
        double s = std::sqrt(1 / (a - c));
        w[i] = (float)((s - (double)m.w0 - kPi * (double)m.fs) / kPi / (double)m.fs);
        if (w[i] < 0)
            pos_index = i + 1;
    }
    const int negcount = pos_index;
    const int poscount = K - pos_index + 1;
    std::vector<unsigned long> segneg(negcount + 1), segpos(poscount + 1);
    for (int i = 0; i < negcount; i++)
        segneg[i] = (unsigned long)std::floor((w[i] + 2) * (float)M); // float product
    segneg[negcount] = N;
    segpos[0] = 0;
    for (int i = 0; i < poscount; i++)
        segpos[i + 1] = (unsigned long)std::floor(w[i + negcount] * (float)M);
    std::vector<unsigned char> set(N, 0);
    for (int j = 0; j < negcount; j++)
        for (unsigned long k = segneg[j]; k < segneg[j + 1] && k < N; k++) {
            dp[k] = j;                  // delayneg[j] = j
            set[k] = 1;
        }
    for (int j = 0; j < poscount; j++)
        for (unsigned long k = segpos[j]; k < segpos[j + 1] && k < N; k++) {
            dp[k] = negcount - 1 + j;   // delaypos[j]
            set[k] = 1;
        }
    int h = 0;
    for (unsigned long k = 0; k < N; k++)
        if (!set[k]) {
            dp[k] = 0;
            h++;
        }
    if (holes)
        *holes = h;
    return K;
}

static bool is_pow2(unsigned long v) { return v && !(v & (v - 1)); }

int build_plan(const msosm_plan_desc &d, Plan &p, std::string &msg)
{
    char buf[256];
    if (d.num_dms < 1 || !d.bw || !d.dm || !d.f0) {
        msg = "plan: num_dms must be >= 1 and bw/dm/f0 arrays non-NULL";
        return MSOSM_ERR_ARG;
    }
    if (d.method != MSOSM_METHOD_MSOSM && d.method != MSOSM_METHOD_OSM) {
        msg = "plan: unknown method";
        return MSOSM_ERR_ARG;
    }
    if (d.chirp != MSOSM_CHIRP_REF_F32W && d.chirp != MSOSM_CHIRP_EXACT_F64) {
        msg = "plan: unknown chirp mode";
        return MSOSM_ERR_ARG;
    }
    if (d.form < MSOSM_FORM_PLAIN_F32 || d.form > MSOSM_FORM_FUSED_F16_MFMA) {
        msg = "plan: unknown execution form";
        return MSOSM_ERR_ARG;
    }
    p.desc = d;
    p.desc.bw = p.desc.dm = p.desc.f0 = nullptr;
    p.num_dms = d.num_dms;
    p.dms.resize(d.num_dms);
    p.m_min.resize(d.num_dms);
    p.K.resize(d.num_dms);
    p.delay_block_size.resize(d.num_dms);
    unsigned long m_common_min = 0;
    for (int i = 0; i < d.num_dms; i++) {
        if (!(d.bw[i] > 0) || !(d.f0[i] > 0) || !(d.dm[i] >= 0) || !std::isfinite(d.dm[i])) {
            snprintf(buf, sizeof buf, "plan: DM trial %d needs bw > 0, f0 > 0 and finite dm >= 0 "
                                      "(got bw=%g f0=%g dm=%g)", i, d.bw[i], d.f0[i], d.dm[i]);
            msg = buf;
            return MSOSM_ERR_ARG;
        }
        p.dms[i] = make_dm_model(d.bw[i], d.dm[i], d.f0[i]);
        p.m_min[i] = (d.method == MSOSM_METHOD_MSOSM) ? msosm_min_order(p.dms[i].nd) : osm_min_order(p.dms[i].nd);
        if (p.m_min[i] > m_common_min)
            m_common_min = p.m_min[i]; // msosm_prepare_dm_concurrent.cpp:333-336
    }
    p.M_min = m_common_min;
    unsigned long M = m_common_min;
    if (d.fftpoint != 0) { // msosm_gpu_dm_concurrent.cpp:29-37
        M = d.fftpoint / 2;
        if (M < m_common_min) {
            fprintf(stdout, "Warning: The filter order is too small for the given FFT point. "
                            "The minimum FFT point %lu is used.\n", 2 * m_common_min);
            fflush(stdout);
            M = m_common_min;
            p.clamped = 1;
        }
    }
    if (M < 2 || !is_pow2(M)) {
        snprintf(buf, sizeof buf, "plan: FFT size 2M = %lu is not supported (the library needs a power of "
                                  "two >= 4; DM = 0 everywhere needs an explicit fftpoint)", 2 * M);
        msg = buf;
        return MSOSM_ERR_UNSUPPORTED;
    }
    if (2 * M > (1ul << 28)) {
        snprintf(buf, sizeof buf, "plan: FFT size %lu exceeds the supported maximum 2^28", 2 * M);
        msg = buf;
        return MSOSM_ERR_UNSUPPORTED;
    }
    p.M = M;
    p.N = 2 * M;
    int count = d.count;
    if (d.input_size != 0) // msosm_gpu_dm_concurrent.cpp:40-41
        count = (int)(d.input_size / M);
    if (count < 1) {
        msg = "plan: count (or input_size / M) must be >= 1";
        return MSOSM_ERR_ARG;
    }
    p.count = count;
    p.delay.assign((size_t)d.num_dms * p.N, 0);
    p.K_max = 1;
    p.R = 1;
    for (int i = 0; i < d.num_dms; i++) {
        int holes = 0;
        if (d.method == MSOSM_METHOD_MSOSM) {
            p.K[i] = segment_delays(p.dms[i], M, p.delay.data() + (size_t)i * p.N, &holes);
        } else {
            p.K[i] = 1; // single segment, zero delay
        }
        if (holes) {
            snprintf(buf, sizeof buf, "plan: segmentation of DM trial %d left %d bins unassigned", i, holes);
            msg = buf;
            return MSOSM_ERR_UNSUPPORTED;
        }
        // msosm_gpu_dm_concurrent.cpp:90-98
        int dbs = p.K[i] + count - 1;
        if (dbs % count != 0)
            dbs += count - dbs % count;
        p.delay_block_size[i] = dbs;
        if (dbs > p.R)
            p.R = dbs;
        if (p.K[i] > p.K_max)
            p.K_max = p.K[i];
    }
    if (p.K_max > 65535) {
        msg = "plan: more than 65535 segments";
        return MSOSM_ERR_UNSUPPORTED;
    }
    return MSOSM_OK;
}

} // namespace msosm
