/*
 * msosm_oracle.c -- CPU restatement of the reference coherent-dedispersion
 * path (MS-OSM and single-segment OSM) and of its fold.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it; the product library
 * (msosm_gpu_amd/) never links or calls it.
 *
 * Build with -ffp-contract=off: the float/double mixing below must round
 * exactly as the reference's x86-64 g++ build (no FMA) does.
 *
 * What pins each part (no reference code can be compiled or run here: it
 * needs CUDA, cuFFT and FFTW, none of which exist on either machine):
 *   ora_nd, ora_*_min_order, ora_clamp_order, ora_delaycount
 *       -- per-config table in tests/golden/plan_tables.json written by an
 *          independent numpy restatement; closed forms (Nd ~ kdm*DM*(1/f0^2 -
 *          1/(f0+bw)^2)*fs, M = 2^ceil(log2(2 sqrt Nd))).
 *   ora_delay_points / ora_segment_w
 *       -- numpy restatement fixtures; invariants (values in [0,K), exactly K
 *          distinct values, monotone in frequency order, full coverage);
 *          brute-force delay floor(tau(f)*fs/M) away from boundaries.
 *   ora_chirp / ora_dedisp_params
 *       -- closed form H[2k] = M*C[k]; textbook O(M^2) DFT on tiny M;
 *          unit-modulus chirp; numpy fixtures.
 *   ora_fft -- brute-force DFT, Parseval, numpy.fft.
 *   ora_unpack_u16 / ora_pack_u16 -- exhaustive over all 65,536 codes.
 *   ora_msosm_run -- K=1 equals direct linear convolution with h (brute
 *       force); impulse response; numpy per-segment restatement fixtures;
 *       block-shift invariance.
 *   ora_fold_* -- crafted periods with known fractional parts (skip every
 *       1/frac periods), periodic impulse trains landing in known bins,
 *       split-invariance of the state machine, numpy restatement fixtures.
 *   ora_msosm_run_blocks_multi -- equal bit for bit, trial by trial, to
 *       ora_msosm_run_blocks.
 *   ora_cpu_msosm_f32 -- against ora_msosm_run (FP64) within FP32 rounding.
 */
#include "msosm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_PI 3.14159265358979323846 /* reference: globals.h:29 pi = M_PI */

/* The reference stores kdm = 4.15 * 1e15 in a float member
 * (msosm_prepare.cpp:6, prepare.h:193). */
static float ref_kdm(void) { return (float)(4.15 * 1e15); }

/* ------------------------------------------------------------------------ */
/* Nd -- msosm_prepare.cpp:8 (and osm_prepare.cpp:8,
 * msosm_prepare_dm_concurrent.cpp:12).  float*float, then double. */
unsigned long ora_nd(float bw, float dm, float f0)
{
    float fs = bw;
    float kdm = ref_kdm();
    float kd = kdm * dm;                     /* float product              */
    double inv_f0sq = 1 / pow((double)f0, 2);
    float f0bw = f0 + bw;                    /* float sum                  */
    double inv_f1sq = 1 / pow((double)f0bw, 2);
    double t = (double)kd * (inv_f0sq - inv_f1sq);
    float inv_fs = 1 / fs;                   /* "1 / fs" is float          */
    t = t / (double)inv_fs;
    if (!(t >= 0))                           /* negative DM: reference UB  */
        return 0;
    return (unsigned long)floor(t);
}

/* msosm_prepare.cpp:17-30: M = 2^ceil(log2(2*sqrt(Nd))). Nd=0 gives 0. */
unsigned long ora_msosm_min_order(unsigned long nd)
{
    double order = 2 * sqrt((double)nd);
    return (unsigned long)pow(2, ceil(log2(order)));
}

/* osm_prepare.cpp:17-30 with next_power_of_2 at :227-240 (int argument). */
unsigned long ora_osm_min_order(unsigned long nd)
{
    int n = (int)nd;
    int order;
    if (n <= 0) {
        order = 1;
    } else {
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

/* msosm_gpu_batch.cpp:16-26: M = fftpoint/2 clamped up to M_min. */
unsigned long ora_clamp_order(unsigned long fftpoint, unsigned long m_min, int *clamped)
{
    unsigned long M = m_min;
    int c = 0;
    if (fftpoint != 0) {
        M = fftpoint / 2;
        if (M < m_min) {
            M = m_min;
            c = 1;
        }
    }
    if (clamped)
        *clamped = c;
    return M;
}

/* msosm_prepare.cpp:88: ceil(Nd / (float)M) in float. */
int ora_delaycount(unsigned long nd, unsigned long M)
{
    return (int)ceilf((float)nd / (float)M);
}

/* Segment boundaries, msosm_prepare.cpp:95-107. */
int ora_segment_w(float bw, float dm, float f0, unsigned long M, float *w, int *negcount)
{
    float fs = bw;
    float kdm = ref_kdm();
    float w0 = (float)(2 * ORA_PI * (double)f0);
    unsigned long nd = ora_nd(bw, dm, f0);
    int K = ora_delaycount(nd, M);
    if (K <= 0) { /* DM=0: defined here as one segment with delay 0 */
        if (w) {
            w[0] = -1;
            w[1] = 1;
        }
        if (negcount)
            *negcount = 1;
        return 1;
    }
    w[0] = -1;
    w[K] = 1;
    int pos_index = 1;
    for (int i = 1; i < K; i++) {
        double a = 1 / pow((double)w0, 2);
        float b = (float)((unsigned long)i * M) / fs;               /* i*M/fs in float */
        double denom = 4 * ORA_PI * ORA_PI * (double)kdm * (double)dm;
        double c = (double)b / denom;
        double s = sqrt(1 / (a - c));
        double e = (s - (double)w0 - ORA_PI * (double)fs) / ORA_PI / (double)fs;
        w[i] = (float)e;
        if (w[i] < 0)
            pos_index = i + 1;
    }
    if (negcount)
        *negcount = pos_index;
    return K;
}

/* delay_points -- msosm_prepare.cpp:86-152. */
int ora_delay_points(float bw, float dm, float f0, unsigned long M, int32_t *dp, int *holes)
{
    unsigned long N = 2 * M;
    unsigned long nd = ora_nd(bw, dm, f0);
    int K = ora_delaycount(nd, M);
    if (K <= 0) {
        for (unsigned long k = 0; k < N; k++)
            dp[k] = 0;
        if (holes)
            *holes = 0;
        return 1;
    }
    float *w = (float *)malloc(sizeof(float) * (size_t)(K + 1));
    int negcount = 0;
    ora_segment_w(bw, dm, f0, M, w, &negcount);
    int poscount = K - negcount + 1;
    unsigned long *segneg = (unsigned long *)malloc(sizeof(unsigned long) * (size_t)(negcount + 1));
    unsigned long *segpos = (unsigned long *)malloc(sizeof(unsigned long) * (size_t)(poscount + 1));
    for (int i = 0; i < negcount; i++)
        segneg[i] = (unsigned long)floorf((w[i] + 2) * (float)M);
    segneg[negcount] = N;
    segpos[0] = 0;
    for (int i = 0; i < poscount; i++)
        segpos[i + 1] = (unsigned long)floorf(w[i + negcount] * (float)M);
    unsigned char *seen = (unsigned char *)calloc(N, 1);
    for (int j = 0; j < negcount; j++)
        for (unsigned long k = segneg[j]; k < segneg[j + 1] && k < N; k++) {
            dp[k] = j;
            seen[k] = 1;
        }
    for (int j = 0; j < poscount; j++)
        for (unsigned long k = segpos[j]; k < segpos[j + 1] && k < N; k++) {
            dp[k] = negcount - 1 + j;
            seen[k] = 1;
        }
    int h = 0;
    for (unsigned long k = 0; k < N; k++)
        if (!seen[k]) {
            dp[k] = 0;
            h++;
        }
    if (holes)
        *holes = h;
    free(seen);
    free(segneg);
    free(segpos);
    free(w);
    return K;
}

/* Chirp samples in IDFT input order -- msosm_prepare.cpp:40-63. */
void ora_chirp(float bw, float dm, float f0, unsigned long M, int mode, double *c)
{
    float fs = bw;
    float kdm = ref_kdm();
    float w0 = (float)(2 * ORA_PI * (double)f0);
    float step = (float)(2 * ORA_PI * (double)fs / (double)M);
    for (unsigned long i = 0; i < M; i++) {
        double ph;
        if (mode == ORA_CHIRP_REF_F32W) {
            float wi = (float)(-ORA_PI * (double)fs + (double)(step * (float)i));
            wi = (float)((double)wi + ORA_PI * (double)fs);
            float wsum = wi + w0; /* float sum, as (w[i] + w0) */
            ph = 4 * ORA_PI * ORA_PI * (double)kdm * (double)dm * (double)wi * (double)wi
                 / (double)wsum / (double)w0 / (double)w0;
        } else {
            /* phi = 2*pi*D*DM*f^2/(f0^2 (f0+f)), f = i*fs/M, all FP64 */
            double f = (double)fs * (double)i / (double)M;
            double F0 = (double)f0;
            ph = 2 * ORA_PI * 4.15e15 * (double)dm * f * f / (F0 * F0 * (F0 + f));
        }
        unsigned long idx = (i < M / 2) ? (M / 2 + i) : (i - M / 2);
        c[2 * idx] = (double)(float)cos(ph);
        c[2 * idx + 1] = (double)(float)(-sin(ph));
        if (mode == ORA_CHIRP_EXACT_F64) {
            c[2 * idx] = cos(ph);
            c[2 * idx + 1] = -sin(ph);
        }
    }
}

/* H = DFT_2M(pad(IDFT_M(C))) -- msosm_prepare.cpp:65-84.  The /M at :68-72 is
 * overwritten by the memcpy at :74, so the IDFT is unnormalised. */
void ora_dedisp_params(float bw, float dm, float f0, unsigned long M, int mode, double *H)
{
    double *c = (double *)malloc(sizeof(double) * 2 * M);
    ora_chirp(bw, dm, f0, M, mode, c);
    ora_fft(c, M, +1);
    memset(H, 0, sizeof(double) * 4 * M);
    memcpy(H, c, sizeof(double) * 2 * M);
    ora_fft(H, 2 * M, -1);
    free(c);
}

/* ------------------------------------------------------------------------ */
/* Iterative radix-2 FFT, FP64.  sign=-1 forward, +1 inverse (unnormalised). */
void ora_fft(double *x, size_t n, int sign)
{
    if (n <= 1)
        return;
    size_t j = 0;
    for (size_t i = 1; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j];
            x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr;
            x[2 * j + 1] = ti;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1;
        for (size_t k = 0; k < half; k++) {
            double ang = sign * 2 * ORA_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (size_t s = 0; s < n; s += len) {
                double *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                double br = b[0] * wr - b[1] * wi;
                double bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br;
                b[1] = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* gpu_kernel.cu:276-299: offset binary, 0 stays 0 ("no data"). */
void ora_unpack_u16(const uint16_t *src, float *dst, size_t n)
{
    for (size_t i = 0; i < 2 * n; i++) {
        uint16_t v = src[i];
        dst[i] = (float)v - 32768.0f * (float)(v != 0);
    }
}

/* gpu_kernel.cu:317-329: truncating cast after adding 32768 to non-zero
 * values.  Out-of-range values saturate to [0,65535], NaN gives 0 (the
 * reference's cast is undefined there). */
void ora_pack_u16(const float *src, uint16_t *dst, size_t n)
{
    for (size_t i = 0; i < 2 * n; i++) {
        float v = src[i];
        float t = v + 32768.0f * (float)(v != 0);
        uint16_t u;
        if (!(t > 0))
            u = 0;
        else if (t >= 65535.0f)
            u = 65535;
        else
            u = (uint16_t)t;
        dst[i] = u;
    }
}

/* gpu_kernel.cu:347-353: sqrt(|a|^2+|b|^2), summed left to right. */
void ora_intensity(const float *a, const float *b, float *out, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        float s = a[2 * i] * a[2 * i];
        s = s + a[2 * i + 1] * a[2 * i + 1];
        s = s + b[2 * i] * b[2 * i];
        s = s + b[2 * i + 1] * b[2 * i + 1];
        out[i] = sqrtf(s);
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 same-plan MS-OSM.  Restates msosm_gpu_batch.cpp:75-108 with the
 * kernels at gpu_kernel.cu:84-100 (multiply by H/N, ring scatter at delay d)
 * and :216-226 (keep the last M of each inverse FFT, divide by 2M).
 * Output block g = last M samples of IDFT(Y_g)/N with
 * Y_g[n] = X_{g-d(n)}[n] * H[n] / N and X_g = DFT of stream samples
 * [(g-1)M, (g+1)M) (zeros before the start). */
int ora_msosm_run_blocks(const double *x, size_t L, unsigned long M, const int32_t *dp,
                         const double *H, size_t b0, size_t nb, double *y)
{
    if (M == 0 || L % M != 0 || b0 + nb > L / M)
        return -1;
    size_t N = 2 * M;
    int K = 1;
    for (size_t n = 0; n < N; n++)
        if (dp[n] + 1 > K)
            K = dp[n] + 1;
    double *hist = (double *)calloc((size_t)K * 2 * N, sizeof(double));
    double *Y = (double *)malloc(sizeof(double) * 2 * N);
    size_t gstart = (b0 >= (size_t)(K - 1)) ? b0 - (size_t)(K - 1) : 0;
    for (size_t g = gstart; g < b0 + nb; g++) {
        double *X = hist + (g % (size_t)K) * 2 * N;
        for (size_t s = 0; s < N; s++) {
            long long idx = (long long)(g * M + s) - (long long)M;
            if (idx >= 0 && (size_t)idx < L) {
                X[2 * s] = x[2 * idx];
                X[2 * s + 1] = x[2 * idx + 1];
            } else {
                X[2 * s] = 0;
                X[2 * s + 1] = 0;
            }
        }
        ora_fft(X, N, -1);
        if (g < b0)
            continue;
        for (size_t n = 0; n < N; n++) {
            long long src = (long long)g - dp[n];
            if (src < (long long)gstart) {
                Y[2 * n] = 0;
                Y[2 * n + 1] = 0;
                continue;
            }
            const double *Xs = hist + ((size_t)src % (size_t)K) * 2 * N;
            double ar = Xs[2 * n] / (double)N, ai = Xs[2 * n + 1] / (double)N;
            Y[2 * n] = ar * H[2 * n] - ai * H[2 * n + 1];
            Y[2 * n + 1] = ar * H[2 * n + 1] + ai * H[2 * n];
        }
        ora_fft(Y, N, +1);
        for (size_t o = 0; o < M; o++) {
            y[2 * ((g - b0) * M + o)] = Y[2 * (M + o)] / (double)N;
            y[2 * ((g - b0) * M + o) + 1] = Y[2 * (M + o) + 1] / (double)N;
        }
    }
    free(hist);
    free(Y);
    return 0;
}

int ora_msosm_run(const double *x, size_t L, unsigned long M, const int32_t *dp, const double *H, double *y)
{
    return ora_msosm_run_blocks(x, L, M, dp, H, 0, L / M, y);
}

/* The same for ndm trials of one stream (msosm_gpu_dm_concurrent.cpp:119-151:
 * one forward spectrum per block shared by every trial): output blocks b0 ..
 * b0+nb-1 of trial t go to y[t][nb*M], delays dp[t][2M], filters H[t][2M].
 * The arithmetic per trial is that of ora_msosm_run_blocks operation for
 * operation (same window FFTs, same products, same inverse), so each row
 * equals ora_msosm_run_blocks for that trial bit for bit; trials run in
 * parallel. */
int ora_msosm_run_blocks_multi(const double *x, size_t L, unsigned long M, int ndm, const int32_t *dp,
                               const double *H, size_t b0, size_t nb, double *y)
{
    if (M == 0 || L % M != 0 || b0 + nb > L / M || ndm < 1)
        return -1;
    const size_t N = 2 * M;
    int K = 1;
    for (size_t i = 0; i < (size_t)ndm * N; i++)
        if (dp[i] + 1 > K)
            K = dp[i] + 1;
    /* spectra of blocks gstart .. b0+nb-1, where gstart is the first block
     * any trial reads: the single-trial routine starts at b0 - (K_t - 1) and
     * treats blocks before that as zero, so each trial below does the same
     * with its own K_t */
    const size_t gstart = (b0 >= (size_t)(K - 1)) ? b0 - (size_t)(K - 1) : 0;
    const size_t nw = b0 + nb - gstart;
    double *spec = (double *)malloc(sizeof(double) * 2 * N * nw);
    if (!spec)
        return -2;
#pragma omp parallel for schedule(dynamic)
    for (long long w = 0; w < (long long)nw; w++) {
        const size_t g = gstart + (size_t)w;
        double *X = spec + (size_t)w * 2 * N;
        for (size_t s = 0; s < N; s++) {
            long long idx = (long long)(g * M + s) - (long long)M;
            if (idx >= 0 && (size_t)idx < L) {
                X[2 * s] = x[2 * idx];
                X[2 * s + 1] = x[2 * idx + 1];
            } else {
                X[2 * s] = 0;
                X[2 * s + 1] = 0;
            }
        }
        ora_fft(X, N, -1);
    }
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < ndm; t++) {
        const int32_t *d = dp + (size_t)t * N;
        const double *Ht = H + (size_t)t * 2 * N;
        int Kt = 1;
        for (size_t n = 0; n < N; n++)
            if (d[n] + 1 > Kt)
                Kt = d[n] + 1;
        const size_t gs = (b0 >= (size_t)(Kt - 1)) ? b0 - (size_t)(Kt - 1) : 0;
        double *Y = (double *)malloc(sizeof(double) * 2 * N);
        for (size_t g = b0; g < b0 + nb; g++) {
            for (size_t n = 0; n < N; n++) {
                long long src = (long long)g - d[n];
                if (src < (long long)gs) {
                    Y[2 * n] = 0;
                    Y[2 * n + 1] = 0;
                    continue;
                }
                const double *Xs = spec + ((size_t)src - gstart) * 2 * N;
                double ar = Xs[2 * n] / (double)N, ai = Xs[2 * n + 1] / (double)N;
                Y[2 * n] = ar * Ht[2 * n] - ai * Ht[2 * n + 1];
                Y[2 * n + 1] = ar * Ht[2 * n + 1] + ai * Ht[2 * n];
            }
            ora_fft(Y, N, +1);
            double *yt = y + (size_t)t * 2 * nb * M;
            for (size_t o = 0; o < M; o++) {
                yt[2 * ((g - b0) * M + o)] = Y[2 * (M + o)] / (double)N;
                yt[2 * ((g - b0) * M + o) + 1] = Y[2 * (M + o) + 1] / (double)N;
            }
        }
        free(Y);
    }
    free(spec);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP32 CPU baseline: same plan, own radix-2 FP32 FFT, OpenMP over blocks.  The
 * forward spectrum of each block is shared by all DMs (as the reference's
 * concurrent engine, msosm_gpu_dm_concurrent.cpp:119-151). */
static void fft_f32(float *x, size_t n, int sign, const float *tw /* n/2 complex, forward */)
{
    size_t j = 0;
    for (size_t i = 1; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j) {
            float tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j];
            x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr;
            x[2 * j + 1] = ti;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1, step = n / len;
        for (size_t s = 0; s < n; s += len) {
            for (size_t k = 0; k < half; k++) {
                float wr = tw[2 * k * step], wi = (float)sign * -tw[2 * k * step + 1];
                float *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                float br = b[0] * wr - b[1] * wi;
                float bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br;
                b[1] = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
        }
    }
}

double ora_cpu_msosm_f32(const uint16_t *u16, size_t L, unsigned long M, int ndm, const int32_t *delay,
                         const float *H32, float *y, int nthreads)
{
    if (M == 0 || L % M != 0)
        return -1;
    size_t N = 2 * M, nblk = L / M;
    int K = 1;
    for (size_t i = 0; i < (size_t)ndm * N; i++)
        if (delay[i] + 1 > K)
            K = delay[i] + 1;
    float *tw = (float *)malloc(sizeof(float) * N);
    for (size_t k = 0; k < N / 2; k++) {
        tw[2 * k] = (float)cos(2 * ORA_PI * (double)k / (double)N);
        tw[2 * k + 1] = (float)(-sin(2 * ORA_PI * (double)k / (double)N));
    }
    /* all forward spectra (shared by every DM) */
    float *X = (float *)malloc(sizeof(float) * 2 * N * nblk);
    double total = 0;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (long long g = 0; g < (long long)nblk; g++) {
        float *Xg = X + (size_t)g * 2 * N;
        for (size_t s = 0; s < N; s++) {
            long long idx = g * (long long)M + (long long)s - (long long)M;
            float re = 0, im = 0;
            if (idx >= 0) {
                uint16_t a = u16[2 * idx], b = u16[2 * idx + 1];
                re = (float)a - 32768.0f * (float)(a != 0);
                im = (float)b - 32768.0f * (float)(b != 0);
            }
            Xg[2 * s] = re;
            Xg[2 * s + 1] = im;
        }
        fft_f32(Xg, N, -1, tw);
    }
    float invN = 1.0f / (float)N;
#pragma omp parallel reduction(+ : total)
    {
        float *Y = (float *)malloc(sizeof(float) * 2 * N);
#pragma omp for schedule(static) collapse(2)
        for (int d = 0; d < ndm; d++) {
            for (long long g = 0; g < (long long)nblk; g++) {
                const int32_t *dp = delay + (size_t)d * N;
                const float *Hd = H32 + (size_t)d * 2 * N;
                for (size_t n = 0; n < N; n++) {
                    long long src = g - dp[n];
                    if (src < 0) {
                        Y[2 * n] = 0;
                        Y[2 * n + 1] = 0;
                        continue;
                    }
                    const float *Xs = X + (size_t)src * 2 * N + 2 * n;
                    float ar = Xs[0] * invN, ai = Xs[1] * invN;
                    Y[2 * n] = ar * Hd[2 * n] - ai * Hd[2 * n + 1];
                    Y[2 * n + 1] = ar * Hd[2 * n + 1] + ai * Hd[2 * n];
                }
                fft_f32(Y, N, +1, tw);
                double acc = 0;
                for (size_t o = 0; o < M; o++) {
                    float re = Y[2 * (M + o)] * invN, im = Y[2 * (M + o) + 1] * invN;
                    if (y) {
                        float *yo = y + ((size_t)d * L + (size_t)g * M + o) * 2;
                        yo[0] = re;
                        yo[1] = im;
                    }
                    acc += (double)re * re + (double)im * im;
                }
                total += acc;
            }
        }
        free(Y);
    }
    free(X);
    free(tw);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Fold -- folding/fold_gpu.cpp:3-202, restated with the same float state. */
struct ora_fold {
    unsigned long size;
    float period_samples_float;
    unsigned long period_samples;
    float diff_samples;
    unsigned long time_bin;
    float *folded_data_raw;
    float *tmp_fold;
    float *folded_data;
    int ready;
    int discard_count;
    int current_discard;
    float current_diff;
    unsigned long current_index;
    int64_t *trace; /* raw bin per sample of the last call, -1 = skipped */
};

ora_fold *ora_fold_create(float period, float fs, unsigned long size, unsigned long time_bin)
{
    ora_fold *f = (ora_fold *)calloc(1, sizeof(ora_fold));
    f->size = size;
    f->period_samples_float = period * fs;                                   /* :6 */
    f->period_samples = (unsigned long)f->period_samples_float;              /* :7 */
    f->diff_samples = f->period_samples_float - (float)f->period_samples;    /* :8 */
    f->time_bin = (time_bin == (unsigned long)-1) ? f->period_samples : time_bin;
    /* the reference leaves these uninitialised (:17-20); zeroed here */
    f->folded_data_raw = (float *)calloc(f->period_samples, sizeof(float));
    f->tmp_fold = (float *)calloc(f->period_samples, sizeof(float));
    f->folded_data = (float *)calloc(f->time_bin ? f->time_bin : 1, sizeof(float));
    f->trace = (int64_t *)malloc(sizeof(int64_t) * (size ? size : 1));
    return f;
}

void ora_fold_destroy(ora_fold *f)
{
    if (!f)
        return;
    free(f->folded_data_raw);
    free(f->tmp_fold);
    free(f->folded_data);
    free(f->trace);
    free(f);
}

/* :29-33.  A discard count of 0 (Nd=0, or discard_samples never called)
 * would leave the reference never "ready" (:113-118 never matches); here it
 * means ready from the start. */
void ora_fold_discard_samples(ora_fold *f, unsigned long nd)
{
    f->discard_count = (int)ceilf((float)nd / f->period_samples_float);
    f->current_discard = 0;
    f->ready = (f->discard_count == 0);
}

static void fold_period_done(ora_fold *f)
{
    if (!f->ready) {
        f->current_discard += 1;
        if (f->current_discard == f->discard_count) {
            f->ready = 1;
            f->current_discard = 0;
        }
    }
}

/* :48-173 */
void ora_fold_fold_data(ora_fold *f, const float *in)
{
    const unsigned long P = f->period_samples;
    const unsigned long NEG1 = (unsigned long)-1;
    unsigned long current_size = f->size;
    unsigned long ci = 0; /* current_input_index */
    if (f->discard_count == 0 && !f->ready)
        f->ready = 1;
    for (unsigned long i = 0; i < f->size; i++)
        f->trace[i] = -1;

    if (f->current_diff > 1.0f) { /* :54-61 deferred skip */
        f->current_diff -= 1.0f;
        ci += 1;
        current_size -= 1;
        if (current_size == 0)
            return;
    }
    if (f->current_index != 0) { /* :62-121 finish the staged period */
        unsigned long ci0 = f->current_index;
        if (current_size < P - ci0) {
            for (unsigned long i = 0; i < current_size; i++) {
                f->trace[ci + i] = (int64_t)(ci0 + i);
                if (f->ready)
                    f->tmp_fold[ci0 + i] = in[ci + i];
            }
            f->current_index += current_size;
            return;
        }
        for (unsigned long i = 0; i < P - ci0; i++) {
            f->trace[ci + i] = (int64_t)(ci0 + i);
            if (f->ready)
                f->tmp_fold[ci0 + i] = in[ci + i];
        }
        if (f->ready)
            for (unsigned long i = 0; i < P; i++)
                f->folded_data_raw[i] += f->tmp_fold[i];
        current_size -= P - ci0;
        ci += P - ci0;
        f->current_index = 0;
        f->current_diff += f->diff_samples;
        if (f->current_diff > 1.0f) {
            f->current_diff -= 1.0f;
            ci += 1;
            current_size -= 1;
            if (current_size == NEG1) {
                f->current_diff += 1.0f;
                return;
            }
        }
        fold_period_done(f);
    }
    while (current_size >= P) { /* :124-152 whole periods */
        for (unsigned long i = 0; i < P; i++) {
            f->trace[ci + i] = (int64_t)i;
            if (f->ready)
                f->folded_data_raw[i] += in[ci + i];
        }
        current_size -= P;
        ci += P;
        f->current_diff += f->diff_samples;
        if (f->current_diff > 1.0f) {
            f->current_diff -= 1.0f;
            ci += 1;
            current_size -= 1;
            /* the reference wraps current_size here when the period ends
             * exactly at the buffer end and keeps reading past it; defined
             * here like the branch at :105-109 (defer the skip). */
            if (current_size == NEG1) {
                f->current_diff += 1.0f;
                return;
            }
        }
        fold_period_done(f);
    }
    if (current_size > 0) { /* :160-172 stage the partial period */
        for (unsigned long i = 0; i < current_size; i++) {
            f->trace[ci + i] = (int64_t)i;
            if (f->ready)
                f->tmp_fold[i] = in[ci + i];
        }
        f->current_index = current_size;
    }
}

/* :175-194 */
void ora_fold_bins(ora_fold *f)
{
    if (f->time_bin == f->period_samples) {
        memcpy(f->folded_data, f->folded_data_raw, sizeof(float) * f->period_samples);
        return;
    }
    int bpt = (int)(f->period_samples / f->time_bin);
    for (unsigned long i = 0; i < f->time_bin; i++) {
        float sum = 0.0f;
        for (unsigned long j = 0; j < (unsigned long)bpt; j++)
            sum += f->folded_data_raw[i * (unsigned long)bpt + j];
        f->folded_data[i] = sum / (float)bpt;
    }
}

unsigned long ora_fold_period_samples(const ora_fold *f) { return f->period_samples; }
unsigned long ora_fold_time_bin(const ora_fold *f) { return f->time_bin; }
const float *ora_fold_raw(const ora_fold *f) { return f->folded_data_raw; }
const float *ora_fold_folded(const ora_fold *f) { return f->folded_data; }
int ora_fold_ready(const ora_fold *f) { return f->ready; }
void ora_fold_trace(ora_fold *f, int64_t *out) { memcpy(out, f->trace, sizeof(int64_t) * f->size); }
