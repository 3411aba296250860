/*
 * msosm_oracle.h -- CPU restatement of the reference MS-OSM / OSM dedispersion
 * path and of its fold.  TEST INFRASTRUCTURE ONLY: it is the checker that the
 * tests, the smoke check and the CPU-baseline leg of the benchmark call; the
 * product library never links or calls it.
 *
 * Every function names the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed in the header comment of
 * msosm_oracle.c.
 */
#ifndef MSOSM_ORACLE_H
#define MSOSM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_CHIRP_REF_F32W 0  /* float-rounded frequency grid, as the reference */
#define ORA_CHIRP_EXACT_F64 1 /* exact FP64 grid and phase                       */

/* ---- planner (host scalars and tables) ---------------------------------- */
unsigned long ora_nd(float bw, float dm, float f0);
unsigned long ora_msosm_min_order(unsigned long nd);
unsigned long ora_osm_min_order(unsigned long nd);
/* fftpoint clamp: returns M (0 = invalid) and sets *clamped when the request
 * was below the minimum. */
unsigned long ora_clamp_order(unsigned long fftpoint, unsigned long m_min, int *clamped);
int ora_delaycount(unsigned long nd, unsigned long M);
/* delay_points[2M]; returns K (>=1).  *holes gets the number of bins the
 * segmentation never assigned (0 for every valid plan). */
int ora_delay_points(float bw, float dm, float f0, unsigned long M, int32_t *delay_points, int *holes);
/* Segment boundary arrays as the reference computes them (w[], negcount). */
int ora_segment_w(float bw, float dm, float f0, unsigned long M, float *w /*K+1*/, int *negcount);
/* Chirp samples C in the M-point IDFT input order (interleaved re,im; M). */
void ora_chirp(float bw, float dm, float f0, unsigned long M, int mode, double *c);
/* H = DFT_2M(zero-pad(IDFT_M(C))) in FP64 (interleaved, 2M complex). */
void ora_dedisp_params(float bw, float dm, float f0, unsigned long M, int mode, double *H);

/* ---- FFT (FP64, power-of-two, in place) --------------------------------- */
void ora_fft(double *x, size_t n, int sign);

/* ---- element kernels ----------------------------------------------------- */
void ora_unpack_u16(const uint16_t *src, float *dst, size_t nsamples);
void ora_pack_u16(const float *src, uint16_t *dst, size_t nsamples);
void ora_intensity(const float *a, const float *b, float *out, size_t n);

/* ---- full filters --------------------------------------------------------- */
/* FP64 same-plan MS-OSM over a whole stream x[L] (interleaved complex double,
 * L a multiple of M).  Output y[L] (interleaved).  OSM is the case delay==0. */
int ora_msosm_run(const double *x, size_t L, unsigned long M, const int32_t *delay_points,
                  const double *H, double *y);
/* Same, for a contiguous sub-range of output blocks [b0, b0+nb) only; x must
 * hold the stream from sample 0 up to (b0+nb)*M. */
/* ora_msosm_run_blocks for ndm trials sharing the forward spectra:
 * delay_points[ndm][2M], H[ndm][2M] complex, y[ndm][nb*M] complex. */
int ora_msosm_run_blocks_multi(const double *x, size_t L, unsigned long M, int ndm, const int32_t *delay_points,
                               const double *H, size_t b0, size_t nb, double *y);
int ora_msosm_run_blocks(const double *x, size_t L, unsigned long M, const int32_t *delay_points,
                         const double *H, size_t b0, size_t nb, double *y);

/* FP32 CPU baseline: multi-DM MS-OSM over a uint16 stream, OpenMP over blocks.
 * H32[ndm][2M] interleaved float, delay[ndm][2M].  Output y[ndm][L] interleaved
 * float (may be NULL: then outputs are reduced into a checksum only).
 * Returns the checksum (sum of |y|^2 over everything). */
double ora_cpu_msosm_f32(const uint16_t *u16, size_t L, unsigned long M, int ndm,
                         const int32_t *delay, const float *H32, float *y, int nthreads);

/* ---- fold (folding/fold_gpu.cpp) ----------------------------------------- */
typedef struct ora_fold ora_fold;
ora_fold *ora_fold_create(float period, float fs, unsigned long size, unsigned long time_bin);
void ora_fold_destroy(ora_fold *f);
void ora_fold_discard_samples(ora_fold *f, unsigned long nd);
void ora_fold_fold_data(ora_fold *f, const float *total_intensity);
void ora_fold_bins(ora_fold *f);
unsigned long ora_fold_period_samples(const ora_fold *f);
unsigned long ora_fold_time_bin(const ora_fold *f);
const float *ora_fold_raw(const ora_fold *f);
const float *ora_fold_folded(const ora_fold *f);
/* Bin index every input sample of the last fold_data call landed in
 * (-1 = skipped / discarded / pending); for bin-assignment checks. */
void ora_fold_trace(ora_fold *f, int64_t *raw_index_per_sample);
int ora_fold_ready(const ora_fold *f);

#ifdef __cplusplus
}
#endif
#endif
