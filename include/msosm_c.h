/*
 * msosm_c.h -- C ABI of the MI355X-native MS-OSM coherent-dedispersion
 * engine (libmsosm.so).
 *
 * Plain C types only: opaque handles, pointers, sizes.  Device pointers are
 * HIP device addresses passed as void*; a HIP stream is passed as void*
 * (NULL = the engine's own stream).  Complex samples are interleaved float
 * pairs {re, im} (layout-compatible with the reference's `Complex`, a
 * cufftComplex, utils/include/globals.h:28) and input samples are packed
 * offset-binary uint16 pairs (`uint16_pair`, globals.h:21-25).
 *
 * Every function returns MSOSM_OK (0) or an error code; msosm_last_error()
 * gives the message of the last failure on the calling thread.  The C++
 * facade in include/msosm/ turns a failure into the reference's
 * print-and-exit(1) behaviour (gpu/include/gpu_common.cuh:41-61).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef MSOSM_C_H
#define MSOSM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define MSOSM_OK 0
#define MSOSM_ERR_ARG 1         /* invalid argument                          */
#define MSOSM_ERR_HIP 2         /* a HIP runtime call failed                 */
#define MSOSM_ERR_STATE 3       /* call out of order                         */
#define MSOSM_ERR_UNSUPPORTED 4 /* valid request the library does not support */
#define MSOSM_ERR_ALLOC 5       /* device or host allocation failed          */

/* ---- plan descriptor ------------------------------------------------------ */
#define MSOSM_METHOD_MSOSM 0 /* multi-segment overlap-save (msosm_prepare.cpp) */
#define MSOSM_METHOD_OSM 1   /* single-segment overlap-save (osm_prepare.cpp) */

#define MSOSM_CHIRP_REF_F32W 0  /* chirp on the reference's float-rounded grid */
#define MSOSM_CHIRP_EXACT_F64 1 /* chirp phase on an exact FP64 grid           */

/* Execution forms (all produce the same output layout [dm][count][M]):      */
#define MSOSM_FORM_PLAIN_F32 0 /* reference structure: per-DM product ring,   */
                               /* separate inverse FFT and discard (FP32)     */
#define MSOSM_FORM_FUSED_F32 1 /* shared spectra ring, gather*H fused into    */
                               /* the inverse FFT, discard fused (FP32)       */
#define MSOSM_FORM_FUSED_F16 2 /* as FUSED_F32 with FP16 spectra ring and H   */
                               /* (pre-scaled), FP32 arithmetic               */
/* Delay-line accumulate as per-bin complex GEMMs on the matrix cores       */
/* (Y_n = A_n B_n, A_n one H per row at the segment delay, B_n the Hankel  */
/* block of ring spectra), then a separate inverse FFT of the products:     */
#define MSOSM_FORM_HANKEL_F16 3    /* FP16 operands, FP32 accumulate         */
#define MSOSM_FORM_HANKEL_BF16 4   /* BF16 operands (single), FP32 accumulate:*/
                                   /* misses the 1e-3 mixed gate, so engines */
                                   /* refuse it; msosm_k_hankel_delay_gemm   */
                                   /* still takes it (GEMM studies)          */
#define MSOSM_FORM_HANKEL_BF16X3 5 /* BF16 hi+lo split operands, 3 products  */
/* As FUSED_F16 with the inverse FFT's DFT stages on the matrix cores (FP16 */
/* data entering each radix pass, FP32 accumulate and twiddles); FFT size   */
/* 16384 only:                                                              */
#define MSOSM_FORM_FUSED_F16_MFMA 6

typedef struct msosm_plan_desc {
    int num_dms;             /* DM trials sharing one input stream (>=1)      */
    const float *bw;         /* [num_dms] complex bandwidth = sample rate, Hz */
    const float *dm;         /* [num_dms] dispersion measure, pc cm^-3 (>=0)  */
    const float *f0;         /* [num_dms] lowest frequency of the band, Hz    */
    unsigned long fftpoint;  /* requested FFT size 2M, 0 = minimum            */
    int count;               /* blocks of M samples per call (>=1)            */
    unsigned long input_size;/* !=0: count = input_size / M                   */
    int method;              /* MSOSM_METHOD_*                                */
    int chirp;               /* MSOSM_CHIRP_*                                 */
    int form;                /* MSOSM_FORM_*                                  */
} msosm_plan_desc;

typedef struct msosm_plan_info {
    unsigned long M;         /* hop (segment length); FFT size is 2M          */
    unsigned long M_min;     /* minimum M over the DMs                        */
    unsigned long fftpoint;  /* 2M                                            */
    int count;               /* blocks per call                               */
    int num_dms;
    int K_max;               /* largest segment count (delaycount)            */
    int R;                   /* ring slots of the reference layout            */
                             /* (max delay_block_size)                         */
    int clamped;             /* 1 if fftpoint was raised to the minimum       */
} msosm_plan_info;

typedef struct msosm_plan msosm_plan;
typedef struct msosm_engine msosm_engine;
typedef struct msosm_fold msosm_fold;

/* Planner (host only; no GPU needed).
 * Replaces Prepare_MSOSM / Prepare_OSM / Prepare_MSOSM_DM_concurrent ctor +
 * calculate_min_order + segmentation (dedispersion/msosm/msosm_prepare.cpp:3-152,
 * msosm_dm_concurrent/msosm_prepare_dm_concurrent.cpp:3-169,
 * overlap_save/osm_prepare.cpp:3-30) and the fftpoint/count logic of
 * initialize_uint16 (msosm_gpu_dm_concurrent.cpp:28-42). */
int msosm_plan_create(const msosm_plan_desc *desc, msosm_plan **out);
/* Scalars of one DM trial without a plan: Nd (Prepare_MSOSM ctor,
 * msosm_prepare.cpp:8) and the minimum order M_min of the method
 * (calculate_min_order, msosm_prepare.cpp:17-30 / osm_prepare.cpp:17-30). */
int msosm_dm_model(float bw, float dm, float f0, int method, unsigned long *nd, unsigned long *m_min);
void msosm_plan_destroy(msosm_plan *plan);
int msosm_plan_get_info(const msosm_plan *plan, msosm_plan_info *info);
/* Per-DM scalars: Nd, M_min, delaycount K, delay_block_size. */
int msosm_plan_dm_info(const msosm_plan *plan, int dm_index, unsigned long *nd, unsigned long *m_min,
                       int *delaycount, int *delay_block_size);
/* Per-bin delay table delay_points[2M] (Prepare_MSOSM::segmentation,
 * msosm_prepare.cpp:133-147). */
int msosm_plan_delay_points(const msosm_plan *plan, int dm_index, int32_t *out);

/* Engine (GPU).
 * Replaces MSOSM_GPU_BATCH / MSOSM_GPU_DM_concurrent / OSM_GPU_BATCH /
 * OSM_GPU_DM_concurrent ::initialize_uint16 device part
 * (msosm_gpu_dm_concurrent.cpp:49-108, msosm_gpu_batch.cpp:32-72). */
int msosm_engine_create(const msosm_plan *plan, void *hip_stream, msosm_engine **out);
int msosm_engine_destroy(msosm_engine *eng);
/* One call: count*M uint16 pairs; where = 0 host pointer, 1 device pointer.
 * Replaces filter_block_uint16 (msosm_gpu_dm_concurrent.cpp:119-151,
 * msosm_gpu_batch.cpp:75-94, osm_gpu_batch.cpp:94-123).  A host buffer may be
 * reused once the call returns; a device buffer must stay unchanged until the
 * engine's stream has run the call.  The kernels run asynchronously on the
 * engine's stream: work on other streams that reads the output must wait for
 * it (msosm_engine_synchronize, or an event on msosm_engine_stream). */
int msosm_engine_filter_u16(msosm_engine *eng, const void *input, int where);
/* Output of the last call, layout [dm][count][M] complex64.
 * Replaces get_output(Complex*) (msosm_gpu_dm_concurrent.cpp:153-165). */
int msosm_engine_output_c64(msosm_engine *eng, void *dst, int where);
/* Same, packed to uint16 pairs (complexToUint16, gpu_kernel.cu:317-329);
 * replaces get_output(uint16_pair*) (msosm_gpu_dm_concurrent.cpp:167-180). */
int msosm_engine_output_u16(msosm_engine *eng, void *dst, int where);
/* Device address of the complex64 output buffer (the reference's public
 * output_buffer_d, msosm_gpu_batch.h:28).  It holds num_dms*count*M + 1
 * elements; the extra element is the reference's sentinel (always 0). */
int msosm_engine_output_ptr(msosm_engine *eng, void **dev_ptr);
/* Filter table H of one DM (complex64, 2M), copied to host (tests). */
int msosm_engine_dedisp_params(msosm_engine *eng, int dm_index, float *host_out);
/* Block until the engine's stream is idle (synchronize()). */
int msosm_engine_synchronize(msosm_engine *eng);
/* Zero the stream history (overlap tail, delay ring, block counter). */
int msosm_engine_reset(msosm_engine *eng);
/* The HIP stream the engine launches on (void* = hipStream_t). */
int msosm_engine_stream(msosm_engine *eng, void **hip_stream);
/* Stream state for checkpoint/resume: the history the next filter call
 * reads -- the overlap tail of the input (the head of the reference's input
 * double buffer, msosm_gpu_batch.cpp:80), the delay ring (products in the
 * reference-structured form, shared spectra otherwise) and the ring cursor
 * (current_index, msosm_gpu_batch.cpp:86,101-108) -- as one byte image of
 * msosm_engine_state_size bytes.  Export waits for the engine's stream; an
 * image imports only into an engine of the same plan (trials, FFT size,
 * blocks per call, method, chirp mode and form), which then continues
 * exactly as the exporting engine would.  where: 0 host, 1 device memory. */
int msosm_engine_state_size(msosm_engine *eng, size_t *bytes);
int msosm_engine_state_export(msosm_engine *eng, void *dst, size_t bytes, int where);
int msosm_engine_state_import(msosm_engine *eng, const void *src, size_t bytes, int where);
/* Named rocTX range on the calling thread (the reference's PUSH_RANGE /
 * POP_RANGE NVTX macros, gpu/include/gpu_common.cuh:16-39).  The library's
 * own work entry points (filter, output, fold, batch launch, state copies)
 * open ranges named after themselves. */
int msosm_range_push(const char *name);
int msosm_range_pop(void);
/* Launch-shape requests of an engine (no reference counterpart; cuFFT picks
 * its own plans).  0 in a field = the built-in rule.
 *   inv_blocks_per_wg: consecutive blocks of one trial that one workgroup of
 *     the inverse transform runs (fused, matrix-core and dense inverses at
 *     4096..16384 points); rule: the largest of 32, 16, ... that leaves at
 *     least 512 workgroups.
 *   hankel_tiles_per_group: 16-trial tiles that share one Hankel GEMM
 *     workgroup's staged delay band; rule: the largest whose band leaves LDS
 *     for three workgroups per CU (reduced until the band fits the LDS).
 *   scratch_chunk_kib: fused forms above 16384 points run their forward
 *     transform over a chunk of windows, and their inverse over a chunk of
 *     (trial, block) units, at a time through an FP32 scratch of about this
 *     size (KiB); rule and upper limit: 128 MiB.
 * get_tuning reports the requests, with the Hankel group size in effect. */
typedef struct msosm_tuning {
    int inv_blocks_per_wg;
    int hankel_tiles_per_group;
    int scratch_chunk_kib;
} msosm_tuning;
int msosm_engine_set_tuning(msosm_engine *eng, const msosm_tuning *tuning);
int msosm_engine_get_tuning(msosm_engine *eng, msosm_tuning *tuning);
/* Per-stage timing with HIP events on the engine stream (0 = off; the
 * host-to-device copy of host input is timed on the engine's copy stream).
 * Stage ids: MSOSM_STAGE_*.  times_ms[i] = summed kernel time of stage i,
 * calls[i] = launches timed.  COPY is the device-side overlap-save history
 * copy of the plain form (the other forms keep that history inside the
 * forward transform), H2D the copy of host input. */
#define MSOSM_STAGE_UNPACK 0
#define MSOSM_STAGE_FFT_FWD 1
#define MSOSM_STAGE_MULTIPLY 2
#define MSOSM_STAGE_FFT_INV 3
#define MSOSM_STAGE_DISCARD 4
#define MSOSM_STAGE_COPY 5
#define MSOSM_STAGE_H2D 6
#define MSOSM_NUM_STAGES 7
int msosm_engine_set_timing(msosm_engine *eng, int on);
int msosm_engine_get_timing(msosm_engine *eng, double *times_ms, long long *calls, int n);

/* Element kernels on device pointers (gpu/include/gpu_kernel.cuh:5-27).
 * stream NULL = default stream. */
int msosm_k_uint16_to_complex(float *dst, const void *src, size_t n, void *stream);    /* :9  */
int msosm_k_complex_to_uint16(void *dst, const float *src, size_t n, void *stream);    /* :10 */
int msosm_k_intensity(const float *a, const float *b, float *out, size_t n, void *stream); /* :11 */
/* ComplexMultiply_concurrent_batch (:22; numDMs=1 is ComplexMultiply_batch). */
int msosm_k_multiply_batch(const float *a, const float *b, int N, float *out, int count, int num_dms,
                           void *stream);
/* ComplexMultiplyDelay_concurrent_batch (:23; numDMs=1 is the batch form). */
int msosm_k_multiply_delay_batch(const float *a, const float *b, int N, float *ring, const int32_t *delay,
                                 int delay_block_size, int index, int count, int num_dms, void *stream);
/* DiscardSamples_concurrent_batch (:24; numDMs=1 is the batch form). */
int msosm_k_discard_batch(const float *a, float *b, int M, int count, int num_dms, void *stream);
/* Batched C2C FFT of power-of-two size n (the build's replacement for the
 * cufftPlanMany/cufftExecC2C call sites, e.g. msosm_gpu_batch.cpp:41-49,84,91).
 * Transform b reads in[b*idist + k], writes out[b*odist + k]; direction -1
 * forward, +1 inverse (unnormalised).  in and out must not overlap. */
int msosm_k_fft_c2c(const float *in, float *out, size_t n, size_t batch, size_t idist, size_t odist,
                    int direction, void *stream);
/* FFT plans (cufftPlanMany / cufftExecC2C / cufftDestroy, e.g.
 * msosm_gpu_batch.cpp:41-49,84,91): a plan fixes the size, batch and row
 * strides, and owns the two-pass scratch (sizes above 2^14 points) that
 * msosm_k_fft_c2c allocates per call; exec transforms in[b*idist + k] to
 * out[b*odist + k], direction -1 forward, +1 inverse (unnormalised).  One
 * plan's execs share its scratch: run them on one stream (or wait between
 * them). */
typedef struct msosm_fft_plan msosm_fft_plan;
int msosm_fft_plan_create(size_t n, size_t batch, size_t idist, size_t odist, msosm_fft_plan **out);
int msosm_fft_plan_exec(msosm_fft_plan *plan, const float *in, float *out, int direction, void *stream);
int msosm_fft_plan_destroy(msosm_fft_plan *plan);
/* As msosm_k_fft_c2c, allocating nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Sizes above
 * 2^14 points use a fixed per-device scratch array of 2^22 complex samples
 * (32 MiB, held by each GPU the library runs on until the process ends)
 * instead of a per-call buffer: batch * n above 2^22 is rejected with
 * MSOSM_ERR_ARG.  The stream must belong to the current device.  Calls that
 * overlap on one GPU share the array and race; the array needs no clearing
 * between calls. */
int msosm_k_fft_c2c_noalloc(const float *in, float *out, size_t n, size_t batch, size_t idist, size_t odist,
                            int direction, void *stream);
/* The multiply-delay of ComplexMultiplyDelay_concurrent_batch (:23,
 * gpu_kernel.cu:160-197) over a shared spectra ring, as per-bin complex GEMMs
 * on the matrix cores: Y[dm][b][n] = H[dm][n] * ring[(g0 + b - d[dm][n]) mod rx][n]
 * (zero when g0 + b < d), for b < count, n < 2^log2n.  form: one of
 * MSOSM_FORM_HANKEL_*.  HANKEL_F16: ring, H and Y are half2 (the engine
 * stores X/N and H/(2M)); BF16 forms: ring, H, Y complex64.  kmax: 1 + the
 * largest delay in the table. */
int msosm_k_hankel_delay_gemm(const void *ring, const void *H, const uint16_t *delay, void *Y, int form,
                              int log2n, int count, int num_dms, long long g0, int rx, int kmax, void *stream);
/* Inverse FFT of rows Y[u][2^log2n] (half2 if y_f16, else complex64), keeping
 * the last 2^(log2n-1) points times scale: out[u][2^(log2n-1)] (the inverse
 * transform and DiscardSamples_concurrent_batch, :24, fused). */
int msosm_k_inverse_discard(const void *Y, float *out, int y_f16, int log2n, size_t units, float scale,
                            void *stream);

/* Fold (folding/include/fold_gpu.h:10-41).  The fold runs on the GPU; bin
 * assignment follows Fold_GPU::fold_data exactly (fold_gpu.cpp:48-173). */
int msosm_fold_create(float period, float fs, unsigned long size, unsigned long time_bin, void *hip_stream,
                      msosm_fold **out);                                     /* ctor :3-27 */
int msosm_fold_destroy(msosm_fold *f);
int msosm_fold_discard_samples(msosm_fold *f, unsigned long nd);             /* :29-33 */
/* Intensity of two polarisations' device outputs (size complex samples each);
 * pol2 may be NULL (single polarisation).  calculate_intensity, :35-43.  Runs
 * on the fold's stream: the engines that produced pol1/pol2 must have
 * finished (or share that stream). */
int msosm_fold_intensity(msosm_fold *f, const void *pol1_dev, const void *pol2_dev);
/* Fold the current intensity buffer (fold_data, :48-173). */
int msosm_fold_fold_data(msosm_fold *f);
/* Average raw bins into time_bin bins (fold_data_bins, :175-194). */
int msosm_fold_bins(msosm_fold *f);
/* Copy results to host: folded[time_bin], raw[period_samples],
 * intensity[size]; any pointer may be NULL. */
int msosm_fold_get(msosm_fold *f, float *folded, float *raw, float *intensity);
int msosm_fold_info(msosm_fold *f, unsigned long *period_samples, unsigned long *time_bin, int *ready);
/* Device address of the intensity buffer (size floats) for callers that
 * produce intensities themselves. */
int msosm_fold_intensity_ptr(msosm_fold *f, void **dev_ptr);
/* Load size intensities from host (where = 0) or device (where = 1) memory
 * into the fold's buffer (the reference's public total_intensity array,
 * fold_gpu.h:31, which fold_data reads). */
int msosm_fold_load_intensity(msosm_fold *f, const float *src, int where);

/* Phase fold of engine output (no reference counterpart:
 * the reference folds on the CPU with integer periods, fold_gpu.cpp:48-173).
 * For each of ndm trials (rows of `stride` complex64 samples at a, and at b
 * for a second polarisation or NULL), sample i with global index n0 + i goes
 * to bin floor(frac((n0+i)/period_samples) * nbins), the phase computed in
 * FP64 on the GPU; value sqrt(|a|^2+|b|^2) (power = 0, the reference's
 * intensity) or |a|^2+|b|^2 (power = 1).  Adds into prof[ndm][nbins] and,
 * when hits != NULL, the per-bin sample counts into hits[nbins] (device
 * pointers; the caller zeroes them).  nbins <= 8192. */
int msosm_fold_phase(const void *a, const void *b, size_t n, int ndm, size_t stride, long long n0,
                     double period_samples, int nbins, int power, float *prof, float *hits, void *stream);

/* Uneven multi-stream batch in one graph launch (no reference
 * counterpart: the reference runs one CUDA stream per engine with
 * host launches per block, msosm_stream/msosm_gpu_batch_stream.cpp:105-149).
 * engines[i] (distinct, created by the caller, timing off) each filter their
 * own resident device input inputs[i] of num_samples[i] packed uint16 pairs
 * (whole calls of count*M; a trailing partial call is ignored).  When
 * period_samples != NULL, every call's output is phase-folded into nbins bins
 * (power, FP64 phase) for streams with period_samples[i] > 0.  The whole
 * pass -- engine resets, all calls, all folds, one graph branch per stream --
 * is captured once into a hipGraph; msosm_batch_launch replays it. */
typedef struct msosm_batch msosm_batch;
int msosm_batch_create(msosm_engine *const *engines, int n, const void *const *inputs,
                       const unsigned long *num_samples, const double *period_samples, int nbins, void *stream,
                       msosm_batch **out);
int msosm_batch_launch(msosm_batch *b);
int msosm_batch_synchronize(msosm_batch *b);
int msosm_batch_info(msosm_batch *b, int *num_streams, size_t *graph_nodes);
/* Host copies of stream i's profiles [num_dms][nbins] and hit counts [nbins]. */
int msosm_batch_profiles(msosm_batch *b, int i, float *prof, float *hits);
int msosm_batch_destroy(msosm_batch *b);

/* Autotuner (no reference counterpart; the reference's callers pick a
 * per-stream fftpoint by hand, msosm_stream/msosm_gpu_batch_stream.cpp:33-103).
 * Among powers of two in [min_fftpoint, max_fftpoint] that are not below the
 * plan's minimum, and for each size among the launch shapes its form offers
 * (msosm_tuning: inverse blocks per workgroup at 4096..16384 points, Hankel
 * tiles per group, scratch chunk above 16384 points), time `calls` filter
 * calls of about target_samples samples on synthetic input and return the
 * fastest in DM-trial*samples per second (cached per descriptor for the life
 * of the process).  The form is desc->form where it runs at that size; the
 * matrix-core and Hankel forms (16384 points at most) run as FUSED_F16 above,
 * the fused forms as PLAIN_F32 above 2^21. */
typedef struct msosm_autotune_result {
    unsigned long fftpoint;
    int form;                /* the form the chosen size runs with           */
    msosm_tuning tuning;     /* launch shapes to give msosm_engine_set_tuning */
    double ns_per_dm_sample;
    int candidates;          /* configurations timed                         */
} msosm_autotune_result;
int msosm_autotune(const msosm_plan_desc *desc, unsigned long min_fftpoint, unsigned long max_fftpoint,
                   unsigned long target_samples, int calls, void *stream, msosm_autotune_result *result);
/* The same, returning only the size, its time and its form. */
int msosm_autotune_fftpoint(const msosm_plan_desc *desc, unsigned long min_fftpoint, unsigned long max_fftpoint,
                            unsigned long target_samples, int calls, void *stream, unsigned long *best_fftpoint,
                            double *best_ns_per_dm_sample, int *best_form);

/* ---- learned DM proposal (no reference counterpart:
 * the reference dedisperses every trial of a grid,
 * test_processing_speed_msosm_dmtrial.cpp) -------------------------------
 * Device pointers throughout; activations are FP32 [B][C][T], convolution
 * weights FP16 [Cout][Cin][KW], linear weights FP16 [O][I]. */
/* Synthetic pulsar without noise: x[i] = sqrt(amp g(phase_i)) z_i, z ~ CN(0,1),
 * Gaussian pulse of FWHM duty (fraction of the period), phase in FP64. */
int msosm_sim_pulsar(void *x_c64, size_t n, double period_samples, double duty, float amp, unsigned int seed,
                     void *stream);
/* X[b] *= exp(sign i phi(f_b)) over the bins of an n-point spectrum. */
int msosm_sim_disperse(void *X_c64, size_t n, double fs, double f0, double dm, int sign, void *stream);
/* u16 pairs = clip(round(scale (x x_scale + CN(0,1) noise)) + 32768, 1, 65535). */
int msosm_sim_quantize(const void *x_c64, size_t n, float x_scale, unsigned int seed, float scale, void *u16_out,
                       void *stream);
/* Power spectrogram out[nchan][T], T = n / (nchan navg): navg consecutive
 * nchan-point spectra per bin (the library supports nchan 256, navg 16), channel 0 =
 * lowest frequency. */
int msosm_spectrogram_u16(const void *u16, size_t n, int nchan, int navg, float *out, void *stream);
int msosm_rows_standardize(float *x, int rows, int T, void *stream);
/* out[k][t] = sum_c S[c][t + delay[k][c]] / sqrt(nchan), t < Tout. */
int msosm_incoherent_dedisperse(const float *S, int nchan, int T, const int *delay, int K, int Tout, float *out,
                                void *stream);
/* out[r][b] = scale * mean of x[r][t] over samples t with floor(frac(t /
 * period) * nbins) = b (period in samples, FP64 phase); 0 where no sample
 * falls. */
int msosm_rows_fold(const float *x, int rows, int T, double period, int nbins, float scale, float *out,
                    void *stream);
/* Folded profiles of K trials from channels folded at `period` seconds:
 * out[k][b] = sum_c Sf[c][(b + rint(delay_s[k][c] / period * nbins)) mod
 * nbins] / sqrt(nchan) (delay_s in seconds). */
int msosm_incoherent_fold(const float *Sf, int nchan, int nbins, const float *delay_s, int K, double period,
                          float *out, void *stream);
/* out[k][r] = |DFT_nbins(prof[r])[k + 1]| / sqrt(nbins), k < K <= nbins/2:
 * harmonic amplitudes of folded profiles, harmonic-major. */
int msosm_rows_harmonics(const float *prof, int rows, int nbins, int K, float *out, void *stream);
int msosm_conv1d_fwd(const float *x, const void *w16, const float *bias, float *y, int B, int Cin, int Cout, int T,
                     int KW, int dil, int relu, void *stream);
/* gx (may be NULL) = data gradient; gw, gb are accumulated (+=). */
int msosm_conv1d_bwd(const float *x, const void *w16, const float *y, const float *gy, float *gx, float *gw,
                     float *gb, int B, int Cin, int Cout, int T, int KW, int dil, int relu, void *stream);
/* feat[B][2C] = (max_t, mean_t); amax[B][C] = argmax used by the backward. */
int msosm_pool_fwd(const float *x, int B, int C, int T, float *feat, int *amax, void *stream);
int msosm_pool_bwd(const float *gfeat, const int *amax, int B, int C, int T, float *gx, void *stream);
int msosm_linear_fwd(const float *x, const void *w16, const float *bias, float *y, int B, int I, int O, void *stream);
int msosm_linear_bwd(const float *x, const void *w16, const float *gy, float *gx, float *gw, float *gb, int B, int I,
                     int O, void *stream);
/* loss (accumulated, +=) = mean_b cross-entropy(softmax(logits_b), target_b);
 * glogits = (softmax - target) / B. */
int msosm_softmax_xent(const float *logits, const float *target, int B, int O, float *loss, float *glogits,
                       void *stream);
/* Adam step on FP32 master weights p, refreshing the FP16 copy p16 (may be NULL). */
int msosm_adam(float *p, void *p16, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
               float eps, int step, void *stream);
int msosm_f32_to_f16(const float *a, void *b, size_t n, void *stream);

/* Output streaming to a host consumer thread (replaces the single host
 * buffer + sentinel element + spin flags of get_output /
 * get_output_thread, msosm_gpu_dm_concurrent.cpp:153-195, and the writer
 * threads of multiple_dm_concurrent_msosm.cpp:84-146).  A ring of `slots`
 * slots of slot_bytes, each a device staging buffer and a page-locked host
 * buffer.  push (producer): the engine's last output (format 0 complex64,
 * 1 uint16 as complexToUint16) is copied to the next free slot -- on the
 * engine's stream into the staging buffer, so the engine's next call may
 * start at once, then to the host on the ring's own stream; blocks while no
 * slot is free (timeout_ms < 0: no limit), *seq = 0, 1, 2, ...  pop
 * (consumer): waits for the oldest pushed slot's host copy, returns its data,
 * byte count and sequence number; release frees it.  One producer and one
 * consumer thread. */
typedef struct msosm_outring msosm_outring;
int msosm_outring_create(size_t slot_bytes, int slots, msosm_outring **out);
int msosm_outring_destroy(msosm_outring *r);
int msosm_outring_push(msosm_outring *r, msosm_engine *eng, int format, int timeout_ms, long long *seq);
int msosm_outring_pop(msosm_outring *r, int timeout_ms, const void **data, size_t *bytes, long long *seq);
int msosm_outring_release(msosm_outring *r);

/* The simulated test signal of the reference's drivers (SimulatedComplexSignal
 * + AWGN, simulated_pulsar/simulated_complex_signal.cpp:12-236, AWGN.cpp:21-58):
 * two unit impulses per period of Np = period*bw samples, dispersed at dm
 * (spread over `range` periods; transforms of fftpoint points), repeated
 * `repeat` times, optionally plus complex Gaussian noise at snr_db against
 * the measured signal power.  signal: host, 2*repeat*Np floats (may be NULL);
 * signal_u16: host, repeat*Np uint16 pairs (may be NULL).  Transforms on the
 * GPU. */
int msosm_refsim_info(float bw, float dm, float f0, float period, unsigned long *Np, unsigned long *Nd,
                      unsigned long *range, unsigned long *fftpoint);
int msosm_refsim_generate(float bw, float dm, float f0, float period, unsigned long repeat, int add_noise,
                          float snr_db, float *signal, void *signal_u16, float *signal_power_db);
/* The noise alone (host only): n complex samples as 2n floats. */
int msosm_refsim_awgn(float snr_db, unsigned long n, float *out);

/* Page-locked host memory for input and output buffers (replaces the
 * reference drivers' cudaHostRegister of their buffers, e.g.
 * test_processing_speed_msosm.cpp:182,196, which has no counterpart on this
 * platform for existing allocations).  msosm_host_free(NULL) is a no-op. */
int msosm_host_alloc(size_t bytes, void **out);
int msosm_host_free(void *p);

const char *msosm_last_error(void);
const char *msosm_version(void);

#ifdef __cplusplus
}
#endif
#endif
