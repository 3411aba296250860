// kernels.h -- host launchers of the HIP kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace msosm {

constexpr int kMaxLog2SinglePass = 14; // one workgroup holds 2^14 complex in LDS

// Batched C2C FFT, any power of two 2^1..2^28 (one pass up to 2^14, two
// passes above).  scale multiplies the output.  scratch: for two-pass sizes a
// buffer of batch*n float2 (NULL: the launcher allocates).
hipError_t launch_fft_c2c(const float2 *in, float2 *out, int log2n, long long batch, long long idist,
                          long long odist, bool inverse, float scale, float2 *scratch, hipStream_t s);
size_t fft_c2c_scratch_elems(int log2n, long long batch);

// Fused forward FFT of count overlapped windows read straight from packed
// uint16 input (cur) and the previous call's last M samples (tail), written
// to slots (g0 + w) mod rx of a spectra ring (FP32, or FP16 scaled by 1/N);
// this call's last M samples go to tail_out (a different buffer from tail);
// mfma_order (FP16, 2^14 points only): bin n at position mfft::ring_pos(n).
hipError_t launch_fwd_u16(const uint32_t *cur, const uint32_t *tail, uint32_t *tail_out, void *ring, bool f16,
                          int log2n, int count, long long g0, int rx, bool mfma_order, hipStream_t s);
// Fused inverse: per (DM, block) gather X_{g-d(n)}[n] * H[n] from the ring,
// inverse FFT, keep the last M samples, scale; out layout [dm][count][M].
// dpk: for sizes with one transform per workgroup (fused_uses_packed_delay),
// the delays in thread-major packed layout dpk[dm][j][t] = d(t + 2j*NT) |
// d(t + (2j+1)*NT) << 16 with NT = fused_delay_threads(log2n); else unused.
// Blocks per workgroup for the inverse kernels that run consecutive blocks
// of one trial per workgroup: the largest power of two up to max_nbpw that
// leaves at least 512 workgroups for rows x count blocks, else 1.
int blocks_per_wg(int count, long long rows, int max_nbpw);
// CUs of the calling thread's current device (cached per device; 256 if the
// query fails)
int device_cu_count();
// Transforms per workgroup for a launch of `items` 2^14-point transforms run
// one workgroup per CU at a time (their LDS): one while they fit in one round
// of workgroups, else runs that spread them over the CUs once.
int runs_per_wg(long long items);
// nbpw: blocks per workgroup for sizes with one transform per workgroup
// (0 = blocks_per_wg(count, ndm, 32)).
hipError_t launch_inv_fused(const void *ring, const void *H, const uint16_t *delay, const uint32_t *dpk, float2 *out,
                            bool f16, int log2n, int count, int ndm, long long g0, int rx, int nbpw, hipStream_t s);
int fused_delay_threads(int log2n);
bool fused_uses_packed_delay(int log2n);
bool fused_supported(int log2n);

// Shared-ring forms at 2^15 .. 2^21 points (bigfft.hip): four-step transforms
// with rows of 2^14 points.  The spectra ring, the filter table and the delay
// table hold bin big_position_bin(p) at position p.  launch_big_permute
// writes rows of natural-order complex64 src into position order (FP16 or
// FP32, times scale).  Forward: count windows, chunk at a time through a
// scratch of chunk << log2n float2, into ring slots (g0 + w) mod rx;
// inverse: all (trial, block) units, chunk at a time through a scratch of
// chunk << log2n float2; output [dm][count][M] as the other fused forms.
// big_chunk: how many of `items` windows or units fit `bytes` of scratch
// (0: the default of 128 MiB), at least 1.
bool big_supported(int log2n);
long long big_position_bin(int log2n, long long p);
int big_chunk(int log2n, long long items, size_t bytes);
hipError_t launch_big_permute(const float2 *src, void *dst, bool f16, float scale, int log2n, long long rows,
                              hipStream_t s);
hipError_t launch_big_forward(const uint32_t *cur, const uint32_t *tail, uint32_t *tail_out, void *ring,
                              float2 *scratch, int chunk, bool f16, int log2n, int count, long long g0, int rx,
                              hipStream_t s);
hipError_t launch_big_inverse(const void *ring, const void *Hp, const uint16_t *dp, float2 *scratch, int chunk,
                              float2 *out, bool f16, int log2n, int count, int ndm, long long g0, int rx,
                              hipStream_t s);

// Fused inverse with the DFT stages on the matrix cores (invmfma.hip), 2^14
// points: ring, Hp and dpk in the gather order of mfft::ring_pos (Hp from
// launch_inv_mfma_filter; dpk[dm][pp] = d(p0) | d(p0 + 64) << 16 with
// p0 = ((pp >> 7) * 4 + 2 ((pp >> 6) & 1)) * 64 + (pp & 63)); tables: a
// device buffer of inv_mfma_tables_bytes() filled once by
// launch_inv_mfma_tables.
size_t inv_mfma_tables_bytes();
hipError_t launch_inv_mfma_tables(void *tables, hipStream_t s);
hipError_t launch_inv_mfma(const __half2 *ring, const __half2 *Hp, const uint32_t *dpk, const void *tables,
                           float2 *out, int count, int ndm, long long g0, int rx, int nbpw, hipStream_t s);
// Filter rows for launch_inv_mfma from the FP16 rows H/(2M): dst[r][ring_pos(n)]
// = src[r][n] / 64 (FP16 headroom; the kernel's last DFT carries the 64)
hipError_t launch_inv_mfma_filter(__half2 *dst, const __half2 *src, long long rows, hipStream_t s);

// Delay-line accumulate as per-bin complex GEMMs on MFMA (hankel.hip):
// Y[dm][b][n] = H[dm][n] X_{g0+b-d(n)}[n] computed over the band of delays a
// tile of 16 trials uses at bin n.  Modes: FP16 operands (ring X/N, H/(2M);
// Y FP16 = product), BF16 operands from an FP32 ring and H (Y FP32 =
// product), BF16 hi+lo split with three products.  kmax: largest delay + 1.
constexpr int HANKEL_F16 = 0;
constexpr int HANKEL_BF16 = 1;
constexpr int HANKEL_BF16X3 = 2;
constexpr size_t kHankelMaxLds = 80 * 1024;
hipError_t launch_hankel(const void *ring, const void *H, const uint16_t *delay, void *Y, int mode, int log2n,
                         int count, int ndm, long long g0, int rx, int kmax, int tpg, int kspan, const int *band_lo,
                         hipStream_t s);
// Trial grouping of launch_hankel from the host delay table [ndm][N]: tpg DM
// tiles of 16 trials per workgroup (grid z), the widest delay band kspan any
// workgroup stages, and band_lo[group][bin group] its lowest delay (a device
// copy goes to launch_hankel).  band_lo null stages the whole delay line.
// tpg_req > 0 asks for that many tiles per group (tuning), 0 = the rule.
void hankel_groups(int mode, const int32_t *delay, int ndm, int N, int kmax, int tpg_req, int *tpg, int *kspan,
                   std::vector<int> *band_lo);
int hankel_max_kmax(int mode);
int hankel_bins_per_wg(int mode);
// Inverse FFT of dense rows Y[unit][2^log2n] (FP16 or FP32 complex), keep the
// last half, times scale: out[unit][2^(log2n-1)].
// nbpw: consecutive rows per workgroup at 2^12..2^14 points (0 = the rule).
hipError_t launch_inv_dense(const void *Y, float2 *out, bool f16y, int log2n, long long units, float scale,
                            int nbpw, hipStream_t s);

// Element kernels (reference gpu_kernel.cu semantics).
hipError_t launch_unpack(float2 *dst, const uint32_t *src, size_t n, hipStream_t s);
hipError_t launch_pack(uint32_t *dst, const float2 *src, size_t n, hipStream_t s);
hipError_t launch_intensity(const float2 *a, const float2 *b, float *out, size_t n, hipStream_t s);
hipError_t launch_multiply(const float2 *a, const float2 *b, int N, float2 *out, int count, int ndm, hipStream_t s);
hipError_t launch_multiply_delay(const float2 *a, const float2 *b, int N, float2 *ring, const int32_t *delay,
                                 int dbs, int index, int count, int ndm, hipStream_t s);
hipError_t launch_discard(const float2 *a, float2 *b, int M, int count, int ndm, hipStream_t s);

// Filter-table generation: chirp samples (FP64 phase) in IDFT input order.
struct ChirpParams {
    float fs, step, w0, kdm, dm, f0;
};
hipError_t launch_chirp(float2 *c, unsigned long M, int mode, const ChirpParams *params_dev, int ndm,
                        hipStream_t s);
// dst[d][0..2M) = (src[d][0..M), 0...)
hipError_t launch_zero_pad(float2 *dst, const float2 *src, unsigned long M, int ndm, hipStream_t s);
// dst = half2(src * scale)
hipError_t launch_to_half(__half2 *dst, const float2 *src, size_t n, float scale, hipStream_t s);
hipError_t launch_i32_to_u16(uint16_t *dst, const int32_t *src, size_t n, hipStream_t s);
// Capturable replacements for hipMemsetAsync / device-to-device copies
// (graph replays stay correct with kernel nodes).
hipError_t launch_fill_u32(void *dst, size_t words, uint32_t value, hipStream_t s);
hipError_t launch_copy_u32(void *dst, const void *src, size_t words, hipStream_t s);
// Deterministic pseudo-random packed uint16 samples (offset binary around
// 32768, never 0) for timing runs.
hipError_t launch_fill_random_u16(uint32_t *dst, size_t n, uint32_t seed, hipStream_t s);

} // namespace msosm
