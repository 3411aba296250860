// mfma_fft.h -- index maps of the matrix-core inverse FFT of 2^14 points
// (invmfma.hip), shared with the forward kernel that fills its spectra ring.
//
// The transform is a Stockham FFT with passes of radix 32, 32 and 16 whose
// DFTs run as v_mfma_f32_16x16x32_f16 products: a pass takes 16 columns
// (butterflies) at a time, lane (q, c) = (lane >> 4, lane & 15) holding
// column c's elements 16 kt + 4 i + q (i < 4) of K-tile kt as four
// interleaved FP16 (re, im) pairs -- the B operand, with the DFT matrix's
// columns in the same order -- and receiving outputs 8 mt + 2 q + e (e < 2)
// of M-tile mt as FP32 (re, im) pairs.  The four quarters of a wave hold
// neighbouring rows, so pass 0's gathers of one instruction span 4 rows
// (bins 512 apart) rather than 16.
//
// Pass 0 reads its inputs straight from the spectra ring, which stores bin n
// of a block at position ring_pos(n): the order in which pass 0's lanes gather
// them (group g = 0..31, K-tile, element i, lane), so every gather
// instruction reads 64 consecutive 4-byte words.
#pragma once
#include <cstdint>

namespace msosm {
namespace mfft {

constexpr int kLog2N = 14;
constexpr int kN = 1 << kLog2N;
constexpr int kM = kN / 2;

// Pass 0 column of group g (0..31) and lane column c: j = 256 (g >> 4) +
// 16 c + (g & 15); its row r is bin j + 512 r.
__host__ __device__ constexpr int p0_column(int g, int c) { return 256 * (g >> 4) + 16 * c + (g & 15); }

// Position of bin n in the ring (and in the permuted filter and delay tables).
__host__ __device__ constexpr int ring_pos(int n)
{
    const int j = n & 511, r = n >> 9;
    const int g = 16 * (j >> 8) + (j & 15), c = (j >> 4) & 15;
    const int kt = r >> 4, i = (r >> 2) & 3, q = r & 3;
    return (((g * 2 + kt) * 4 + i) * 64) + q * 16 + c;
}

// Inverse of ring_pos: the bin stored at position p.
__host__ __device__ constexpr int ring_bin(int p)
{
    const int lane = p & 63, i = (p >> 6) & 3, kt = (p >> 8) & 1, g = p >> 9;
    const int q = lane >> 4, c = lane & 15;
    return p0_column(g, c) + 512 * (16 * kt + 4 * i + q);
}

// LDS layouts (in 4-byte half2 units) of the pass 1 input A1[j'][r'] (512
// columns of 32 rows) and the pass 2 input A2[j''][r''] (1024 columns of 16
// rows): the four rows a lane's B operand holds (16 kt + 4 i + q, i < 4)
// are one 16-byte chunk (one ds_read_b128), and the XOR swizzles make both
// the 16-byte reads of 16 consecutive lanes and the 4-byte writes of a
// wave's 64 lanes hit distinct banks.
__host__ __device__ constexpr int a1_addr(int jp, int rp)
{
    const int col = jp ^ ((jp >> 8) & 1) ^ ((jp >> 2) & 1);
    const int ch = (4 * (rp >> 4) + (rp & 3)) ^ ((jp >> 5) & 7) ^ (((jp >> 1) & 1) << 2);
    return col * 32 + ch * 4 + ((rp >> 2) & 3);
}
__host__ __device__ constexpr int a2_addr(int jpp, int rpp)
{
    const int col = jpp ^ ((jpp >> 6) & 3);
    const int ch = (rpp & 3) ^ ((jpp >> 2) & 3);
    return col * 16 + ch * 4 + (rpp >> 2);
}

} // namespace mfft
} // namespace msosm
