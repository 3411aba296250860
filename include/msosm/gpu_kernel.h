// gpu_kernel.h -- drop-in facade for the reference's kernel launchers
// (gpu/include/gpu_kernel.cuh:5-27) over libmsosm.so.  Pointers are device
// pointers; the stream argument takes a hipStream_t (nullptr = default).
#pragma once
#include "globals.h"

inline void uint16ToComplex(Complex *dst, uint16_t *src, size_t size, void *stream = nullptr)
{
    msosm_facade::check(msosm_k_uint16_to_complex(&dst->x, src, size, stream), "uint16ToComplex");
}
inline void complexToUint16(uint16_t *dst, Complex *src, size_t size, void *stream = nullptr)
{
    msosm_facade::check(msosm_k_complex_to_uint16(dst, &src->x, size, stream), "complexToUint16");
}
inline void calculateIntensity(Complex *a, Complex *b, float *total_intensity, unsigned long size,
                               void *stream = nullptr)
{
    msosm_facade::check(msosm_k_intensity(&a->x, b ? &b->x : nullptr, total_intensity, size, stream),
                        "calculateIntensity");
}
inline void ComplexMultiply_batch(Complex *a, Complex *b, int N, Complex *block, int count, void *stream = nullptr)
{
    msosm_facade::check(msosm_k_multiply_batch(&a->x, &b->x, N, &block->x, count, 1, stream), "ComplexMultiply");
}
inline void ComplexMultiply_concurrent_batch(Complex *a, Complex *b, int N, Complex *block, int count, int numDMs,
                                             void *stream = nullptr)
{
    msosm_facade::check(msosm_k_multiply_batch(&a->x, &b->x, N, &block->x, count, numDMs, stream),
                        "ComplexMultiply_concurrent");
}
inline void ComplexMultiplyDelay_batch(Complex *a, Complex *b, int N, Complex *block, int *delay,
                                       int delay_block_size, int index, int count, void *stream = nullptr)
{
    msosm_facade::check(msosm_k_multiply_delay_batch(&a->x, &b->x, N, &block->x, delay, delay_block_size, index,
                                                     count, 1, stream),
                        "ComplexMultiplyDelay");
}
inline void ComplexMultiplyDelay_concurrent_batch(Complex *a, Complex *b, int N, Complex *block, int *delay,
                                                  int delay_block_size, int index, int count, int numDMs,
                                                  void *stream = nullptr)
{
    msosm_facade::check(msosm_k_multiply_delay_batch(&a->x, &b->x, N, &block->x, delay, delay_block_size, index,
                                                     count, numDMs, stream),
                        "ComplexMultiplyDelay_concurrent");
}
inline void DiscardSamples_batch(Complex *a, Complex *b, int M, int count, void *stream = nullptr)
{
    msosm_facade::check(msosm_k_discard_batch(&a->x, &b->x, M, count, 1, stream), "DiscardSamples");
}
inline void DiscardSamples_concurrent_batch(Complex *a, Complex *b, int M, int count, int numDMs,
                                            void *stream = nullptr)
{
    msosm_facade::check(msosm_k_discard_batch(&a->x, &b->x, M, count, numDMs, stream), "DiscardSamples_concurrent");
}
