/*
 * totals.h — FP64 sums over FluidInterior nodes of the five conserved fields (conservation monitor), shared by
 * ns_totals (the handle's state in the device layout) and ns_totals_noalloc (caller-held dense fields).
 *
 * Nodes are visited in dense x-fastest order n = (k*NJ + j)*NI + i by a grid-stride loop of kTotalsBlocks
 * blocks of kTotalsThreads threads, whatever the storage layout: both entry points then form the same
 * per-block partials, and summing them in block order gives the same bits.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_common.h"

namespace nsfsi {

constexpr int kTotalsBlocks = 256;
constexpr int kTotalsThreads = 256;

// Partial sums of block blockIdx.x into partials[5 * blockIdx.x + q]. offOf(n): element offset of node n in
// field 0 (field q at + q * fstride) and in cls; cls holds the node class in its low three bits.
template <class T, class OffOf>
__device__ __forceinline__ void totalsBlock(size_t N, OffOf offOf, size_t fstride, const T* S, const uint8_t* cls,
                                            double* partials) {
    __shared__ double sh[kTotalsThreads / 64];
    double sm[5] = {0, 0, 0, 0, 0};
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (size_t)gridDim.x * blockDim.x) {
        const size_t o = offOf(n);
        if ((cls[o] & CLS_TYPE_MASK) != NS_FLUID_INTERIOR) continue;
        for (int q = 0; q < 5; ++q) sm[q] += (double)S[o + q * fstride];
    }
    for (int q = 0; q < 5; ++q) {
        double v = wave_sum(sm[q]);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        __syncthreads();
        if (lane == 0) sh[wv] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0;
            for (int b = 0; b < (int)(blockDim.x >> 6); ++b) r += sh[b];
            partials[(size_t)blockIdx.x * 5 + q] = r;
        }
    }
}

}  // namespace nsfsi
