// Block-wide exclusive scans from wave scans (HIP device code for gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtbvh {

// Inclusive scan over the 64 lanes of a wave (shuffles, no barrier).
__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= uint32_t(off)) v += t;
    }
    return v;
}

// Exclusive scan over a block of N threads: wave scans, then the wave
// totals through sh (N / 64 words); *total (when not null) gets the sum.
// Two barriers; sh may be reused as soon as it returns.
template <int N>
__device__ __forceinline__ uint32_t block_exclusive_n(uint32_t v, uint32_t* sh, uint32_t* total) {
    constexpr int kWaves = N / 64;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive(v);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t t = sh[w];
        before += w < int(wave) ? t : 0u;
        all += t;
    }
    if (total) *total = all;
    __syncthreads();
    return before + incl - v;
}

}  // namespace rtbvh
