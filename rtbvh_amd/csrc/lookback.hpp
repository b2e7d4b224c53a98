// Decoupled look-back over tile words, shared by the builders' single-pass
// scans (HIP device code for gfx950).  A tile's id comes from a counter that
// running blocks take, and a tile waits only on lower ids, whose blocks are
// already running: every wait ends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtbvh {

// Look-back tile words: bits 63:62 the status (0 not yet written, 1 the
// tile's own aggregate, 2 the inclusive prefix through the tile), the rest
// the payload.  Agent-scope atomics: tiles run on every XCD.
constexpr uint64_t kAggregate = 1ull << 62;
constexpr uint64_t kInclusive = 2ull << 62;
constexpr uint64_t kPayload = (1ull << 62) - 1;

__device__ __forceinline__ void publish(uint64_t* w, uint64_t v) {
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t wait_word(uint64_t* w) {
    uint64_t v;
    while (((v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 62) == 0)
        __builtin_amdgcn_s_sleep(1);
    return v;
}

// Wave-wide sum (all 64 lanes take part).
__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Exclusive prefix of tile `tile`: the payloads of the words before it down
// to the nearest inclusive one, by the 64 lanes of one wave (lane k looks at
// word end - 1 - k of each window of 64).
__device__ inline uint64_t lookback(uint64_t* words, uint32_t tile, uint32_t lane) {
    uint64_t excl = 0;
    int64_t end = tile;
    while (end > 0) {
        const int64_t p = end - 1 - int64_t(lane);
        const uint64_t w = p >= 0 ? wait_word(&words[p]) : kInclusive;
        const uint64_t inc = __ballot((w >> 62) == 2);
        const uint32_t first = inc ? uint32_t(__builtin_ctzll(inc)) : 64u;
        excl += wave_sum(lane <= first && p >= 0 ? (w & kPayload) : 0);
        if (inc) break;
        end -= 64;
    }
    return excl;
}

}  // namespace rtbvh
