// Rays traced in the order of their origins (no counterpart in the
// reference, which traces rays in the caller's order, one packet at a time).
//
// A large set of rays in random order — shadow rays from random surface
// points, say — gives every wavefront 64 unrelated paths through the tree:
// the lanes diverge at once and every node a wave loads serves one lane.
// Grouping the rays by the cell of their origin in a grid over the origins'
// bounds (cells in Morton order; 2^b cells per axis, b from 2 to 8, so that
// a cell holds 4 to 16 rays on average up to 2^28 rays) puts rays that start
// close together into the same wavefront and the same stretch of time, so
// that rays towards one light, or out of one surface patch, walk the same
// nodes.
// A ray's result does not depend on the order in which the rays are traced,
// so the results are exactly those of an unordered traversal; they are
// written back in the caller's order.
//
// Passes (stream-ordered, no host reads, no allocation, capturable in a
// graph):
//   bounds_part   per-block min/max over a sample of at most 2^20 origins
//                 spread evenly over the rays (the grid's frame is a
//                 grouping heuristic: rays outside it fall into the border
//                 cells)
//   bounds_final  one block reduces the parts to the frame (lo, 2^b / extent)
//   keys          (cell of ray i, i) for every ray
//   radix sort    stable, by the 3b bits of the cells, 8 a pass (the
//                 Morton-code sort's kernels); the sorted values are the
//                 caller's ray index of each slot, ids[j]
// Compressed 8-wide trees trace through ids (the traversal takes position
// j's ray from rays[ids[j]] and writes its result at ids[j]). For 4-wide
// trees:
//   copy          sorted[j] = rays[ids[j]]
//   (the traversal of sorted[] into results[])
//   gather        out[ids[j]] = results[j]
#include <math.h>

#include "internal.hpp"

namespace rtbvh {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMinAxisBits = 2, kMaxAxisBits = 8;
constexpr uint32_t kParts = 1024;      // bounds_part blocks
constexpr uint32_t kSampleRounds = 4;  // origins per thread of bounds_part

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// 2^b cells per axis for n rays: about 4 to 16 rays per cell, b in [2, 8]
uint32_t axis_bits(uint64_t n) {
    uint32_t lg = 0;
    while (lg < 63 && (uint64_t(1) << lg) < n) ++lg;
    const uint32_t b = lg >= 4 ? (lg - 2) / 3 : 0;
    return b < kMinAxisBits ? kMinAxisBits : (b > kMaxAxisBits ? kMaxAxisBits : b);
}

// bits 0..9 of x to bits 0, 3, 6, ..., 27
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// frame: lo x, y, z, then cells per unit length along x, y, z (0 for an
// axis of zero or unbounded extent), then the cells per axis.  NaN and
// infinite coordinates land in a border cell (fmaxf returns its non-NaN
// operand).
__device__ __forceinline__ uint32_t origin_cell(const float4& o, const float* frame) {
    const float top = frame[6] - 1.f;
    const uint32_t qx = uint32_t(fminf(fmaxf((o.x - frame[0]) * frame[3], 0.f), top));
    const uint32_t qy = uint32_t(fminf(fmaxf((o.y - frame[1]) * frame[4], 0.f), top));
    const uint32_t qz = uint32_t(fminf(fmaxf((o.z - frame[2]) * frame[5], 0.f), top));
    return spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
}

__device__ __forceinline__ float wave_min(float x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fminf(x, __shfl_xor(x, off));
    return x;
}

__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off));
    return x;
}

// parts[b * 6 + k]: min x, y, z, max x, y, z of block b's sample.
__global__ void __launch_bounds__(kBlock)
    bounds_part_kernel(const float4* rays, uint64_t n, uint64_t step, float* parts) {
    float mn0 = INFINITY, mn1 = INFINITY, mn2 = INFINITY;
    float mx0 = -INFINITY, mx1 = -INFINITY, mx2 = -INFINITY;
    for (uint32_t r = 0; r < kSampleRounds; ++r) {
        const uint64_t j = uint64_t(r) * kParts * kBlock + blockIdx.x * kBlock + threadIdx.x;
        const uint64_t i = j * step;
        if (i < n) {
            const float4 o = rays[2 * i];
            mn0 = fminf(mn0, o.x);
            mn1 = fminf(mn1, o.y);
            mn2 = fminf(mn2, o.z);
            mx0 = fmaxf(mx0, o.x);
            mx1 = fmaxf(mx1, o.y);
            mx2 = fmaxf(mx2, o.z);
        }
    }
    __shared__ float red[kBlock / 64][6];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    mn0 = wave_min(mn0);
    mn1 = wave_min(mn1);
    mn2 = wave_min(mn2);
    mx0 = wave_max(mx0);
    mx1 = wave_max(mx1);
    mx2 = wave_max(mx2);
    if (lane == 0) {
        red[wave][0] = mn0;
        red[wave][1] = mn1;
        red[wave][2] = mn2;
        red[wave][3] = mx0;
        red[wave][4] = mx1;
        red[wave][5] = mx2;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t k = threadIdx.x;
        float v = red[0][k];
        for (int w = 1; w < kBlock / 64; ++w)
            v = k < 3 ? fminf(v, red[w][k]) : fmaxf(v, red[w][k]);
        parts[blockIdx.x * 6 + k] = v;
    }
}

__global__ void __launch_bounds__(kBlock)
    bounds_final_kernel(const float* parts, float axis_cells, float* frame) {
    float mn0 = INFINITY, mn1 = INFINITY, mn2 = INFINITY;
    float mx0 = -INFINITY, mx1 = -INFINITY, mx2 = -INFINITY;
    for (uint32_t b = threadIdx.x; b < kParts; b += kBlock) {
        const float* q = parts + b * 6;
        mn0 = fminf(mn0, q[0]);
        mn1 = fminf(mn1, q[1]);
        mn2 = fminf(mn2, q[2]);
        mx0 = fmaxf(mx0, q[3]);
        mx1 = fmaxf(mx1, q[4]);
        mx2 = fmaxf(mx2, q[5]);
    }
    __shared__ float red[kBlock / 64][6];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    mn0 = wave_min(mn0);
    mn1 = wave_min(mn1);
    mn2 = wave_min(mn2);
    mx0 = wave_max(mx0);
    mx1 = wave_max(mx1);
    mx2 = wave_max(mx2);
    if (lane == 0) {
        red[wave][0] = mn0;
        red[wave][1] = mn1;
        red[wave][2] = mn2;
        red[wave][3] = mx0;
        red[wave][4] = mx1;
        red[wave][5] = mx2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint32_t k = threadIdx.x;
        float lo = red[0][k], hi = red[0][k + 3];
        for (int w = 1; w < kBlock / 64; ++w) {
            lo = fminf(lo, red[w][k]);
            hi = fmaxf(hi, red[w][k + 3]);
        }
        const float ext = hi - lo;
        const bool usable = isfinite(lo) && isfinite(ext) && ext > 0.f;
        frame[k] = usable ? lo : 0.f;
        frame[k + 3] = usable ? axis_cells / ext : 0.f;
    }
    if (threadIdx.x == 0) frame[6] = axis_cells;
}

__global__ void __launch_bounds__(kBlock)
    keys_kernel(const float4* rays, uint64_t n, const float* frame, uint32_t* keys,
                uint32_t* vals) {
    const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    keys[i] = origin_cell(rays[2 * i], frame);
    vals[i] = uint32_t(i);
}

__global__ void __launch_bounds__(kBlock)
    copy_kernel(const float4* rays, const uint32_t* ids, uint64_t n, float4* sorted) {
    const uint64_t j = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint64_t i = ids[j];
    sorted[2 * j] = rays[2 * i];
    sorted[2 * j + 1] = rays[2 * i + 1];
}

typedef float HitVec __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kBlock)
    gather_hits_kernel(const HitVec* results, const uint32_t* ids, uint64_t n, HitVec* hits) {
    const uint64_t j = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    __builtin_nontemporal_store(results[j], hits + ids[j]);
}

__global__ void __launch_bounds__(kBlock)
    gather_occluded_kernel(const uint8_t* results, const uint32_t* ids, uint64_t n,
                           uint8_t* occluded) {
    const uint64_t j = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    occluded[ids[j]] = results[j];
}

dim3 grid_for(uint64_t n) { return dim3(uint32_t((n + kBlock - 1) / kBlock)); }

}  // namespace

size_t ray_order_bytes(uint64_t ray_count, bool copies) {
    const uint32_t n32 = uint32_t(ray_count < kRadixPairsMax ? ray_count : kRadixPairsMax);
    size_t b = align256(8 * sizeof(float)) + align256(size_t(kParts) * 6 * sizeof(float)) +
               4 * align256(ray_count * sizeof(uint32_t)) +
               align256(radix_pairs_scratch_bytes(n32));
    if (copies)
        b += align256(ray_count * sizeof(DevRay)) + align256(ray_count * sizeof(DevHit));
    return b;
}

RayOrder ray_order_carve(void* base, uint64_t ray_count, bool copies) {
    const uint32_t n32 = uint32_t(ray_count < kRadixPairsMax ? ray_count : kRadixPairsMax);
    RayOrder r{};
    char* p = static_cast<char*>(base);
    r.frame = reinterpret_cast<float*>(p);
    p += align256(8 * sizeof(float));
    r.parts = reinterpret_cast<float*>(p);
    p += align256(size_t(kParts) * 6 * sizeof(float));
    uint32_t** words[4] = {&r.k0, &r.v0, &r.k1, &r.v1};
    for (uint32_t** w : words) {
        *w = reinterpret_cast<uint32_t*>(p);
        p += align256(ray_count * sizeof(uint32_t));
    }
    r.radix_scratch = p;
    r.radix_bytes = align256(radix_pairs_scratch_bytes(n32));
    p += r.radix_bytes;
    if (copies) {
        r.sorted = reinterpret_cast<DevRay*>(p);
        p += align256(ray_count * sizeof(DevRay));
        r.results = reinterpret_cast<DevHit*>(p);
    }
    r.capacity = ray_count;
    return r;
}

hipError_t ray_order_sort(const RayOrder& r, const DevRay* rays, uint64_t n, hipStream_t stream,
                          const uint32_t** ids) {
    *ids = r.v0;
    if (n == 0) return hipSuccess;
    if (n > r.capacity || n > kRadixPairsMax) return hipErrorInvalidValue;
    const float4* rv = reinterpret_cast<const float4*>(rays);
    const uint64_t samples = uint64_t(kSampleRounds) * kParts * kBlock;
    const uint64_t step = (n + samples - 1) / samples;
    hipLaunchKernelGGL(bounds_part_kernel, dim3(kParts), dim3(kBlock), 0, stream, rv, n, step,
                       r.parts);
    hipLaunchKernelGGL(bounds_final_kernel, dim3(1), dim3(kBlock), 0, stream, r.parts,
                       float(1u << axis_bits(n)), r.frame);
    hipLaunchKernelGGL(keys_kernel, grid_for(n), dim3(kBlock), 0, stream, rv, n, r.frame, r.k0,
                       r.v0);
    bool second = false;
    hipError_t e = radix_sort_pairs(r.k0, r.v0, r.k1, r.v1, uint32_t(n), 3 * axis_bits(n),
                                    r.radix_scratch, r.radix_bytes, stream, &second);
    if (e != hipSuccess) return e;
    *ids = second ? r.v1 : r.v0;
    if (r.sorted)
        hipLaunchKernelGGL(copy_kernel, grid_for(n), dim3(kBlock), 0, stream, rv, *ids, n,
                           reinterpret_cast<float4*>(r.sorted));
    return hipGetLastError();
}

hipError_t ray_order_gather(const RayOrder& r, const uint32_t* ids, uint64_t n, DevHit* hits,
                            uint8_t* occluded, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (n > r.capacity || r.results == nullptr) return hipErrorInvalidValue;
    if (hits)
        hipLaunchKernelGGL(gather_hits_kernel, grid_for(n), dim3(kBlock), 0, stream,
                           reinterpret_cast<const HitVec*>(r.results), ids, n,
                           reinterpret_cast<HitVec*>(hits));
    else
        hipLaunchKernelGGL(gather_occluded_kernel, grid_for(n), dim3(kBlock), 0, stream,
                           reinterpret_cast<const uint8_t*>(r.results), ids, n, occluded);
    return hipGetLastError();
}

}  // namespace rtbvh
