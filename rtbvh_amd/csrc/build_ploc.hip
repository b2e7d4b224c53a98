// GPU locally-ordered clustering (PLOC) builder for gfx950.
//
// Restates LocallyOrderedClusteringBuilder (src/builders/locb.rs) so that the
// node array is the one the reference produces for the same input:
//   world bounds  Aabb::union_of_list (+1e-4)          src/aabb.rs:125-131
//   Morton codes  MortonEncoder::new/encode             src/morton.rs:36-61
//   stable sort   get_sorted_indices (par_sort_by)      src/morton.rs:63-102
//   leaves        [n-1, 2n-1), bounds +1e-4, count 1    locb.rs:285-294
//   iterations    cluster()                             locb.rs:48-245
// The sort is an LSD radix sort (stable, 4 passes of 8 bits), so ties keep
// their input order as rayon's stable merge sort does.
//
// Each clustering iteration over the N active clusters is three kernels
// while N is large:
//   nn_mark   nearest neighbour over |i-j| <= r (ascending j, strict '<', so
//             the lowest j wins ties, locb.rs:92-142) for a tile plus an r
//             halo on each side, then the mutual-pair mark at the lower index
//             (locb.rs:159-167) and a per-tile mark count;
//   tile_scan exclusive scan of the tile counts (the prefix_sum of
//             locb.rs:170-176) and the next iteration's bookkeeping;
//   merge     parents at the upper index's compacted slot, children written
//             to their final positions end-2M.. (locb.rs:178-215).
// Once N <= max(kTailEnter, scenes) the remaining iterations run in one
// workgroup with the clusters resident in LDS (the long tail of small
// iterations costs one launch).
#include <algorithm>
#include <cmath>
#include <vector>

#include "block_scan.hpp"
#include "internal.hpp"

namespace rtbvh {

hipError_t scan_exclusive_u32(uint32_t* data, uint32_t count, uint32_t* total, void* scratch,
                              size_t scratch_bytes, hipStream_t stream);
size_t scan_scratch_bytes(uint32_t count);

namespace {

constexpr int kRadius = 14;            // locb.rs:27
// clusters per tile in the multi-block phase: one per thread, so that the
// many iterations with few clusters still spread over many waves (1024
// clusters per tile built the 1M scene 0.8 ms slower)
constexpr int kTile = 256;
constexpr int kBlock = 256;
constexpr int kTailThreads = 1024;
constexpr int kTailMax = 4096;         // clusters held in LDS by the tail kernel
// The tail takes over at this many clusters (or at one per scene, when a
// batch has more scenes): measured on the 1M scene, 2048 builds 0.3 ms
// faster than 4096 (the single workgroup's share shrinks) and 1024 0.3 ms
// slower (the multi-block iterations' launches cost more than they save).
constexpr uint32_t kTailEnter = 2048;

__host__ __device__ __forceinline__ uint32_t tail_threshold(uint32_t roots) {
    return roots > kTailEnter ? roots : kTailEnter;
}
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;

struct Bounds6 {
    float mn[3], mx[3];
};

// half_area(a.union_of(b)) (src/aabb.rs:74-86, 343-346) as a clustering
// distance.  IEEE min/max give the same values as f32::min/max here: both
// ignore a NaN operand, and where they differ (the sign of a zero) the
// resulting distance compares equal.  Stored bounds use rs_min/rs_max.
__device__ __forceinline__ float dist6(float an0, float an1, float an2, float ax0, float ax1,
                                       float ax2, float bn0, float bn1, float bn2, float bx0,
                                       float bx1, float bx2) {
    const float d0 = fmaxf(ax0, bx0) - fminf(an0, bn0);
    const float d1 = fmaxf(ax1, bx1) - fminf(an1, bn1);
    const float d2 = fmaxf(ax2, bx2) - fminf(an2, bn2);
    return (d0 + d1) * d2 + d0 * d1;
}

__device__ __forceinline__ Bounds6 bounds_of(const DevAabb& a) {
    Bounds6 b;
    b.mn[0] = a.min[0];
    b.mn[1] = a.min[1];
    b.mn[2] = a.min[2];
    b.mx[0] = a.max[0];
    b.mx[1] = a.max[1];
    b.mx[2] = a.max[2];
    return b;
}

__device__ __forceinline__ DevAabb union_node(const DevAabb& s, const DevAabb& b) {
    DevAabb r;
    for (int k = 0; k < 3; ++k) {
        r.min[k] = rs_min(s.min[k], b.min[k]);
        r.max[k] = rs_max(s.max[k], b.max[k]);
    }
    r.extra1 = 0;
    r.extra2 = 0;
    return r;
}

__device__ __forceinline__ void load_node(const DevAabb* p, DevAabb& out) {
    const float4* q = reinterpret_cast<const float4*>(p);
    const float4 a = q[0], b = q[1];
    out.min[0] = a.x;
    out.min[1] = a.y;
    out.min[2] = a.z;
    out.extra1 = __float_as_int(a.w);
    out.max[0] = b.x;
    out.max[1] = b.y;
    out.max[2] = b.z;
    out.extra2 = __float_as_int(b.w);
}

__device__ __forceinline__ void store_node(DevAabb* p, const DevAabb& n) {
    float4* q = reinterpret_cast<float4*>(p);
    q[0] = make_float4(n.min[0], n.min[1], n.min[2], __int_as_float(n.extra1));
    q[1] = make_float4(n.max[0], n.max[1], n.max[2], __int_as_float(n.extra2));
}

// ------------------------------------------------------- input preparation --

// Vector3/Vector4 Primitive::aabb (rtbvh_ffi/src/lib.rs:401-421): an empty
// box grown by the centre.
__global__ void point_aabbs_kernel(const float* centers, uint32_t stride, uint32_t n,
                                   DevAabb* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DevAabb a;
    for (int k = 0; k < 3; ++k) {
        const float c = centers[size_t(i) * stride + k];
        a.min[k] = rs_min(1e34f, c);
        a.max[k] = rs_max(-1e34f, c);
    }
    a.extra1 = 0;
    a.extra2 = 0;
    store_node(out + i, a);
}

// World bounds, first level: per-block min/max with the grow_bb operators.
__global__ void bounds_partial_kernel(const DevAabb* aabbs, uint32_t n, Bounds6* partial) {
    __shared__ Bounds6 sh[kBlock];
    Bounds6 acc;
    for (int k = 0; k < 3; ++k) {
        acc.mn[k] = 1e34f;
        acc.mx[k] = -1e34f;
    }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        DevAabb a;
        load_node(aabbs + i, a);
        for (int k = 0; k < 3; ++k) {
            acc.mn[k] = rs_min(acc.mn[k], a.min[k]);
            acc.mx[k] = rs_max(acc.mx[k], a.max[k]);
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            for (int k = 0; k < 3; ++k) {
                sh[threadIdx.x].mn[k] = rs_min(sh[threadIdx.x].mn[k], sh[threadIdx.x + s].mn[k]);
                sh[threadIdx.x].mx[k] = rs_max(sh[threadIdx.x].mx[k], sh[threadIdx.x + s].mx[k]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// World bounds, second level (one block over the partials, in order), then
// Aabb::with_offset(1e-4).
// One block per scene: block b reduces partial[b*parts, (b+1)*parts).
__global__ void bounds_final_kernel(const Bounds6* partial, uint32_t parts, Bounds6* world) {
    __shared__ Bounds6 sh[kBlock];
    partial += size_t(blockIdx.x) * parts;
    world += blockIdx.x;
    Bounds6 acc;
    for (int k = 0; k < 3; ++k) {
        acc.mn[k] = 1e34f;
        acc.mx[k] = -1e34f;
    }
    for (uint32_t p = threadIdx.x; p < parts; p += kBlock)
        for (int k = 0; k < 3; ++k) {
            acc.mn[k] = rs_min(acc.mn[k], partial[p].mn[k]);
            acc.mx[k] = rs_max(acc.mx[k], partial[p].mx[k]);
        }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int k = 0; k < 3; ++k) {
                sh[threadIdx.x].mn[k] = rs_min(sh[threadIdx.x].mn[k], sh[threadIdx.x + s].mn[k]);
                sh[threadIdx.x].mx[k] = rs_max(sh[threadIdx.x].mx[k], sh[threadIdx.x + s].mx[k]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) {
            acc.mn[k] = sh[0].mn[k] - 0.0001f;
            acc.mx[k] = sh[0].mx[k] + 0.0001f;
        }
        *world = acc;
    }
}

// morton_split for 10-bit inputs: bit b moves to bit 3b (the reference's
// mask loop, src/morton.rs:10-25, is this spreading).
__device__ __forceinline__ uint32_t part1by2(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Rust `f as i32` then max(0), min(1023) (src/morton.rs:57-59).
__device__ __forceinline__ uint32_t grid_coord(float g) {
    int32_t q;
    if (g != g)
        q = 0;
    else if (g >= 2147483648.0f)
        q = 2147483647;
    else if (g <= -2147483904.0f)
        q = -2147483647 - 1;
    else
        q = int32_t(g);
    q = q < 0 ? 0 : q;
    q = q > 1023 ? 1023 : q;
    return uint32_t(q);
}

template <bool SEG>
__global__ void morton_kernel(const float* centers, uint32_t stride, uint32_t n,
                              const Bounds6* world, const uint32_t* prim_scene, uint32_t* keys,
                              uint32_t* vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Bounds6 w = world[SEG ? prim_scene[i] : 0u];
    uint32_t x[3];
    for (int k = 0; k < 3; ++k) {
        // MortonEncoder::new: wtg = 1024 * (1 / diag); off = (-min) * wtg
        const float wtg = 1024.0f * (1.0f / (w.mx[k] - w.mn[k]));
        const float off = (-w.mn[k]) * wtg;
        const float g = centers[size_t(i) * stride + k] * wtg + off;
        x[k] = grid_coord(g);
    }
    keys[i] = part1by2(x[0]) | (part1by2(x[1]) << 1) | (part1by2(x[2]) << 2);
    vals[i] = i;
}

// ------------------------------------------------------------ scene batches --
// A batch is scene_count scenes whose primitives are concatenated; scene s
// owns primitives [off[s], off[s+1]).  Each scene gets its own world bounds
// and Morton grid, the sort orders by (scene, code), and clustering never
// pairs clusters of different scenes, so every scene's tree is the one a
// build of that scene alone produces (node numbering aside).

// prim_scene[i] = s with off[s] <= i < off[s+1].
__global__ void scene_of_kernel(const uint32_t* off, uint32_t scenes, uint32_t n,
                                uint32_t* prim_scene) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = scenes;  // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    prim_scene[i] = lo;
}

// Per-scene world bounds, first level: blockIdx.y is the scene, blockIdx.x
// one of gridDim.x slices of it.
__global__ void bounds_seg_partial_kernel(const DevAabb* aabbs, const uint32_t* off,
                                          Bounds6* partial) {
    __shared__ Bounds6 sh[kBlock];
    Bounds6 acc;
    for (int k = 0; k < 3; ++k) {
        acc.mn[k] = 1e34f;
        acc.mx[k] = -1e34f;
    }
    const uint32_t s = blockIdx.y;
    const uint32_t b = off[s], e = off[s + 1];
    for (uint32_t i = b + blockIdx.x * kBlock + threadIdx.x; i < e; i += gridDim.x * kBlock) {
        DevAabb a;
        load_node(aabbs + i, a);
        for (int k = 0; k < 3; ++k) {
            acc.mn[k] = rs_min(acc.mn[k], a.min[k]);
            acc.mx[k] = rs_max(acc.mx[k], a.max[k]);
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            for (int k = 0; k < 3; ++k) {
                sh[threadIdx.x].mn[k] = rs_min(sh[threadIdx.x].mn[k], sh[threadIdx.x + h].mn[k]);
                sh[threadIdx.x].mx[k] = rs_max(sh[threadIdx.x].mx[k], sh[threadIdx.x + h].mx[k]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[size_t(s) * gridDim.x + blockIdx.x] = sh[0];
}

// Scene keys for the final sort passes: keys[i] = scene of vals[i].
__global__ void scene_keys_kernel(const uint32_t* vals, const uint32_t* prim_scene, uint32_t n,
                                  uint32_t* keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = prim_scene[vals[i]];
}

__global__ void copy_words_kernel(const uint32_t* src, uint32_t* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------ radix sort --
// Stable LSD radix sort of (key, value) pairs, 8-bit digits.
__global__ void radix_hist_kernel(const uint32_t* keys, uint32_t n, uint32_t shift,
                                  uint32_t tiles, uint32_t* counts) {
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// A tile's pairs are ranked into LDS in digit order (stable: rounds, then
// waves, then lanes in order), then written out so that consecutive threads
// write consecutive slots of one digit's run.  The tile's count of digit d is
// the difference of consecutive entries of the scanned counts (digit-major,
// the last entry's end is n).
__global__ void __launch_bounds__(kSortBlock)
    radix_scatter_kernel(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t n,
                         uint32_t shift, uint32_t tiles, const uint32_t* offsets,
                         uint32_t* keys_out, uint32_t* vals_out) {
    constexpr int kWaves = kSortBlock / 64;
    static_assert(kSortBlock == 256, "one thread per digit");
    __shared__ uint32_t run[256];    // next local slot of each digit
    __shared__ uint32_t start[256];  // first local slot of each digit
    __shared__ uint32_t gbase[256];  // first global slot of each digit's run of this tile
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t scan_sh[kWaves];
    __shared__ uint32_t skeys[kSortTile];
    __shared__ uint32_t svals[kSortTile];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t base = blockIdx.x * kSortTile;
    {
        const uint32_t idx = threadIdx.x * tiles + blockIdx.x;
        const uint32_t first = offsets[idx];
        const uint32_t next = idx + 1 < 256u * tiles ? offsets[idx + 1] : n;
        gbase[threadIdx.x] = first;
        const uint32_t s = block_exclusive_n<kSortBlock>(next - first, scan_sh, nullptr);
        start[threadIdx.x] = s;
        run[threadIdx.x] = s;
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kSortItems; ++r) {
        for (int w = 0; w < kWaves; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const uint32_t i = base + r * kSortBlock + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? vals_in[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        const uint32_t cnt = __popcll(peers);
        if (valid && rank == 0) wcnt[wave][d] = cnt;
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            pos = run[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wcnt[w][d];
        }
        __syncthreads();
        {
            uint32_t add = 0;
            for (int w = 0; w < kWaves; ++w) add += wcnt[w][threadIdx.x];
            run[threadIdx.x] += add;
        }
        if (valid) {
            skeys[pos] = key;
            svals[pos] = val;
        }
        __syncthreads();
    }
    const uint32_t count = n - base < uint32_t(kSortTile) ? n - base : uint32_t(kSortTile);
    for (uint32_t k = threadIdx.x; k < count; k += kSortBlock) {
        const uint32_t key = skeys[k];
        const uint32_t d = (key >> shift) & 255u;
        const uint32_t g = gbase[d] + (k - start[d]);
        keys_out[g] = key;
        vals_out[g] = svals[k];
    }
}

// ---------------------------------------------------------------- leaves --
// locb.rs:289-294: leaf i = aabbs[prim_indices[i]].with_offset(1e-4), count 1,
// left_first i.
template <bool SEG>
__global__ void leaves_kernel(const DevAabb* aabbs, const uint32_t* sorted, uint32_t n,
                              const uint32_t* prim_scene, DevAabb* clusters,
                              uint32_t* cluster_scene) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t prim = sorted[i];
    DevAabb a;
    load_node(aabbs + prim, a);
    for (int k = 0; k < 3; ++k) {
        a.min[k] = a.min[k] - 0.0001f;
        a.max[k] = a.max[k] + 0.0001f;
    }
    a.extra1 = 1;
    a.extra2 = int32_t(i);
    store_node(clusters + i, a);
    if (SEG) cluster_scene[i] = prim_scene[prim];
}

// n <= 2: a single root leaf with the world bounds (locb.rs:258-269).
__global__ void small_root_kernel(const Bounds6* world, uint32_t n, DevAabb* nodes,
                                  uint32_t* indices) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DevAabb r;
    for (int k = 0; k < 3; ++k) {
        r.min[k] = world->mn[k];
        r.max[k] = world->mx[k];
    }
    r.extra1 = int32_t(n);
    r.extra2 = 0;
    store_node(nodes, r);
    for (uint32_t i = 0; i < n; ++i) indices[i] = i;
}

// ------------------------------------------------------ clustering state --
// One state record per iteration: iteration k reads states[k] and writes
// states[k+1], so batches of iterations are enqueued without host round
// trips.  An iteration whose input already fits the tail kernel passes its
// state through unchanged.
struct PlocState {
    uint32_t n_active;  // clusters in [begin, end) of the reference's node array
    uint32_t begin;
    uint32_t end;
    uint32_t merged;    // pairs merged by this iteration (M)
    uint32_t buf;       // cluster buffer holding the active clusters
    uint32_t tail_at;   // the tail kernel takes over at this many clusters
    uint32_t pad[2];
};

__device__ __forceinline__ void nn_window(uint32_t i, uint32_t n, uint32_t& sb, uint32_t& se) {
    // search_range, locb.rs:36-45 (relative to begin = 0)
    sb = i > uint32_t(kRadius) ? i - kRadius : 0u;
    se = (i + kRadius + 1) < n ? i + kRadius + 1 : n;
}

// Nearest neighbour + mutual marks for one tile.  Bounds are staged in LDS
// as six arrays so that lanes reading consecutive clusters hit consecutive
// banks.  kNnBlock threads cover the tile and its halo (kTile + 2 kRadius
// neighbour searches) in one pass, and each search loads its candidates four
// at a time and compares them in index order (the first lowest distance
// wins, as in the sequential loop): later iterations are short, and their
// time is the latency of this chain.
constexpr int kNnBlock = 320;
static_assert(kNnBlock >= kTile + 2 * kRadius, "one neighbour search per thread");
static_assert(kNnBlock >= kTile + 4 * kRadius, "staging in one pass");

template <bool SEG>
__global__ void __launch_bounds__(kNnBlock)
    nn_mark_kernel(const PlocState* state, const DevAabb* cl0, const DevAabb* cl1,
                   const uint32_t* cs0, const uint32_t* cs1, uint32_t* nb_out,
                   uint32_t* mark_out, uint32_t* tile_count) {
    constexpr int kHalo = 2 * kRadius;
    constexpr int kSpan = kTile + 2 * kHalo;
    __shared__ float s_b[6][kSpan];
    __shared__ uint32_t s_sc[SEG ? kSpan : 1];
    __shared__ uint32_t s_nb[kTile + 2 * kRadius];
    __shared__ uint32_t s_total;
    const PlocState stt = *state;
    if (stt.n_active <= stt.tail_at) return;
    const uint32_t n = stt.n_active;
    const DevAabb* cl = stt.buf ? cl1 : cl0;
    const uint32_t* cs = stt.buf ? cs1 : cs0;
    const uint32_t ts = blockIdx.x * kTile;
    if (ts >= n) return;
    const uint32_t te = ts + kTile < n ? ts + kTile : n;
    const uint32_t lb = ts > uint32_t(kHalo) ? ts - kHalo : 0u;
    const uint32_t le = te + kHalo < n ? te + kHalo : n;
    if (threadIdx.x == 0) s_total = 0;
    for (uint32_t k = lb + threadIdx.x; k < le; k += kNnBlock) {
        DevAabb a;
        load_node(cl + k, a);
        s_b[0][k - lb] = a.min[0];
        s_b[1][k - lb] = a.min[1];
        s_b[2][k - lb] = a.min[2];
        s_b[3][k - lb] = a.max[0];
        s_b[4][k - lb] = a.max[1];
        s_b[5][k - lb] = a.max[2];
        if (SEG) s_sc[k - lb] = cs[k];
    }
    __syncthreads();
    const uint32_t nb0 = ts > uint32_t(kRadius) ? ts - kRadius : 0u;
    const uint32_t nb1 = te + kRadius < n ? te + kRadius : n;
    for (uint32_t i = nb0 + threadIdx.x; i < nb1; i += kNnBlock) {
        uint32_t s0, s1;
        nn_window(i, n, s0, s1);
        const uint32_t li = i - lb;
        const float a0 = s_b[0][li], a1 = s_b[1][li], a2 = s_b[2][li];
        const float a3 = s_b[3][li], a4 = s_b[4][li], a5 = s_b[5][li];
        float best = 3.40282347e38f;  // f32::MAX (locb.rs:95)
        uint32_t best_j = 0xFFFFFFFFu;
        for (uint32_t j0 = s0; j0 < s1; j0 += 4) {
            float d[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t j = j0 + uint32_t(u);
                // a position past the window reads a staged cluster and is
                // dropped (NaN never compares less)
                const uint32_t lj = (j < s1 ? j : s1 - 1) - lb;
                const bool skip = j >= s1 || j == i || (SEG && s_sc[lj] != s_sc[li]);
                const float dd = dist6(a0, a1, a2, a3, a4, a5, s_b[0][lj], s_b[1][lj],
                                       s_b[2][lj], s_b[3][lj], s_b[4][lj], s_b[5][lj]);
                d[u] = skip ? __builtin_nanf("") : dd;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (d[u] < best) {
                    best = d[u];
                    best_j = j0 + uint32_t(u);
                }
        }
        s_nb[i - nb0] = best_j;
    }
    __syncthreads();
    uint32_t cnt = 0;
    for (uint32_t i = ts + threadIdx.x; i < te; i += kNnBlock) {
        const uint32_t j = s_nb[i - nb0];
        uint32_t m = 0;
        if (j > i && j < nb1) m = (s_nb[j - nb0] == i) ? 1u : 0u;
        nb_out[i] = j;
        mark_out[i] = m;
        cnt += m;
    }
    if (cnt) atomicAdd(&s_total, cnt);
    __syncthreads();
    if (threadIdx.x == 0) tile_count[blockIdx.x] = s_total;
}

// Exclusive scan of the tile counts; writes M into states[k] and the next
// iteration's bookkeeping into states[k+1] (locb.rs:178-185).
__global__ void __launch_bounds__(1024) tile_scan_kernel(PlocState* states, uint32_t* tile_count) {
    __shared__ uint32_t sums[1024 / 64];
    const PlocState stt = states[0];
    if (stt.n_active <= stt.tail_at) {
        if (threadIdx.x == 0) {
            PlocState nx = stt;
            nx.merged = 0;
            states[1] = nx;
        }
        return;
    }
    const uint32_t n = stt.n_active;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    const uint32_t per = (tiles + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = b + per < tiles ? b + per : tiles;
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += tile_count[i];
    uint32_t m = 0;
    uint32_t run = block_exclusive_n<1024>(s, sums, &m);
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = tile_count[i];
        tile_count[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) {
        states[0].merged = m;
        PlocState nx;
        nx.n_active = n - m;
        nx.begin = stt.begin - m;
        nx.end = stt.end - 2 * m;
        nx.merged = 0;
        nx.buf = stt.buf ^ 1u;
        nx.tail_at = stt.tail_at;
        nx.pad[0] = nx.pad[1] = 0;
        states[1] = nx;
    }
}

// Parents, children and compaction for one tile (locb.rs:193-215).  m[k] is
// the inclusive prefix of marks; the tile keeps m for [ts - r, te).
template <bool SEG>
__global__ void __launch_bounds__(kBlock)
    merge_kernel(const PlocState* state, const DevAabb* cl0, DevAabb* cl1, const uint32_t* cs0,
                 uint32_t* cs1, const uint32_t* nb, const uint32_t* mark,
                 const uint32_t* tile_offset, DevAabb* nodes) {
    __shared__ uint32_t s_m[kRadius + kTile];
    __shared__ uint32_t s_scan[kBlock / 64];
    const PlocState stt = *state;
    if (stt.n_active <= stt.tail_at) return;
    const uint32_t n = stt.n_active;
    const DevAabb* cl = stt.buf ? cl1 : cl0;
    DevAabb* cl_out = stt.buf ? const_cast<DevAabb*>(cl0) : cl1;
    const uint32_t* cs = stt.buf ? cs1 : cs0;
    uint32_t* cs_out = stt.buf ? const_cast<uint32_t*>(cs0) : cs1;
    const uint32_t ts = blockIdx.x * kTile;
    if (ts >= n) return;
    const uint32_t te = ts + kTile < n ? ts + kTile : n;
    const uint32_t children_begin = stt.end - 2 * stt.merged;
    static_assert(kTile == kBlock, "one cluster of the tile per thread");
    // every load that does not wait for another first: the cluster, its
    // neighbour and mark, the marks of the left halo (one per lane of wave
    // 0) and the tile's offset; then the neighbour's neighbour and box
    const uint32_t i = ts + threadIdx.x;
    const bool live = i < te;
    const uint32_t mk = live ? mark[i] : 0u;
    const uint32_t j = live ? nb[i] : 0xFFFFFFFFu;
    DevAabb ci;
    if (live) load_node(cl + i, ci);
    const uint32_t hm =
        threadIdx.x < uint32_t(kRadius) && ts > threadIdx.x ? mark[ts - 1 - threadIdx.x] : 0u;
    const uint32_t excl = tile_offset[blockIdx.x];  // marks in [0, ts)
    const uint32_t nbj = live && j < n ? nb[j] : 0xFFFFFFFFu;
    const bool mutual = nbj == i;
    const bool upper = mutual && j < i;  // i is the upper element of its pair
    DevAabb cj;
    if (upper) load_node(cl + j, cj);
    // m = inclusive prefix of marks over [ts - r, te)
    const uint32_t run = excl + block_exclusive_n<kBlock>(mk, s_scan, nullptr) + mk;
    s_m[kRadius + threadIdx.x] = run;
    if (threadIdx.x < 64) {
        // left halo: m[ts - 1 - l] = m[ts - 1] - (mark[ts - 1] + .. + mark[ts - l])
        const uint32_t hincl = wave_inclusive(hm);
        if (threadIdx.x < uint32_t(kRadius) && ts > threadIdx.x)
            s_m[kRadius - 1 - threadIdx.x] = excl - (hincl - hm);
    }
    __syncthreads();
    if (!live) return;
    const uint32_t mi = run;
    if (mutual) {
        if (upper) {
            // the parent takes the upper element's compacted slot
            const uint32_t mj = s_m[kRadius + j - ts];  // j >= ts - r
            const uint32_t first_child = children_begin + (mj - 1) * 2;
            DevAabb parent = union_node(ci, cj);  // input[j].union_of(input[i]), locb.rs:202
            parent.extra1 = -1;
            parent.extra2 = int32_t(first_child);
            store_node(cl_out + (i - mi), parent);
            if (SEG) cs_out[i - mi] = cs[i];
            store_node(nodes + first_child, cj);
            store_node(nodes + first_child + 1, ci);
        }
    } else {
        store_node(cl_out + (i - mi), ci);
        if (SEG) cs_out[i - mi] = cs[i];
    }
}

// All remaining iterations in one workgroup, clusters resident in LDS as
// six bound arrays plus the two node fields.  Thread t owns clusters
// t, t + 1024, t + 2048, t + 3072, so lanes touch consecutive LDS words; the
// mark prefix comes from wave ballots.
constexpr int kTailWaves = kTailThreads / 64;

// CAP clusters at most: kTailEnter for the common case, kTailMax for a
// batch of more scenes (the smaller instance keeps half the registers per
// thread and builds the 1M scene 0.3 ms faster).
template <bool SEG, int CAP>
struct TailShared {
    static constexpr int kPer = CAP / kTailThreads;  // clusters per thread
    float b[6][CAP];
    int32_t e1[CAP];
    int32_t e2[CAP];
    uint16_t nb[CAP];
    uint16_t m[CAP];
    uint16_t sc[SEG ? CAP : 2];  // scene of each cluster (batches)
    uint32_t wsum[kPer * kTailWaves];
};

template <typename TS>
__device__ __forceinline__ DevAabb tail_node(const TS& sh, uint32_t i) {
    DevAabb a;
    a.min[0] = sh.b[0][i];
    a.min[1] = sh.b[1][i];
    a.min[2] = sh.b[2][i];
    a.max[0] = sh.b[3][i];
    a.max[1] = sh.b[4][i];
    a.max[2] = sh.b[5][i];
    a.extra1 = sh.e1[i];
    a.extra2 = sh.e2[i];
    return a;
}

template <typename TS>
__device__ __forceinline__ void tail_store(TS& sh, uint32_t i, const DevAabb& a) {
    sh.b[0][i] = a.min[0];
    sh.b[1][i] = a.min[1];
    sh.b[2][i] = a.min[2];
    sh.b[3][i] = a.max[0];
    sh.b[4][i] = a.max[1];
    sh.b[5][i] = a.max[2];
    sh.e1[i] = a.extra1;
    sh.e2[i] = a.extra2;
}

// Runs until n_final clusters remain (1, or one per scene of a batch), stores
// them at nodes[0, n_final) and reports the count it reached in *result.
template <bool SEG, int CAP>
__global__ void __launch_bounds__(kTailThreads)
    tail_kernel(const PlocState* state, const DevAabb* cl0, const DevAabb* cl1,
                const uint32_t* cs0, const uint32_t* cs1, DevAabb* nodes, uint32_t n_final,
                PlocState* result) {
    extern __shared__ unsigned char smem_raw[];
    using TS = TailShared<SEG, CAP>;
    constexpr int kTailPer = TS::kPer;
    TS& sh = *reinterpret_cast<TS*>(smem_raw);
    const PlocState stt = *state;
    uint32_t n = stt.n_active;
    uint32_t end = stt.end;
    if (n > uint32_t(CAP)) {  // the iterations before did not get here
        if (threadIdx.x == 0) *result = stt;
        return;
    }
    const DevAabb* cl_in = stt.buf ? cl1 : cl0;
    const uint32_t* cs_in = stt.buf ? cs1 : cs0;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    for (uint32_t i = tid; i < n; i += kTailThreads) {
        DevAabb a;
        load_node(cl_in + i, a);
        tail_store(sh, i, a);
        if (SEG) sh.sc[i] = uint16_t(cs_in[i]);
    }
    __syncthreads();
    while (n > n_final) {
        // nearest neighbours
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const uint32_t i = tid + uint32_t(k) * kTailThreads;
            if (i >= n) continue;
            uint32_t s0, s1;
            nn_window(i, n, s0, s1);
            const float a0 = sh.b[0][i], a1 = sh.b[1][i], a2 = sh.b[2][i];
            const float a3 = sh.b[3][i], a4 = sh.b[4][i], a5 = sh.b[5][i];
            const uint32_t si = SEG ? sh.sc[i] : 0u;
            float best = 3.40282347e38f;
            uint32_t best_j = 0xFFFFu;
            // four candidates in flight, compared in index order (as
            // nn_mark_kernel), where registers allow: the batch instances
            // (scene ids, or 4096 clusters) would spill, and keep the
            // one-at-a-time loop
            constexpr uint32_t kStep = (!SEG && CAP <= int(kTailEnter)) ? 4u : 1u;
            for (uint32_t j0 = s0; j0 < s1; j0 += kStep) {
                float d[kStep];
#pragma unroll
                for (int u = 0; u < int(kStep); ++u) {
                    const uint32_t j = j0 + uint32_t(u);
                    const uint32_t jc = j < s1 ? j : s1 - 1;
                    const bool skip = j >= s1 || j == i || (SEG && sh.sc[jc] != si);
                    const float dd = dist6(a0, a1, a2, a3, a4, a5, sh.b[0][jc], sh.b[1][jc],
                                           sh.b[2][jc], sh.b[3][jc], sh.b[4][jc], sh.b[5][jc]);
                    d[u] = skip ? __builtin_nanf("") : dd;
                }
#pragma unroll
                for (int u = 0; u < int(kStep); ++u)
                    if (d[u] < best) {
                        best = d[u];
                        best_j = j0 + uint32_t(u);
                    }
            }
            sh.nb[i] = uint16_t(best_j);
        }
        __syncthreads();
        // mutual marks at the lower index and their inclusive prefix
        uint32_t mk[kTailPer], pre[kTailPer];
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const uint32_t i = tid + uint32_t(k) * kTailThreads;
            bool mark = false;
            if (i < n) {
                const uint32_t j = sh.nb[i];
                mark = j > i && j < n && sh.nb[j] == i;
            }
            const uint64_t bal = __ballot(mark);
            mk[k] = mark ? 1u : 0u;
            pre[k] = __popcll(bal & lt);
            if (lane == 0) sh.wsum[k * kTailWaves + wave] = __popcll(bal);
        }
        __syncthreads();
        uint32_t off[kTailPer];
        uint32_t running = 0;
        for (int c = 0; c < kTailPer * kTailWaves; ++c) {
#pragma unroll
            for (int k = 0; k < kTailPer; ++k)
                if (c == k * kTailWaves + int(wave)) off[k] = running;
            running += sh.wsum[c];
        }
        const uint32_t merged = running;
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const uint32_t i = tid + uint32_t(k) * kTailThreads;
            if (i < n) sh.m[i] = uint16_t(off[k] + pre[k] + mk[k]);
        }
        __syncthreads();
        if (merged == 0) break;  // no mutual pair: only with NaN bounds
        const uint32_t children_begin = end - 2 * merged;
        DevAabb keep[kTailPer];
        uint32_t dst[kTailPer];
        uint16_t keep_sc[kTailPer];
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const uint32_t i = tid + uint32_t(k) * kTailThreads;
            dst[k] = 0xFFFFFFFFu;
            if (i >= n) continue;
            keep_sc[k] = SEG ? sh.sc[i] : uint16_t(0);
            const uint32_t j = sh.nb[i];
            const bool mutual = j < n && sh.nb[j] == i;
            const uint32_t mi = sh.m[i];
            if (mutual) {
                if (j < i) {
                    const uint32_t first_child = children_begin + (uint32_t(sh.m[j]) - 1) * 2;
                    const DevAabb ci = tail_node(sh, i);
                    const DevAabb cj = tail_node(sh, j);
                    DevAabb parent = union_node(ci, cj);
                    parent.extra1 = -1;
                    parent.extra2 = int32_t(first_child);
                    store_node(nodes + first_child, cj);
                    store_node(nodes + first_child + 1, ci);
                    keep[k] = parent;
                    dst[k] = i - mi;
                }
            } else {
                keep[k] = tail_node(sh, i);
                dst[k] = i - mi;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTailPer; ++k)
            if (dst[k] != 0xFFFFFFFFu) {
                tail_store(sh, dst[k], keep[k]);
                if (SEG) sh.sc[dst[k]] = keep_sc[k];
            }
        __syncthreads();
        n -= merged;
        end -= 2 * merged;
    }
    if (n == n_final)
        for (uint32_t i = tid; i < n; i += kTailThreads) store_node(nodes + i, tail_node(sh, i));
    if (tid == 0) {
        PlocState r = stt;
        r.n_active = n;
        r.end = end;
        *result = r;
    }
}

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

constexpr uint32_t kMaxIterations = 1u << 14;
constexpr uint32_t kBatch = 6;  // iterations before the first host check

// The first iteration's state (a kernel, so that the whole build can be
// recorded into a graph).
__global__ void init_state_kernel(PlocState* st, uint32_t n, uint32_t roots) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    PlocState s;
    s.n_active = n;
    s.begin = n - roots;
    s.end = 2 * n - roots;
    s.merged = 0;
    s.buf = 0;
    s.tail_at = tail_threshold(roots);
    s.pad[0] = s.pad[1] = 0;
    st[0] = s;
}

// fail = 1 when clustering stopped before one cluster per scene remained.
__global__ void check_kernel(const PlocState* result, uint32_t n_final, uint32_t* fail) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    fail[0] = result->n_active == n_final ? 0u : 1u;
}

}  // namespace

// Stable LSD radix sort of n (key, value) pairs by the low `bits` bits of
// the keys, 8 bits a pass, with the kernels of the Morton-code sort.  The
// pairs ping-pong between (k0, v0) and (k1, v1); *second is set when the
// sorted pairs end in (k1, v1).  n <= kRadixPairsMax (the kernels index in
// 32 bits, a tile past n included).
size_t radix_pairs_scratch_bytes(uint32_t n) {
    const size_t tiles = (size_t(n) + kSortTile - 1) / kSortTile;
    return align_up(256 * tiles * 4) + align_up(scan_scratch_bytes(uint32_t(256 * tiles)));
}

hipError_t radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t n,
                            uint32_t bits, void* scratch, size_t scratch_bytes,
                            hipStream_t stream, bool* second) {
    *second = false;
    if (n == 0 || bits == 0) return hipSuccess;
    if (n > kRadixPairsMax || scratch_bytes < radix_pairs_scratch_bytes(n))
        return hipErrorInvalidValue;
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    const size_t counts_bytes = align_up(size_t(256) * tiles * 4);
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    void* scan = static_cast<char*>(scratch) + counts_bytes;
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(radix_hist_kernel, dim3(tiles), dim3(256), 0, stream, ki, n, shift,
                           tiles, counts);
        const hipError_t e = scan_exclusive_u32(counts, 256u * tiles, nullptr, scan,
                                                scratch_bytes - counts_bytes, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(tiles), dim3(kSortBlock), 0, stream, ki,
                           vi, n, shift, tiles, counts, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
        *second = !*second;
    }
    return hipGetLastError();
}

// Every buffer of one PLOC build of n primitives in `scenes` scenes.  The
// tree (nodes, indices) has allocations of its own so that an eager build
// can hand it over and free the rest.
struct PlocPlanImpl {
    uint32_t n = 0, scenes = 1, stride = 3;
    bool seg = false, point_aabbs = false;
    uint32_t iterations = 0;  // clustering iterations recorded by enqueue
    DevAabb* nodes = nullptr;
    uint32_t* indices = nullptr;
    void* ws = nullptr;
    Workspace wsh;  // ws from the workspace cache (eager builds)
    // carved from ws
    DevAabb* in_aabbs = nullptr;  // point boxes (aabbs == null)
    Bounds6* part = nullptr;
    Bounds6* world = nullptr;
    uint32_t *k0 = nullptr, *k1 = nullptr, *v1 = nullptr, *counts = nullptr;
    void* scan_scratch = nullptr;
    size_t scan_bytes = 0;
    DevAabb *cl0 = nullptr, *cl1 = nullptr;
    uint32_t *nbuf = nullptr, *mbuf = nullptr, *tcount = nullptr;
    PlocState* st = nullptr;
    uint32_t *soff = nullptr, *pscene = nullptr, *cs0 = nullptr, *cs1 = nullptr;
    uint32_t* fail = nullptr;
    uint32_t parts = 256, sort_tiles = 0, ploc_tiles = 0, count_entries = 0;

    uint32_t node_count() const { return seg ? 2 * n - scenes : (n <= 2 ? 1u : 2 * n - 1); }
};

namespace {

void free_plan(PlocPlanImpl* p, bool keep_tree) {
    if (!p) return;
    if (!keep_tree) {
        if (p->nodes) (void)hipFree(p->nodes);
        if (p->indices) (void)hipFree(p->indices);
    }
    if (p->wsh.ptr)
        // a failed build may have work in flight; a finished one has waited
        workspace_release(&p->wsh, nullptr, !keep_tree);
    else if (p->ws)
        (void)hipFree(p->ws);
    delete p;
}

void init_plan(PlocPlanImpl* p, uint32_t n, uint32_t scenes, bool point_aabbs,
               size_t stride_floats) {
    p->n = n;
    p->scenes = scenes;
    p->seg = scenes > 1;
    p->point_aabbs = point_aabbs;
    p->stride = uint32_t(stride_floats);
    p->sort_tiles = (n + kSortTile - 1) / kSortTile;
    p->ploc_tiles = (n + kTile - 1) / kTile;
    p->parts = p->seg ? 16u : 256u;  // bounds partials (per scene when batched)
    p->count_entries = 256u * p->sort_tiles;
    p->scan_bytes = scan_scratch_bytes(p->count_entries);
}

// Carves the build's workspace out of `base` (when not null) and returns its
// size in bytes.
size_t carve_plan(PlocPlanImpl* p, char* base) {
    const uint32_t n = p->n, scenes = p->scenes;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += align_up(bytes);
        return o;
    };
    const size_t o_aabb = p->point_aabbs ? carve(size_t(n) * sizeof(DevAabb)) : 0;
    const size_t o_part = carve(size_t(p->parts) * scenes * sizeof(Bounds6));
    const size_t o_world = carve(size_t(scenes) * sizeof(Bounds6));
    const size_t o_k0 = carve(size_t(n) * 4), o_k1 = carve(size_t(n) * 4);
    const size_t o_v1 = carve(size_t(n) * 4);
    const size_t o_cnt = carve(size_t(p->count_entries) * 4);
    const size_t o_scan = carve(p->scan_bytes);
    const size_t o_cl0 = carve(size_t(n) * sizeof(DevAabb));
    const size_t o_cl1 = carve(size_t(n) * sizeof(DevAabb));
    const size_t o_nb = carve(size_t(n) * 4), o_mark = carve(size_t(n) * 4);
    const size_t o_tc = carve(size_t(p->ploc_tiles) * 4);
    const size_t o_state = carve(size_t(kMaxIterations + 2) * sizeof(PlocState));
    const size_t o_soff = carve(size_t(scenes + 1) * 4);
    const size_t o_pscene = p->seg ? carve(size_t(n) * 4) : 0;
    const size_t o_cs0 = p->seg ? carve(size_t(n) * 4) : 0;
    const size_t o_cs1 = p->seg ? carve(size_t(n) * 4) : 0;
    const size_t o_fail = carve(4);
    if (base == nullptr) return off;
    char* b = base;
    p->in_aabbs = p->point_aabbs ? reinterpret_cast<DevAabb*>(b + o_aabb) : nullptr;
    p->part = reinterpret_cast<Bounds6*>(b + o_part);
    p->world = reinterpret_cast<Bounds6*>(b + o_world);
    p->k0 = reinterpret_cast<uint32_t*>(b + o_k0);
    p->k1 = reinterpret_cast<uint32_t*>(b + o_k1);
    p->v1 = reinterpret_cast<uint32_t*>(b + o_v1);
    p->counts = reinterpret_cast<uint32_t*>(b + o_cnt);
    p->scan_scratch = b + o_scan;
    p->cl0 = reinterpret_cast<DevAabb*>(b + o_cl0);
    p->cl1 = reinterpret_cast<DevAabb*>(b + o_cl1);
    p->nbuf = reinterpret_cast<uint32_t*>(b + o_nb);
    p->mbuf = reinterpret_cast<uint32_t*>(b + o_mark);
    p->tcount = reinterpret_cast<uint32_t*>(b + o_tc);
    p->st = reinterpret_cast<PlocState*>(b + o_state);
    p->soff = reinterpret_cast<uint32_t*>(b + o_soff);
    p->pscene = p->seg ? reinterpret_cast<uint32_t*>(b + o_pscene) : nullptr;
    p->cs0 = p->seg ? reinterpret_cast<uint32_t*>(b + o_cs0) : nullptr;
    p->cs1 = p->seg ? reinterpret_cast<uint32_t*>(b + o_cs1) : nullptr;
    p->fail = reinterpret_cast<uint32_t*>(b + o_fail);
    return off;
}

// cached: the workspace comes from the per-device cache (eager builds, which
// release it when they return; plans keep theirs for their lifetime).
ResultCode alloc_plan(uint32_t n, uint32_t scenes, bool point_aabbs, size_t stride_floats,
                      PlocPlanImpl** out, bool cached) {
    PlocPlanImpl* p = new PlocPlanImpl();
    init_plan(p, n, scenes, point_aabbs, stride_floats);
    const size_t bytes = carve_plan(p, nullptr);
    hipError_t e;
    if ((e = hipMalloc(&p->nodes, size_t(p->node_count()) * sizeof(DevAabb))) != hipSuccess ||
        (e = hipMalloc(&p->indices, size_t(n) * sizeof(uint32_t))) != hipSuccess ||
        (e = cached ? workspace_acquire(kWsPloc, bytes, &p->wsh) : hipMalloc(&p->ws, bytes)) !=
            hipSuccess) {
        free_plan(p, false);
        return hip_fail(e, "hipMalloc(ploc)");
    }
    if (cached) p->ws = p->wsh.ptr;
    carve_plan(p, static_cast<char*>(p->ws));
    *out = p;
    return Ok;
}

// Point boxes, world bounds, Morton codes, sort, leaves and the first state.
// Needs p->soff on the device for batches.
template <bool SEG>
hipError_t enqueue_front(const PlocPlanImpl* p, const RTAabb* aabbs, const float* centers,
                         hipStream_t stream) {
    const uint32_t n = p->n, g = (n + kBlock - 1) / kBlock;
    const DevAabb* in_aabbs = aabbs ? reinterpret_cast<const DevAabb*>(aabbs) : p->in_aabbs;
    if (!aabbs)
        hipLaunchKernelGGL(point_aabbs_kernel, dim3(g), dim3(kBlock), 0, stream, centers,
                           p->stride, n, p->in_aabbs);
    if (SEG) {
        hipLaunchKernelGGL(scene_of_kernel, dim3(g), dim3(kBlock), 0, stream, p->soff, p->scenes,
                           n, p->pscene);
        hipLaunchKernelGGL(bounds_seg_partial_kernel, dim3(p->parts, p->scenes), dim3(kBlock), 0,
                           stream, in_aabbs, p->soff, p->part);
        hipLaunchKernelGGL(bounds_final_kernel, dim3(p->scenes), dim3(kBlock), 0, stream, p->part,
                           p->parts, p->world);
    } else {
        hipLaunchKernelGGL(bounds_partial_kernel, dim3(p->parts), dim3(kBlock), 0, stream,
                           in_aabbs, n, p->part);
        hipLaunchKernelGGL(bounds_final_kernel, dim3(1), dim3(kBlock), 0, stream, p->part,
                           p->parts, p->world);
    }
    hipLaunchKernelGGL(morton_kernel<SEG>, dim3(g), dim3(kBlock), 0, stream, centers, p->stride,
                       n, p->world, p->pscene, p->k0, p->indices);
    // 4 stable passes over the codes, then (batches) 1 or 2 passes over the
    // scene ids; the sorted order must end in p->indices.
    uint32_t *ki = p->k0, *vi = p->indices, *ko = p->k1, *vo = p->v1;
    const uint32_t scene_bits = SEG ? 32u - uint32_t(__builtin_clz(p->scenes - 1)) : 0u;
    const uint32_t scene_passes = (scene_bits + 7) / 8;
    hipError_t e;
    for (uint32_t pass = 0; pass < 4 + scene_passes; ++pass) {
        uint32_t shift = pass * 8;
        if (pass == 4)
            hipLaunchKernelGGL(scene_keys_kernel, dim3(g), dim3(kBlock), 0, stream, vi, p->pscene,
                               n, ki);
        if (pass >= 4) shift = (pass - 4) * 8;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(p->sort_tiles), dim3(256), 0, stream, ki, n,
                           shift, p->sort_tiles, p->counts);
        if ((e = scan_exclusive_u32(p->counts, p->count_entries, nullptr, p->scan_scratch,
                                    p->scan_bytes, stream)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(p->sort_tiles), dim3(kSortBlock), 0, stream,
                           ki, vi, n, shift, p->sort_tiles, p->counts, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (vi != p->indices)
        hipLaunchKernelGGL(copy_words_kernel, dim3(g), dim3(kBlock), 0, stream, vi, p->indices, n);
    hipLaunchKernelGGL(leaves_kernel<SEG>, dim3(g), dim3(kBlock), 0, stream, in_aabbs, p->indices,
                       n, p->pscene, p->cl0, p->cs0);
    hipLaunchKernelGGL(init_state_kernel, dim3(1), dim3(1), 0, stream, p->st, n,
                       SEG ? p->scenes : 1u);
    return hipGetLastError();
}

// One multi-block clustering iteration reading states[iter]; a no-op once
// the active set fits the tail kernel.
template <bool SEG>
void enqueue_iteration(const PlocPlanImpl* p, uint32_t iter, uint32_t tiles, hipStream_t stream) {
    hipLaunchKernelGGL(nn_mark_kernel<SEG>, dim3(tiles), dim3(kNnBlock), 0, stream, p->st + iter,
                       p->cl0, p->cl1, p->cs0, p->cs1, p->nbuf, p->mbuf, p->tcount);
    hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, stream, p->st + iter, p->tcount);
    hipLaunchKernelGGL(merge_kernel<SEG>, dim3(tiles), dim3(kBlock), 0, stream, p->st + iter,
                       p->cl0, p->cl1, p->cs0, p->cs1, p->nbuf, p->mbuf, p->tcount, p->nodes);
}

template <bool SEG>
hipError_t set_tail_lds() {
    static bool done[64] = {false};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (done[dev]) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(tail_kernel<SEG, int(kTailEnter)>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(TailShared<SEG, int(kTailEnter)>)));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(tail_kernel<SEG, kTailMax>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                int(sizeof(TailShared<SEG, kTailMax>)));
    if (e == hipSuccess) done[dev] = true;
    return e;
}

// The remaining iterations in one workgroup from states[iter]; the final
// state lands in states[iter + 1] and p->fail tells whether it is complete.
template <bool SEG>
hipError_t enqueue_tail(const PlocPlanImpl* p, uint32_t iter, hipStream_t stream) {
    const uint32_t roots = SEG ? p->scenes : 1u;
    if (tail_threshold(roots) <= kTailEnter)
        hipLaunchKernelGGL((tail_kernel<SEG, int(kTailEnter)>), dim3(1), dim3(kTailThreads),
                           sizeof(TailShared<SEG, int(kTailEnter)>), stream, p->st + iter, p->cl0,
                           p->cl1, p->cs0, p->cs1, p->nodes, roots, p->st + iter + 1);
    else
        hipLaunchKernelGGL((tail_kernel<SEG, kTailMax>), dim3(1), dim3(kTailThreads),
                           sizeof(TailShared<SEG, kTailMax>), stream, p->st + iter, p->cl0,
                           p->cl1, p->cs0, p->cs1, p->nodes, roots, p->st + iter + 1);
    hipLaunchKernelGGL(check_kernel, dim3(1), dim3(1), 0, stream, p->st + iter + 1, roots,
                       p->fail);
    return hipGetLastError();
}

// Eager build with host checks between batches of iterations; *iterations
// receives the number of multi-block iterations the data needed.
template <bool SEG>
ResultCode run_eager(PlocPlanImpl* p, const RTAabb* aabbs, const float* centers,
                     const uint32_t* scene_off_host, hipStream_t stream, uint32_t* iterations) {
    hipError_t e;
    if (SEG) {
        if ((e = hipMemcpyAsync(p->soff, scene_off_host, size_t(p->scenes + 1) * 4,
                                hipMemcpyHostToDevice, stream)) != hipSuccess)
            return hip_fail(e, "scene offsets");
    }
    if ((e = set_tail_lds<SEG>()) != hipSuccess) return hip_fail(e, "tail LDS size");
    if ((e = enqueue_front<SEG>(p, aabbs, centers, stream)) != hipSuccess)
        return hip_fail(e, "ploc sort");
    uint32_t known = p->n;
    uint32_t iter = 0;
    PlocState host_state;
    const uint32_t tail_at = tail_threshold(SEG ? p->scenes : 1u);
    uint32_t batch = kBatch;
    while (known > tail_at) {
        const uint32_t tiles = (known + kTile - 1) / kTile;
        for (uint32_t bt = 0; bt < batch; ++bt) {
            if (iter + 1 >= kMaxIterations) {
                set_error("clustering needed more than %u iterations", kMaxIterations);
                return Error;
            }
            enqueue_iteration<SEG>(p, iter, tiles, stream);
            ++iter;
        }
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "ploc iteration");
        if ((e = hipMemcpyAsync(&host_state, p->st + iter, sizeof(host_state),
                                hipMemcpyDeviceToHost, stream)) != hipSuccess)
            return hip_fail(e, "state read");
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "ploc iteration");
        if (host_state.n_active >= known) {
            set_error("clustering made no progress (non-finite bounds?)");
            return Error;
        }
        // the next batch: as many iterations as the shrink rate of this one
        // predicts to reach the tail, plus one (surplus iterations return at
        // once), so that a build needs few host round trips
        const double rate = std::pow(double(host_state.n_active) / double(known), 1.0 / batch);
        known = host_state.n_active;
        batch = kBatch;
        if (known > tail_at && rate > 0.0 && rate < 0.999) {
            const double need = std::log(double(tail_at) / double(known)) / std::log(rate);
            batch = uint32_t(std::min(64.0, std::max(1.0, std::ceil(need)))) + 1u;
        }
    }
    if ((e = enqueue_tail<SEG>(p, iter, stream)) != hipSuccess) return hip_fail(e, "tail launch");
    uint32_t fail = 1;
    if ((e = hipMemcpyAsync(&fail, p->fail, 4, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(&host_state, p->st + iter + 1, sizeof(host_state),
                            hipMemcpyDeviceToHost, stream)) != hipSuccess)
        return hip_fail(e, "state read");
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "ploc tail");
    if (fail) {
        set_error("clustering stopped at %u clusters (non-finite bounds?)", host_state.n_active);
        return Error;
    }
    if (iterations) {
        // iterations that did work: the first state small enough for the tail
        uint32_t need = 0;
        if (iter > 0) {
            std::vector<PlocState> states(iter + 1);
            if ((e = hipMemcpy(states.data(), p->st, sizeof(PlocState) * (iter + 1),
                               hipMemcpyDeviceToHost)) != hipSuccess)
                return hip_fail(e, "state read");
            while (need < iter && states[need].n_active > tail_at) ++need;
        }
        *iterations = need;
    }
    return Ok;
}

// n <= 2, single scene: one root leaf with the world bounds, into p->nodes
// and p->indices.
hipError_t enqueue_small(const PlocPlanImpl* p, const RTAabb* aabbs, const float* centers,
                         hipStream_t stream) {
    const DevAabb* in_aabbs = aabbs ? reinterpret_cast<const DevAabb*>(aabbs) : p->in_aabbs;
    if (!aabbs)
        hipLaunchKernelGGL(point_aabbs_kernel, dim3(1), dim3(kBlock), 0, stream, centers,
                           p->stride, p->n, p->in_aabbs);
    hipLaunchKernelGGL(bounds_partial_kernel, dim3(1), dim3(kBlock), 0, stream, in_aabbs, p->n,
                       p->part);
    hipLaunchKernelGGL(bounds_final_kernel, dim3(1), dim3(kBlock), 0, stream, p->part, 1u,
                       p->world);
    hipLaunchKernelGGL(small_root_kernel, dim3(1), dim3(1), 0, stream, p->world, p->n, p->nodes,
                       p->indices);
    return hipGetLastError();
}

ResultCode build_small(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                       uint32_t n, hipStream_t stream, DeviceBvh* out) {
    PlocPlanImpl* p = nullptr;
    ResultCode rc = alloc_plan(n, 1, aabbs == nullptr, stride_floats, &p, true);
    if (rc != Ok) return rc;
    hipError_t e = enqueue_small(p, aabbs, centers, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        free_plan(p, false);
        return hip_fail(e, "small build");
    }
    out->nodes = p->nodes;
    out->indices = p->indices;
    out->node_count = 1;
    out->prim_count = n;
    out->root_count = 1;
    free_plan(p, true);
    return Ok;
}

// Eager build of one scene (scene_off null) or a batch (scene_off: host
// offsets, scenes >= 2); the tree's arrays pass to *out.
ResultCode build_eager(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                       uint32_t n, const uint32_t* scene_off, uint32_t scenes, hipStream_t stream,
                       DeviceBvh* out) {
    if (scenes == 1 && n <= 2) return build_small(aabbs, centers, stride_floats, n, stream, out);
    PlocPlanImpl* p = nullptr;
    ResultCode rc = alloc_plan(n, scenes, aabbs == nullptr, stride_floats, &p, true);
    if (rc != Ok) return rc;
    rc = p->seg ? run_eager<true>(p, aabbs, centers, scene_off, stream, nullptr)
                : run_eager<false>(p, aabbs, centers, nullptr, stream, nullptr);
    if (rc != Ok) {
        free_plan(p, false);
        return rc;
    }
    out->nodes = p->nodes;
    out->indices = p->indices;
    out->node_count = p->node_count();
    out->prim_count = n;
    out->root_count = p->seg ? scenes : 1u;
    free_plan(p, true);
    return Ok;
}

}  // namespace

// ---- capturable builds -------------------------------------------------------

ResultCode ploc_plan_create(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                            const uint32_t* scene_off, uint32_t scenes, hipStream_t stream,
                            PlocPlanImpl** out, DeviceBvh* tree) {
    const uint32_t n = scene_off[scenes];
    PlocPlanImpl* p = nullptr;
    ResultCode rc = alloc_plan(n, scenes, aabbs == nullptr, stride_floats, &p, false);
    if (rc != Ok) return rc;
    uint32_t need = 0;
    hipError_t e = hipMemcpy(p->soff, scene_off, size_t(scenes + 1) * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_plan(p, false);
        return hip_fail(e, "scene offsets");
    }
    rc = p->seg ? run_eager<true>(p, aabbs, centers, scene_off, stream, &need)
                : run_eager<false>(p, aabbs, centers, nullptr, stream, &need);
    if (rc != Ok) {
        free_plan(p, false);
        return rc;
    }
    // headroom for later inputs of the same sizes (animated geometry)
    p->iterations = need + (need / 4 > 4 ? need / 4 : 4);
    if (p->iterations + 2 > kMaxIterations) p->iterations = kMaxIterations - 2;
    tree->nodes = p->nodes;
    tree->indices = p->indices;
    tree->parents = nullptr;
    tree->node_count = p->node_count();
    tree->prim_count = n;
    tree->root_count = p->seg ? scenes : 1u;
    *out = p;
    return Ok;
}

hipError_t ploc_plan_enqueue(const PlocPlanImpl* p, const RTAabb* aabbs, const float* centers,
                             hipStream_t stream) {
    if (aabbs == nullptr && !p->point_aabbs) return hipErrorInvalidValue;
    hipError_t e;
    if (p->seg) {
        if ((e = enqueue_front<true>(p, aabbs, centers, stream)) != hipSuccess) return e;
        for (uint32_t k = 0; k < p->iterations; ++k)
            enqueue_iteration<true>(p, k, p->ploc_tiles, stream);
        return enqueue_tail<true>(p, p->iterations, stream);
    }
    if ((e = enqueue_front<false>(p, aabbs, centers, stream)) != hipSuccess) return e;
    for (uint32_t k = 0; k < p->iterations; ++k)
        enqueue_iteration<false>(p, k, p->ploc_tiles, stream);
    return enqueue_tail<false>(p, p->iterations, stream);
}

const uint32_t* ploc_plan_fail_word(const PlocPlanImpl* p) { return p->fail; }
const uint32_t* ploc_plan_scene_offsets(const PlocPlanImpl* p) { return p->soff; }
uint32_t ploc_plan_iterations(const PlocPlanImpl* p) { return p->iterations; }
void ploc_plan_destroy(PlocPlanImpl* p) { free_plan(p, false); }

ResultCode gpu_build_ploc(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                          size_t n64, int radius, hipStream_t stream, DeviceBvh* out) {
    if (radius != kRadius) {
        set_error("PLOC search radius %d not supported (built for %d)", radius, kRadius);
        return InvalidArgument;
    }
    if (n64 == 0) return NoPrimitives;
    if (n64 > (1u << 30)) {
        set_error("prim_count %zu too large", n64);
        return InvalidArgument;
    }
    return build_eager(aabbs, centers, stride_floats, uint32_t(n64), nullptr, 1u, stream, out);
}

size_t ploc_workspace_bytes(size_t n, bool point_aabbs) {
    PlocPlanImpl p;
    init_plan(&p, uint32_t(n), 1, point_aabbs, 3);
    return carve_plan(&p, nullptr);
}

ResultCode gpu_build_ploc_into(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                               uint32_t n, void* ws, size_t ws_bytes, RTBvhNode* nodes,
                               uint32_t* indices, hipStream_t stream) {
    PlocPlanImpl p;  // every buffer is the caller's: nothing to free
    init_plan(&p, n, 1, aabbs == nullptr, stride_floats);
    const size_t need = carve_plan(&p, nullptr);
    if (need > ws_bytes) {
        set_error("the build of %u primitives needs %zu bytes of workspace (%zu given)", n, need,
                  ws_bytes);
        return InvalidArgument;
    }
    carve_plan(&p, static_cast<char*>(ws));
    p.nodes = reinterpret_cast<DevAabb*>(nodes);
    p.indices = indices;
    if (n <= 2) {
        hipError_t e = enqueue_small(&p, aabbs, centers, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        return e == hipSuccess ? Ok : hip_fail(e, "small build");
    }
    return run_eager<false>(&p, aabbs, centers, nullptr, stream, nullptr);
}

ResultCode check_batch_offsets(const uint32_t* scene_off, uint32_t scenes) {
    if (scenes == 0 || scene_off == nullptr) {
        set_error("no scenes");
        return NoPrimitives;
    }
    if (scenes > kMaxBatchScenes) {
        set_error("a batch holds at most %u scenes (got %u)", kMaxBatchScenes, scenes);
        return InvalidArgument;
    }
    if (scene_off[0] != 0) {
        set_error("scene_offsets[0] must be 0");
        return InvalidArgument;
    }
    for (uint32_t s = 0; s < scenes; ++s) {
        if (scene_off[s + 1] < scene_off[s] || scene_off[s + 1] - scene_off[s] < 3) {
            set_error("scene %u has %lld primitives; a batched scene needs at least 3", s,
                      (long long)scene_off[s + 1] - (long long)scene_off[s]);
            return InvalidArgument;
        }
    }
    if (scene_off[scenes] > (1u << 30)) {
        set_error("batch of %u primitives too large", scene_off[scenes]);
        return InvalidArgument;
    }
    return Ok;
}

ResultCode gpu_build_ploc_batch(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                                const uint32_t* scene_off, uint32_t scenes, hipStream_t stream,
                                DeviceBvh* out) {
    ResultCode rc = check_batch_offsets(scene_off, scenes);
    if (rc != Ok) return rc;
    const uint32_t n = scene_off[scenes];
    return build_eager(aabbs, centers, stride_floats, n, scenes > 1 ? scene_off : nullptr, scenes,
                       stream, out);
}

}  // namespace rtbvh
