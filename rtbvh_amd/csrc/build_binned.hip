// GPU binned-SAH builder for gfx950, breadth-first.
//
// Restates BinnedSahBuildTask::run (src/builders/binned_sah.rs:132-282) per
// node, one tree level per step, all nodes of a level in parallel:
//   pad the node box by 1e-4                                  :133
//   leaf when it holds <= 1 primitive or depth >= 64          :140-143
//   16 bins per axis over the padded node box; a primitive's
//   bin is min(15, sat(max(c*ctb + bo, 0)))                   :119-128,145-172
//   SAH sweep (right costs, then left + right, strict '<')    :80-114
//   axis ties keep the lower axis                             :179-185
//   no split / cost >= A*(n-1): approximate median on the
//   longest axis when n > max_leaf, else leaf                 :187-210
//   a side left empty makes a leaf                            :221-222
//   child boxes are unions of the bins on either side         :228-247
// Two reference defects are fixed so that every child box lies inside its
// parent: the left box uses the split actually taken (the reference uses the
// pre-fallback bin count, :232-235) and leaves are not padded a second time
// (:135).  The partition is stable on both sides (the reference's Lomuto
// partition, src/utils.rs:76-96, permutes the right side), so primitive order
// inside multi-primitive leaves may differ; the sets, the topology and the
// bounds do not.  Nodes are numbered breadth-first (the reference's numbering
// depends on thread timing, builders/mod.rs:59-76).
//
// A level is four kernels.  Its task count, first node id and number of
// binned tasks sit in a device record that the level before writes, so the
// host enqueues levels in batches and reads one record per batch:
//   bin_kernel        primitives of tasks with more than 16 into the task's
//                     48 bins (LDS atomics per block, then one global merge)
//   decide_kernel     per task the SAH decision (a 16-lane group per task)
//   children_kernel   per tile of 1024 tasks the splitting tasks and binned
//                     children before the tile (a decoupled look-back over
//                     tile words), then node records, child tasks and the
//                     next level's record
//   partition_kernel  the stable partition of every splitting task's range:
//                     ranks of left primitives from a segmented scan over
//                     tiles of 1024 primitives (decoupled look-back across
//                     tiles of one task); the primitives of a leaf are
//                     written to both index buffers and drop out
// Look-back waits are bounded: tile ids come from a counter taken by running
// blocks, a tile waits only on lower ids, and every tile publishes a word
// before it waits on any.
#include <algorithm>

#include "internal.hpp"
#include "lookback.hpp"

namespace rtbvh {

namespace {

constexpr int kBins = 16;           // binned_sah.rs:314
constexpr uint32_t kMaxDepth = 64;  // binned_sah.rs:313
// Levels that hold tasks: depths 0..64 (a task at depth 64 is a leaf).
constexpr uint32_t kMaxLevels = kMaxDepth + 1;
constexpr int kBlock = 256;
// Tasks of at most kSmallTask primitives are binned by their split group
// (through LDS) instead of by bin_kernel's atomics.
constexpr uint32_t kSmallTask = 64;
// The first kHotSlots slots (the levels near the root, where every block of
// bin_kernel merges into the same few bins) have kCopies copies, one per
// XCD's blocks, so that fewer atomics queue on one address; the decision
// folds the copies.
constexpr uint32_t kHotSlots = 64;
constexpr uint32_t kCopies = 8;
constexpr uint32_t kDecideGrid = 4096;      // most decide_kernel blocks (grid stride)
constexpr uint32_t kTaskTile = 1024;        // children_kernel: tasks per tile, one per thread
constexpr uint32_t kChildGrid = 512;        // most children_kernel blocks
constexpr int kPartBlock = 1024;            // partition_kernel
constexpr uint32_t kChunks = 4;             // partition tile = kChunks runs of kPartBlock
constexpr uint32_t kPrimTile = kChunks * kPartBlock;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kLeaf = 0xFFFFFFFFu;

// A task of one level: its node's padded box and bin parameters.
struct BTask {
    float mn[3];     // padded node box, binned_sah.rs:133
    uint32_t begin;  // primitives [begin, end) of the index array
    float mx[3];
    uint32_t end;
    float ctb[3];    // (1 / extent) * 16, :146
    uint32_t depth;
    float bo[3];     // -min * ctb, :147
    uint32_t slot;   // bins of bin_kernel, or kNoSlot
};

// Device record of one level, written by the level before (the root's by
// root_bounds_kernel); records past the last level hold no tasks and the
// final node count.
struct BLevel {
    uint32_t num_tasks;
    uint32_t node_base;     // task t is node node_base + t
    uint32_t slots;         // tasks binned by bin_kernel: slots [0, slots)
    uint32_t split_ticket;  // children_kernel's next tile
    uint32_t part_ticket;   // partition_kernel's next tile
    uint32_t pad[3];
};

// What partition_kernel needs of a task.
struct BPart {
    uint32_t begin;
    uint32_t left_count;
    uint32_t axis;  // kLeaf: the task is a leaf
    uint32_t split_index;
    float ctb, bo;   // of the split axis
    uint32_t child;  // next-level task of the left child; the right one is child + 1
    uint32_t pad;
};

// One bin, all zero when empty: the count and the bounds as order-preserving
// encodings, the minimum complemented so that both fold by maximum.
struct BinAcc {
    uint32_t cnt;
    uint32_t mn[3];  // ~fenc(min)
    uint32_t mx[3];  // fenc(max)
    uint32_t pad;
};

// A task's decision, held in LDS until the tile's prefix is known.
struct SDec {
    uint32_t split, axis, split_index, left_count;
    float lmn[3], lmx[3], rmn[3], rmx[3];
};

__device__ __forceinline__ uint32_t fenc(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t e) {
    return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}
// Aabb::empty (1e34, -1e34) joined with the bounds folded into a word
// (min and max over the encodings, NaN bounds never folded in).
__device__ __forceinline__ float dec_min(uint32_t w) {
    const float v = fdec(~w);
    return (w != 0 && v < 1e34f) ? v : 1e34f;
}
__device__ __forceinline__ float dec_max(uint32_t w) {
    const float v = fdec(w);
    return (w != 0 && v > -1e34f) ? v : -1e34f;
}

__device__ __forceinline__ uint32_t bin_index(float c, float ctb, float bo) {
    // (bin_count - 1).min((c * ctb + bo).max(0.0) as usize)
    const float b = c * ctb + bo;
    const float m = rs_max(b, 0.0f);
    uint32_t u;
    if (!(m > 0.0f))
        u = 0;
    else if (m >= 4294967296.0f)
        u = 0xFFFFFFFFu;
    else
        u = uint32_t(m);
    return u < uint32_t(kBins - 1) ? u : uint32_t(kBins - 1);
}

__device__ __forceinline__ float half_area6(const float* mn, const float* mx) {
    const float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return (d0 + d1) * d2 + d0 * d1;
}

// a[i] for i in 0..2 by selects (an indexed register array would go to
// scratch memory)
template <typename T>
__device__ __forceinline__ T pick3(const T (&a)[3], uint32_t i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}

__device__ __forceinline__ bool binned_by_atomics(uint32_t count, uint32_t depth) {
    return count > kSmallTask && depth < kMaxDepth;
}

// The 48 bins of a slot (copy: the caller's copy of a hot slot).
__device__ __forceinline__ BinAcc* slot_bins(BinAcc* bins, BinAcc* hot, uint32_t slot,
                                             uint32_t copy) {
    return slot < kHotSlots ? hot + (size_t(slot) * kCopies + copy) * 3 * kBins
                            : bins + size_t(slot) * 3 * kBins;
}

// A task from its unpadded box: the padding and bin parameters of
// binned_sah.rs:133,146-147.
__device__ __forceinline__ BTask make_task(const float* mn, const float* mx, uint32_t begin,
                                          uint32_t end, uint32_t depth, uint32_t slot) {
    BTask t;
    for (int k = 0; k < 3; ++k) {
        t.mn[k] = mn[k] - 0.0001f;
        t.mx[k] = mx[k] + 0.0001f;
        t.ctb[k] = (1.0f / (t.mx[k] - t.mn[k])) * float(kBins);
        t.bo[k] = (-t.mn[k]) * t.ctb[k];
    }
    t.begin = begin;
    t.end = end;
    t.depth = depth;
    t.slot = slot;
    return t;
}

// Row (16-lane) moves by DPP: a group of 16 lanes is one DPP row, so the
// scans and reductions of a group stay in the VALU.  row_shr:k gives lane i
// the value of lane i - k, row_shl:k that of lane i + k (lanes without a
// source keep their own value, which the callers never use);
// quad_perm [1,0,3,2] and [2,3,0,1] pair lanes i ^ 1 and i ^ 2,
// row_half_mirror lanes i and 7 - i of each half, row_mirror i and 15 - i.
constexpr int kDppShl = 0x100, kDppShr = 0x110, kDppXor1 = 0xB1, kDppXor2 = 0x4E,
              kDppHalfMirror = 0x141, kDppMirror = 0x140;

template <int kCtrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), kCtrl, 0xF, 0xF, false));
}
template <int kCtrl>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), kCtrl, 0xF, 0xF, false));
}
// Maximum and minimum over a row of 16 lanes (every lane of the row gets it).
__device__ __forceinline__ uint32_t row_max(uint32_t v) {
    v = max(v, dpp_u32<kDppXor1>(v));
    v = max(v, dpp_u32<kDppXor2>(v));
    v = max(v, dpp_u32<kDppHalfMirror>(v));
    return max(v, dpp_u32<kDppMirror>(v));
}
__device__ __forceinline__ uint32_t row_min(uint32_t v) {
    v = min(v, dpp_u32<kDppXor1>(v));
    v = min(v, dpp_u32<kDppXor2>(v));
    v = min(v, dpp_u32<kDppHalfMirror>(v));
    return min(v, dpp_u32<kDppMirror>(v));
}

// A block's next tile of a persistent loop.
__device__ __forceinline__ uint32_t next_tile(uint32_t* ticket, uint32_t* s_tile) {
    if (threadIdx.x == 0) *s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    return *s_tile;
}

// Bin every primitive of every task with a bin slot (more than kSmallTask
// primitives; smaller tasks with two or more are binned in decide_kernel).
// Slots are numbered in task order and tasks in range order, so the slots
// met along the index array never decrease; finished primitives carry -1
// and sit between them.  A block walks its run of the index array in chunks
// of kBlock and accumulates in an LDS window of kLdsTasks slots, merged into
// the global bins (one set of atomics per slot and bin) when a chunk reaches
// past it and at the end; a chunk that alone spans more slots than the
// window goes to the global bins directly.
constexpr int kLdsTasks = 16;  // 4: 41 ms spatial, 22 ms binned at 1M; 16: 36, 19.5; 32: no better
constexpr uint32_t kBinGrid = 1024;  // blocks of bin_kernel (runs of n / 1024)
constexpr uint32_t kNone = 0xFFFFFFFFu;

__global__ void __launch_bounds__(kBlock)
    bin_kernel(const BLevel* levels, uint32_t level, uint32_t n, uint32_t run,
               const int32_t* ptask, const uint32_t* idx, const BTask* tasks,
               const float* centers, uint32_t stride, const DevAabb* aabbs, BinAcc* bins,
               BinAcc* hot) {
    if (levels[level].slots == 0) return;
    constexpr int kSlots = kLdsTasks * 3 * kBins;
    __shared__ uint32_t s_cnt[kSlots];
    __shared__ uint32_t s_mn[kSlots][3];
    __shared__ uint32_t s_mx[kSlots][3];
    // first and last slot of each wave's slotted lanes, two chunks in turn
    __shared__ uint32_t s_first[2][kBlock / 64], s_last[2][kBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    auto clear = [&]() {
        for (uint32_t k = threadIdx.x; k < uint32_t(kSlots); k += kBlock) {
            s_cnt[k] = 0;
            for (int q = 0; q < 3; ++q) {
                s_mn[k][q] = 0;
                s_mx[k][q] = 0;
            }
        }
    };
    const uint32_t copy = blockIdx.x % kCopies;
    auto merge = [&](uint32_t wlo) {
        for (uint32_t k = threadIdx.x; k < uint32_t(kSlots); k += kBlock) {
            if (s_cnt[k] == 0) continue;
            BinAcc* g = slot_bins(bins, hot, wlo + k / (3 * kBins), copy) + k % (3 * kBins);
            atomicAdd(&g->cnt, s_cnt[k]);
            for (int q = 0; q < 3; ++q) {
                if (s_mn[k][q]) atomicMax(&g->mn[q], s_mn[k][q]);
                if (s_mx[k][q]) atomicMax(&g->mx[q], s_mx[k][q]);
            }
        }
    };
    clear();
    __syncthreads();
    uint32_t wlo = kNone;  // first slot of the LDS window (kNone: empty)
    const uint32_t b0 = blockIdx.x * run;
    const uint32_t b1 = b0 + run < n ? b0 + run : n;
    for (uint32_t c0 = b0; c0 < b1; c0 += kBlock) {
        const uint32_t i = c0 + threadIdx.x;
        const int32_t t = i < b1 ? ptask[i] : -1;
        BTask tk;
        uint32_t slot = kNoSlot;
        if (t >= 0) {
            tk = tasks[t];
            slot = tk.slot;
        }
        // the chunk's slot range, from the first and last slotted lane of
        // each wave (slots never decrease along the array)
        const uint64_t m = __ballot(slot != kNoSlot);
        const uint32_t f = __shfl(slot, m ? int(__builtin_ctzll(m)) : 0, 64);
        const uint32_t l = __shfl(slot, m ? int(63 - __builtin_clzll(m)) : 0, 64);
        const uint32_t par = ((c0 - b0) / kBlock) & 1u;
        if (lane == 0) {
            s_first[par][wave] = m ? f : kNone;
            s_last[par][wave] = m ? l : kNone;
        }
        __syncthreads();
        uint32_t lo = kNone, hi = kNone;
        for (uint32_t w = 0; w < kBlock / 64; ++w) {
            if (lo == kNone) lo = s_first[par][w];
            if (s_last[par][w] != kNone) hi = s_last[par][w];
        }
        if (lo == kNone) continue;  // no slotted primitive in the chunk
        if (wlo != kNone && hi - wlo >= uint32_t(kLdsTasks)) {
            merge(wlo);
            __syncthreads();
            clear();
            __syncthreads();
            wlo = kNone;
        }
        if (wlo == kNone && hi - lo < uint32_t(kLdsTasks)) wlo = lo;
        const bool local = wlo != kNone;  // else the chunk alone spans more than the window
        // a primitive's bounds as bin words (0: nothing, NaN bounds skipped)
        uint32_t emn[3] = {0, 0, 0}, emx[3] = {0, 0, 0};
        float cc[3] = {0.0f, 0.0f, 0.0f};
        if (slot != kNoSlot) {
            const uint32_t p = idx[i];
            const float* c = centers + size_t(p) * stride;
            const DevAabb a = aabbs[p];
            for (int k = 0; k < 3; ++k) {
                cc[k] = c[k];
                if (a.min[k] == a.min[k]) emn[k] = ~fenc(a.min[k]);
                if (a.max[k] == a.max[k]) emx[k] = fenc(a.max[k]);
            }
        }
        // A row of 16 lanes whose primitives share slot and bin on an axis
        // (neighbours in the array often do near the root) adds its row
        // totals with one lane; other rows add per lane.
        uint32_t rmn[3], rmx[3];
        for (int k = 0; k < 3; ++k) {
            rmn[k] = row_max(emn[k]);
            rmx[k] = row_max(emx[k]);
        }
        const uint32_t rowm = uint32_t(m >> (lane & 48u)) & 0xFFFFu;  // slotted lanes of the row
        const uint32_t rlead = rowm ? uint32_t(__builtin_ctz(rowm)) : 16u;
        for (int ax = 0; ax < 3; ++ax) {
            const uint32_t b = slot != kNoSlot ? bin_index(cc[ax], tk.ctb[ax], tk.bo[ax]) : 0u;
            const uint32_t key = slot != kNoSlot ? (slot << 4) | b : 0u;  // slot < 2^24
            const bool uni = rowm != 0 &&
                             row_min(slot != kNoSlot ? key : 0xFFFFFFFFu) == row_max(key);
            const bool one = uni ? (lane & 15u) == rlead : slot != kNoSlot;
            if (one) {
                const uint32_t add = uni ? uint32_t(__builtin_popcount(rowm)) : 1u;
                const uint32_t* wmn = uni ? rmn : emn;
                const uint32_t* wmx = uni ? rmx : emx;
                const uint32_t s = ax * kBins + b;
                if (local) {
                    const uint32_t ls = (slot - wlo) * 3 * kBins + s;
                    atomicAdd(&s_cnt[ls], add);
                    for (int k = 0; k < 3; ++k) {
                        if (wmn[k]) atomicMax(&s_mn[ls][k], wmn[k]);
                        if (wmx[k]) atomicMax(&s_mx[ls][k], wmx[k]);
                    }
                } else {
                    BinAcc* g = slot_bins(bins, hot, slot, copy) + s;
                    atomicAdd(&g->cnt, add);
                    for (int k = 0; k < 3; ++k) {
                        if (wmn[k]) atomicMax(&g->mn[k], wmn[k]);
                        if (wmx[k]) atomicMax(&g->mx[k], wmx[k]);
                    }
                }
            }
        }
    }
    __syncthreads();
    if (wlo != kNone) merge(wlo);
}

// A run of bins: its box and its primitive count.
struct GBox {
    float mn[3], mx[3];
    uint32_t cnt;
};

// lo := lo joined with hi, where hi holds higher bins.  The minimum keeps
// lo's bound unless hi's is strictly lower (rs_min/rs_max keep the first
// argument on equal values), which is what a fold in increasing bin order
// gives: joins in any grouping return the same bits.
__device__ __forceinline__ void gbox_join(GBox& lo, const GBox& hi) {
    for (int k = 0; k < 3; ++k) {
        lo.mn[k] = rs_min(lo.mn[k], hi.mn[k]);
        lo.mx[k] = rs_max(lo.mx[k], hi.mx[k]);
    }
    lo.cnt += hi.cnt;
}

template <int kCtrl>
__device__ __forceinline__ GBox gbox_dpp(const GBox& b) {
    GBox r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = dpp_f32<kCtrl>(b.mn[k]);
        r.mx[k] = dpp_f32<kCtrl>(b.mx[k]);
    }
    r.cnt = dpp_u32<kCtrl>(b.cnt);
    return r;
}

// One step of the prefix (p, bins 0..lane) and suffix (q, bins lane..15)
// scans of a group.
template <int kOff>
__device__ __forceinline__ void scan_step(GBox& p, GBox& q, int lane) {
    GBox lower = gbox_dpp<kDppShr + kOff>(p);
    const GBox higher = gbox_dpp<kDppShl + kOff>(q);
    if (lane >= kOff) {
        gbox_join(lower, p);
        p = lower;
    }
    if (lane + kOff < kBins) gbox_join(q, higher);
}

// One step of the group's (cost, bin) minimum, ties to the lower bin.
template <int kCtrl>
__device__ __forceinline__ void argmin_step(float& key, int& at) {
    const float ok = dpp_f32<kCtrl>(key);
    const int oi = int(dpp_u32<kCtrl>(uint32_t(at)));
    if (ok < key || (ok == key && oi < at)) {
        key = ok;
        at = oi;
    }
}

__device__ __forceinline__ GBox gbox_shfl(const GBox& b, int src) {
    GBox r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = __shfl(b.mn[k], src, kBins);
        r.mx[k] = __shfl(b.mx[k], src, kBins);
    }
    r.cnt = __shfl(b.cnt, src, kBins);
    return r;
}

__device__ __forceinline__ float gbox_half_area(const GBox& b) { return half_area6(b.mn, b.mx); }

// decide() for a task of two primitives, restated for two bins without the
// group's scans (every lane computes it; the values are those of the scans
// bit for bit).  A primitive's bin box is its box, a NaN bound and one past
// 1e34 in magnitude giving Aabb::empty's (dec_min / dec_max of a one-value
// fold).  On an axis where the two bins differ, every plane between them
// costs area(lower) + area(upper) and the first one, after the lower bin,
// wins; where they share a bin no plane has both sides.  The fallback's
// first bin whose prefix reaches 2/5 of two primitives (one) is the lower
// bin on the longest axis, when it is below the last bin.
__device__ __forceinline__ SDec decide_pair(const BTask& tk, uint32_t max_leaf,
                                            const uint32_t* idx, const float* centers,
                                            uint32_t stride, const DevAabb* aabbs) {
    SDec d;
    d.split = 0;
    d.axis = 0;
    d.split_index = 0;
    d.left_count = 0;
    uint32_t bin[2][3];
    float mn[2][3], mx[2][3];
    for (int q = 0; q < 2; ++q) {
        const uint32_t p = idx[tk.begin + uint32_t(q)];
        const float* c = centers + size_t(p) * stride;
        const DevAabb a = aabbs[p];
        for (int k = 0; k < 3; ++k) {
            bin[q][k] = bin_index(c[k], tk.ctb[k], tk.bo[k]);
            mn[q][k] = (a.min[k] == a.min[k] && a.min[k] < 1e34f) ? a.min[k] : 1e34f;
            mx[q][k] = (a.max[k] == a.max[k] && a.max[k] > -1e34f) ? a.max[k] : -1e34f;
        }
    }
    const float area0 = half_area6(mn[0], mx[0]), area1 = half_area6(mn[1], mx[1]);
    float cost[3];
    uint32_t cnt[3];
    for (int k = 0; k < 3; ++k) {
        cost[k] = 3.40282347e38f;
        cnt[k] = kBins;
        if (bin[0][k] != bin[1][k]) {
            const bool first_low = bin[0][k] < bin[1][k];
            const float c = first_low ? area0 * 1.0f + area1 * 1.0f : area1 * 1.0f + area0 * 1.0f;
            if (c < 3.40282347e38f) {
                cost[k] = c;
                cnt[k] = (first_low ? bin[0][k] : bin[1][k]) + 1u;
            }
        }
    }
    uint32_t axis = 0;
    if (cost[0] > cost[1]) axis = 1;
    if (pick3(cost, axis) > cost[2]) axis = 2;
    uint32_t split_index = pick3(cnt, axis);
    const float max_split_cost = half_area6(tk.mn, tk.mx) * (2.0f - 1.0f);
    if (pick3(cnt, axis) == uint32_t(kBins) || pick3(cost, axis) >= max_split_cost) {
        if (2u <= max_leaf) return d;
        // Aabb::longest_axis of the padded box, aabb.rs:354-363
        const float e0 = tk.mx[0] - tk.mn[0], e1 = tk.mx[1] - tk.mn[1], e2 = tk.mx[2] - tk.mn[2];
        axis = 0;
        if (e1 > e0) axis = 1;
        if (e2 > (axis == 1 ? e1 : e0)) axis = 2;
        const uint32_t lowest = min(pick3(bin[0], axis), pick3(bin[1], axis));
        if (lowest < uint32_t(kBins) - 1u) split_index = lowest + 1u;
    }
    const uint32_t lim = split_index < uint32_t(kBins) ? split_index : uint32_t(kBins);
    const bool l0 = pick3(bin[0], axis) < lim, l1 = pick3(bin[1], axis) < lim;
    if (l0 != l1) {  // one primitive on each side
        d.split = 1;
        d.axis = axis;
        d.split_index = split_index;
        d.left_count = 1;
        for (int k = 0; k < 3; ++k) {
            d.lmn[k] = l0 ? mn[0][k] : mn[1][k];
            d.lmx[k] = l0 ? mx[0][k] : mx[1][k];
            d.rmn[k] = l0 ? mn[1][k] : mn[0][k];
            d.rmx[k] = l0 ? mx[1][k] : mx[0][k];
        }
    }
    return d;
}

// The decision for one task by its group of 16 lanes (every lane returns
// it): the best SAH plane per axis (find_split, binned_sah.rs:80-114), the
// axis, the fallback and the child boxes (:150-230).  Lane i holds bin i:
// inclusive prefix (bins 0..i) and suffix (bins i..15) runs by DPP row
// scans, the cost of the plane after bin i as the one-thread sweep computes
// it, half_area(left) * left count + half_area(right) * right count, and the
// first lowest cost below FLT_MAX by a reduction.  The bins' bounds hold no
// NaN (the binning skips NaN bounds), so every bound equals the sequential
// fold's, up to the sign of a zero in the right runs of the sweep, which no
// comparison tells apart; the child boxes come from increasing-order runs,
// as the sweep builds them.  A task's bins are left empty for the slot's
// next owner.
__device__ __forceinline__ SDec decide(const BTask& tk, int lane, BinAcc* bins, BinAcc* hot,
                                       uint32_t max_leaf, const uint32_t* idx,
                                       const float* centers, uint32_t stride,
                                       const DevAabb* aabbs, uint4* gather) {
    SDec d;
    d.split = 0;
    d.axis = 0;
    d.split_index = 0;
    d.left_count = 0;
    const uint32_t work = tk.end - tk.begin;
    if (!(work > 1 && tk.depth < kMaxDepth)) return d;
    if (work == 2) return decide_pair(tk, max_leaf, idx, centers, stride, aabbs);
    BinAcc own[3];  // this lane's bin of each axis
    if (tk.slot == kNoSlot) {
        // at most kSmallTask primitives: bin_kernel's accumulation without
        // atomics; the group stages its primitives in LDS (encoded minimum
        // with the three bins packed in w, encoded maximum) and lane j
        // gathers bin j
        for (uint32_t q = uint32_t(lane); q < work; q += kBins) {
            const uint32_t p = idx[tk.begin + q];
            const float* c = centers + size_t(p) * stride;
            const DevAabb a = aabbs[p];
            uint4 pmn = make_uint4(0, 0, 0, 0), pmx = pmn;
            pmn.w = bin_index(c[0], tk.ctb[0], tk.bo[0]) |
                    (bin_index(c[1], tk.ctb[1], tk.bo[1]) << 4) |
                    (bin_index(c[2], tk.ctb[2], tk.bo[2]) << 8);
            if (a.min[0] == a.min[0]) pmn.x = ~fenc(a.min[0]);
            if (a.min[1] == a.min[1]) pmn.y = ~fenc(a.min[1]);
            if (a.min[2] == a.min[2]) pmn.z = ~fenc(a.min[2]);
            if (a.max[0] == a.max[0]) pmx.x = fenc(a.max[0]);
            if (a.max[1] == a.max[1]) pmx.y = fenc(a.max[1]);
            if (a.max[2] == a.max[2]) pmx.z = fenc(a.max[2]);
            gather[2 * q] = pmn;
            gather[2 * q + 1] = pmx;
        }
        // the group's lanes are one wave's: its LDS operations run in order
        __builtin_amdgcn_wave_barrier();
        for (int ax = 0; ax < 3; ++ax) {
            own[ax].cnt = 0;
            for (int k = 0; k < 3; ++k) {
                own[ax].mn[k] = 0;
                own[ax].mx[k] = 0;
            }
        }
        for (uint32_t q = 0; q < work; ++q) {
            const uint4 qmn = gather[2 * q], qmx = gather[2 * q + 1];
            const uint32_t qmnv[3] = {qmn.x, qmn.y, qmn.z};
            const uint32_t qmxv[3] = {qmx.x, qmx.y, qmx.z};
            for (int ax = 0; ax < 3; ++ax)
                if (((qmn.w >> (4 * ax)) & 15u) == uint32_t(lane)) {
                    ++own[ax].cnt;
                    for (int k = 0; k < 3; ++k) {
                        own[ax].mn[k] = max(own[ax].mn[k], qmnv[k]);
                        own[ax].mx[k] = max(own[ax].mx[k], qmxv[k]);
                    }
                }
        }
        __builtin_amdgcn_wave_barrier();
    } else {
        // fold the slot's copies; leave them empty for the slot's next owner
        BinAcc empty;
        empty.cnt = 0;
        for (int k = 0; k < 3; ++k) empty.mn[k] = empty.mx[k] = 0;
        empty.pad = 0;
        for (int ax = 0; ax < 3; ++ax) own[ax] = empty;
        const uint32_t copies = tk.slot < kHotSlots ? kCopies : 1u;
        for (uint32_t cp = 0; cp < copies; ++cp) {
            BinAcc* tb = slot_bins(bins, hot, tk.slot, cp);
            for (int ax = 0; ax < 3; ++ax) {
                const BinAcc v = tb[ax * kBins + lane];
                own[ax].cnt += v.cnt;
                for (int k = 0; k < 3; ++k) {
                    own[ax].mn[k] = max(own[ax].mn[k], v.mn[k]);
                    own[ax].mx[k] = max(own[ax].mx[k], v.mx[k]);
                }
                tb[ax * kBins + lane] = empty;
            }
        }
    }
    // per axis: prefix and suffix runs, best plane cost and bin count
    auto scan_axis = [&](const BinAcc& own_bin, GBox& pre, GBox& suf, float& cost,
                         uint32_t& cnt) {
        GBox b;
        for (int k = 0; k < 3; ++k) {
            b.mn[k] = dec_min(own_bin.mn[k]);
            b.mx[k] = dec_max(own_bin.mx[k]);
        }
        b.cnt = own_bin.cnt;
        GBox p = b, q = b;
        scan_step<1>(p, q, lane);
        scan_step<2>(p, q, lane);
        scan_step<4>(p, q, lane);
        scan_step<8>(p, q, lane);
        pre = p;
        suf = q;
        const GBox right = gbox_dpp<kDppShl + 1>(q);
        float key = __builtin_inff();
        if (lane < kBins - 1) {
            const float c = gbox_half_area(p) * float(p.cnt) +
                            gbox_half_area(right) * float(right.cnt);
            if (c < 3.40282347e38f) key = c;
        }
        int at = lane;
        argmin_step<kDppXor1>(key, at);
        argmin_step<kDppXor2>(key, at);
        argmin_step<kDppHalfMirror>(key, at);
        argmin_step<kDppMirror>(key, at);
        if (key == __builtin_inff()) {
            cost = 3.40282347e38f;
            cnt = kBins;
        } else {
            cost = key;
            cnt = uint32_t(at + 1);
        }
    };
    GBox pre0, pre1, pre2, suf0, suf1, suf2;
    float cost[3];
    uint32_t cnt[3];
    scan_axis(own[0], pre0, suf0, cost[0], cnt[0]);
    scan_axis(own[1], pre1, suf1, cost[1], cnt[1]);
    scan_axis(own[2], pre2, suf2, cost[2], cnt[2]);
    // member-wise selects keep the runs in registers
    auto gsel = [](uint32_t i, const GBox& a0, const GBox& a1, const GBox& a2) {
        GBox r;
        for (int k = 0; k < 3; ++k) {
            r.mn[k] = i == 0 ? a0.mn[k] : (i == 1 ? a1.mn[k] : a2.mn[k]);
            r.mx[k] = i == 0 ? a0.mx[k] : (i == 1 ? a1.mx[k] : a2.mx[k]);
        }
        r.cnt = i == 0 ? a0.cnt : (i == 1 ? a1.cnt : a2.cnt);
        return r;
    };
    uint32_t axis = 0;
    if (cost[0] > cost[1]) axis = 1;
    if (pick3(cost, axis) > cost[2]) axis = 2;
    uint32_t split_index = pick3(cnt, axis);
    const float max_split_cost = half_area6(tk.mn, tk.mx) * (float(work) - 1.0f);
    if (pick3(cnt, axis) == uint32_t(kBins) || pick3(cost, axis) >= max_split_cost) {
        if (work <= max_leaf) return d;
        // Aabb::longest_axis of the padded box, aabb.rs:354-363
        const float e0 = tk.mx[0] - tk.mn[0], e1 = tk.mx[1] - tk.mn[1], e2 = tk.mx[2] - tk.mn[2];
        axis = 0;
        if (e1 > e0) axis = 1;
        if (e2 > (axis == 1 ? e1 : e0)) axis = 2;
        // the first bin whose prefix count reaches 2/5 of the work
        const uint32_t run = axis == 0 ? pre0.cnt : (axis == 1 ? pre1.cnt : pre2.cnt);
        const bool reach = lane < kBins - 1 && uint64_t(run) >= uint64_t(work * 2 / 5 + 1);
        const uint32_t base = threadIdx.x & 63u & ~uint32_t(kBins - 1);  // group's lane 0
        const uint32_t m = uint32_t(__ballot(reach) >> base) & 0xFFFFu;
        if (m) split_index = uint32_t(__builtin_ctz(m) + 1);
    }
    const uint32_t lim = split_index < uint32_t(kBins) ? split_index : uint32_t(kBins);
    const GBox lb = gbox_shfl(gsel(axis, pre0, pre1, pre2), lim > 0 ? int(lim) - 1 : 0);
    GBox rb = gbox_shfl(gsel(axis, suf0, suf1, suf2), lim < uint32_t(kBins) ? int(lim) : 0);
    const uint32_t left = lim > 0 ? lb.cnt : 0u;
    if (lim >= uint32_t(kBins))
        for (int k = 0; k < 3; ++k) {
            rb.mn[k] = 1e34f;  // Aabb::empty
            rb.mx[k] = -1e34f;
        }
    if (left > 0 && left < work) {
        d.split = 1;
        d.axis = axis;
        d.split_index = split_index;
        d.left_count = left;
        for (int k = 0; k < 3; ++k) {
            d.lmn[k] = lb.mn[k];
            d.lmx[k] = lb.mx[k];
            d.rmn[k] = rb.mn[k];
            d.rmx[k] = rb.mx[k];
        }
    }
    return d;
}

// The decision of every task of the level, one group of 16 lanes per task
// (grid stride: decisions are independent of each other).
__global__ void __launch_bounds__(kBlock)
    decide_kernel(const BLevel* levels, uint32_t level, const BTask* tasks, BinAcc* bins,
                  BinAcc* hot, uint32_t max_leaf, const uint32_t* idx, const float* centers,
                  uint32_t stride, const DevAabb* aabbs, SDec* decs) {
    __shared__ uint4 s_gather[kBlock / kBins][2 * kSmallTask];  // per group
    const uint32_t num = levels[level].num_tasks;
    const uint32_t group = threadIdx.x / kBins;
    const int lane = int(threadIdx.x % kBins);
    // whole groups: kBlock is a multiple of kBins
    for (uint32_t t = blockIdx.x * (kBlock / kBins) + group; t < num;
         t += gridDim.x * (kBlock / kBins)) {
        const SDec d = decide(tasks[t], lane, bins, hot, max_leaf, idx, centers, stride, aabbs,
                              s_gather[group]);
        if (lane == 0) decs[t] = d;
    }
}

// Per tile of kTaskTile tasks, one per thread: the prefix of splitting tasks
// and of binned children over the tiles before (payload: splits in bits
// 61:31, binned children in bits 30:0), each task's node record, its
// children as tasks of the next level and what the partition needs; the last
// tile writes the next level's record.  A block takes one tile when the grid
// covers the tiles, else tiles until none is left.
__global__ void __launch_bounds__(kTaskTile)
    children_kernel(BLevel* levels, uint32_t level, const BTask* tasks, const SDec* decs,
                    uint64_t* tile_words, DevAabb* nodes, BTask* next, BPart* parts) {
    constexpr uint32_t kWaves = kTaskTile / 64;
    __shared__ uint32_t s_wsplit[kWaves], s_wslot[kWaves];
    __shared__ uint32_t s_tile, s_split_excl, s_slot_excl;
    const BLevel lv = levels[level];
    const uint32_t num = lv.num_tasks;
    if (num == 0) {
        // past the last level: carry the node count forward
        if (blockIdx.x == 0 && threadIdx.x == 0) levels[level + 1].node_base = lv.node_base;
        return;
    }
    const uint32_t tiles = (num + kTaskTile - 1) / kTaskTile;
    if (blockIdx.x >= tiles) return;  // no ticket taken: nothing would be left for it
    // this level's words: after every earlier level's, as levels occupy
    // consecutive node ids
    uint64_t* words = tile_words + lv.node_base / kTaskTile + level;
    const uint32_t wlane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (;;) {
        const uint32_t tile = next_tile(&levels[level].split_ticket, &s_tile);
        if (tile >= tiles) break;
        const uint32_t t = tile * kTaskTile + threadIdx.x;
        const bool valid = t < num;
        BTask tk;
        SDec d;
        d.split = 0;
        bool lbin = false, rbin = false;
        if (valid) {
            tk = tasks[t];
            d = decs[t];
            if (d.split) {
                lbin = binned_by_atomics(d.left_count, tk.depth + 1);
                rbin = binned_by_atomics(tk.end - tk.begin - d.left_count, tk.depth + 1);
            }
        }
        const bool sp = valid && d.split;
        const uint64_t ms = __ballot(sp), ml = __ballot(lbin), mr = __ballot(rbin);
        const uint64_t below = (1ull << wlane) - 1;
        uint32_t below_split = uint32_t(__popcll(ms & below));
        uint32_t below_slots = uint32_t(__popcll(ml & below) + __popcll(mr & below));
        if (wlane == 0) {
            s_wsplit[wave] = uint32_t(__popcll(ms));
            s_wslot[wave] = uint32_t(__popcll(ml) + __popcll(mr));
        }
        __syncthreads();
        if (wave == 0) {
            uint32_t ts = 0, tl = 0;
            for (uint32_t w = 0; w < kWaves; ++w) {
                ts += s_wsplit[w];
                tl += s_wslot[w];
            }
            const uint64_t agg = (uint64_t(ts) << 31) | uint64_t(tl);
            uint64_t excl = 0;
            if (tile == 0) {
                if (wlane == 0) publish(&words[0], kInclusive | agg);
            } else {
                if (wlane == 0) publish(&words[tile], kAggregate | agg);
                excl = lookback(words, tile, wlane);
                if (wlane == 0) publish(&words[tile], kInclusive | (excl + agg));
            }
            if (wlane == 0) {
                s_split_excl = uint32_t(excl >> 31);
                s_slot_excl = uint32_t(excl & 0x7FFFFFFFu);
                if (tile == tiles - 1) {
                    const uint64_t total = excl + agg;
                    BLevel* nx = &levels[level + 1];
                    nx->num_tasks = 2 * uint32_t(total >> 31);
                    nx->node_base = lv.node_base + num;
                    nx->slots = uint32_t(total & 0x7FFFFFFFu);
                }
            }
        }
        __syncthreads();
        if (valid) {
            for (uint32_t w = 0; w < wave; ++w) {
                below_split += s_wsplit[w];
                below_slots += s_wslot[w];
            }
            const uint32_t work = tk.end - tk.begin;
            DevAabb nd;
            for (int k = 0; k < 3; ++k) {
                nd.min[k] = tk.mn[k];
                nd.max[k] = tk.mx[k];
            }
            BPart pt;
            pt.begin = tk.begin;
            pt.left_count = 0;
            pt.axis = kLeaf;
            pt.split_index = 0;
            pt.ctb = 0.0f;
            pt.bo = 0.0f;
            pt.child = 0;
            pt.pad = 0;
            if (d.split) {
                const uint32_t child = 2 * (s_split_excl + below_split);  // next-level task
                const uint32_t slot = s_slot_excl + below_slots;
                nd.extra1 = -1;
                nd.extra2 = int32_t(lv.node_base + num + child);
                next[child] = make_task(d.lmn, d.lmx, tk.begin, tk.begin + d.left_count,
                                        tk.depth + 1, lbin ? slot : kNoSlot);
                next[child + 1] = make_task(d.rmn, d.rmx, tk.begin + d.left_count, tk.end,
                                            tk.depth + 1, rbin ? slot + (lbin ? 1u : 0u) : kNoSlot);
                pt.left_count = d.left_count;
                pt.axis = d.axis;
                pt.split_index = d.split_index;
                pt.ctb = pick3(tk.ctb, d.axis);
                pt.bo = pick3(tk.bo, d.axis);
                pt.child = child;
            } else {
                nd.extra1 = int32_t(work);
                nd.extra2 = int32_t(tk.begin);
            }
            nodes[lv.node_base + t] = nd;
            parts[t] = pt;
        }
        if (gridDim.x >= tiles) break;  // one tile per block
        __syncthreads();  // s_w* and s_tile are rewritten for the next tile
    }
}

// Stable partition of every splitting task's range, one tile of kPrimTile
// primitives per block (kChunks runs of kPartBlock; lane order = index
// order).  A left primitive goes to begin + (lefts before it in the task), a
// right one to begin + left_count + (rights before it).  Lefts before come
// from the tile's own ranks and, for the one task that may begin in an
// earlier tile, a carry over the tiles before: a tile's word holds the lefts
// of its last task from the task's first primitive in the tile (aggregate)
// or in the index array (inclusive); a tile whose last task begins inside
// it, or whose last primitive is no splitting task's, publishes inclusive at
// once.
__global__ void __launch_bounds__(kPartBlock)
    partition_kernel(BLevel* levels, uint32_t level, uint32_t n, int32_t* ptask,
                     const uint32_t* idx, const BPart* parts, const float* centers,
                     uint32_t stride, uint64_t* tile_words, uint32_t* idx_out,
                     int32_t* ptask_out) {
    constexpr uint32_t kWaves = kPartBlock / 64;
    constexpr uint32_t kRuns = kChunks * kWaves;  // (chunk, wave) runs of 64, in index order
    static_assert(kRuns == 64, "one lane per run");
    __shared__ uint32_t s_rank[kPrimTile];  // lefts before each primitive in the tile
    __shared__ uint32_t s_run[kRuns];
    __shared__ uint32_t s_tile, s_carry, s_total;
    __shared__ int32_t s_last;
    __shared__ uint32_t s_last_begin, s_last_split;
    if (levels[level].num_tasks == 0) return;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t tiles = (n + kPrimTile - 1) / kPrimTile;
    uint64_t* words = tile_words + size_t(level) * tiles;
    const uint32_t tile = next_tile(&levels[level].part_ticket, &s_tile);  // grid = tiles
    const uint32_t base = tile * kPrimTile;
    const uint32_t last = (base + kPrimTile < n ? base + kPrimTile : n) - 1;
    int32_t tt[kChunks];
    uint32_t pp[kChunks];
    bool ff[kChunks];
    BPart pb[kChunks];
    for (uint32_t c = 0; c < kChunks; ++c) {
        const uint32_t i = base + c * kPartBlock + threadIdx.x;
        const int32_t t = i < n ? ptask[i] : -1;
        uint32_t p = 0;
        bool f = false;
        BPart pt;
        pt.axis = kLeaf;
        pt.begin = 0;
        if (t >= 0) {
            p = idx[i];
            pt = parts[t];
            if (pt.axis != kLeaf)
                f = bin_index(centers[size_t(p) * stride + pt.axis], pt.ctb, pt.bo) <
                    pt.split_index;
        }
        if (i == last) {
            s_last = t;
            s_last_begin = pt.begin;
            s_last_split = pt.axis != kLeaf ? 1u : 0u;
        }
        tt[c] = t;
        pp[c] = p;
        ff[c] = f;
        pb[c] = pt;
        const uint64_t m = __ballot(f);
        if (lane == 0) s_run[c * kWaves + wave] = uint32_t(__popcll(m));
        s_rank[c * kPartBlock + threadIdx.x] = uint32_t(__popcll(m & ((1ull << lane) - 1)));
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive prefix of the 64 run counts, lane k holding run k
        const uint32_t v = s_run[lane];
        uint32_t incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(incl, off, 64);
            if (lane >= uint32_t(off)) incl += u;
        }
        s_run[lane] = incl - v;
        if (lane == 63) s_total = incl;
    }
    __syncthreads();
    for (uint32_t c = 0; c < kChunks; ++c)
        s_rank[c * kPartBlock + threadIdx.x] += s_run[c * kWaves + wave];
    __syncthreads();
    if (wave == 0) {
        // the tile's word: lefts of its last task within the tile
        const uint32_t total = s_total;
        uint64_t word = kInclusive;  // no splitting task continues past the tile
        if (s_last >= 0 && s_last_split) {
            const uint32_t b = s_last_begin;
            word = b >= base ? kInclusive | (total - s_rank[b - base])
                             : kAggregate | total;  // one task covers the tile
        }
        if (lane == 0) publish(&words[tile], word);
        // carry: lefts of the first task in the tiles before (thread 0 holds
        // primitive `base`)
        const bool need = __shfl(tt[0] >= 0 && pb[0].axis != kLeaf && pb[0].begin < base ? 1 : 0,
                                 0, 64) != 0;
        const uint32_t carry = need ? uint32_t(lookback(words, tile, lane)) : 0u;
        if (lane == 0) {
            if ((word >> 62) == 1) publish(&words[tile], kInclusive | (carry + total));
            s_carry = carry;
        }
    }
    __syncthreads();
    for (uint32_t c = 0; c < kChunks; ++c) {
        const int32_t t = tt[c];
        if (t < 0) continue;
        const uint32_t i = base + c * kPartBlock + threadIdx.x;
        const BPart& pt = pb[c];
        if (pt.axis == kLeaf) {
            // final place: both buffers, out of later levels
            idx_out[i] = pp[c];
            ptask_out[i] = -1;
            ptask[i] = -1;
            continue;
        }
        const uint32_t rank = s_rank[c * kPartBlock + threadIdx.x];
        const uint32_t lefts = pt.begin >= base ? rank - s_rank[pt.begin - base] : s_carry + rank;
        const uint32_t pos = ff[c] ? pt.begin + lefts
                                   : pt.begin + pt.left_count + (i - pt.begin - lefts);
        if (pos >= n) continue;  // never taken: ranks stay inside the task's range
        idx_out[pos] = pp[c];
        ptask_out[pos] = int32_t(pt.child + (ff[c] ? 0u : 1u));
    }
}

__global__ void iota_kernel(uint32_t* idx, int32_t* ptask, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        idx[i] = i;
        ptask[i] = 0;
    }
}

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
    out[i] = a;
}

// Root task: Aabb::union_of_list (binned_sah.rs:361), sequential fold order
// per block then across blocks.
// Root box in two passes: kRootParts blocks read the boxes coalesced
// (grid stride) and leave one partial box each; one block reduces those.
// The minimum and maximum ignore NaN as Rust's f32::min/max do and do not
// depend on the order, up to the sign of a zero bound, which the 0.0001
// padding below removes.
constexpr uint32_t kRootParts = 512;

__device__ __forceinline__ void block_box(float (&acc)[6], float (*sh)[6]) {
    for (int k = 0; k < 6; ++k) sh[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (uint32_t h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 3; ++k) {
                sh[threadIdx.x][k] = rs_min(sh[threadIdx.x][k], sh[threadIdx.x + h][k]);
                sh[threadIdx.x][3 + k] = rs_max(sh[threadIdx.x][3 + k], sh[threadIdx.x + h][3 + k]);
            }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock)
    root_partial_kernel(const DevAabb* aabbs, uint32_t n, float* parts) {
    __shared__ float sh[kBlock][6];
    float acc[6] = {1e34f, 1e34f, 1e34f, -1e34f, -1e34f, -1e34f};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRootParts * kBlock) {
        const DevAabb a = aabbs[i];
        for (int k = 0; k < 3; ++k) {
            acc[k] = rs_min(acc[k], a.min[k]);
            acc[3 + k] = rs_max(acc[3 + k], a.max[k]);
        }
    }
    block_box(acc, sh);
    if (threadIdx.x < 6) parts[size_t(blockIdx.x) * 6 + threadIdx.x] = sh[0][threadIdx.x];
}

// The root task (its box padded by 1e-4 before the node padding) and the
// first level's record.
__global__ void __launch_bounds__(kBlock)
    root_bounds_kernel(const float* parts, uint32_t n, BTask* task, BLevel* levels) {
    __shared__ float sh[kBlock][6];
    float acc[6] = {1e34f, 1e34f, 1e34f, -1e34f, -1e34f, -1e34f};
    for (uint32_t p = threadIdx.x; p < kRootParts; p += kBlock)
        for (int k = 0; k < 3; ++k) {
            acc[k] = rs_min(acc[k], parts[p * 6 + k]);
            acc[3 + k] = rs_max(acc[3 + k], parts[p * 6 + 3 + k]);
        }
    block_box(acc, sh);
    if (threadIdx.x == 0) {
        float mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            mn[k] = sh[0][k] - 0.0001f;
            mx[k] = sh[0][3 + k] + 0.0001f;
        }
        const bool binned = binned_by_atomics(n, 0);
        *task = make_task(mn, mx, 0, n, 0, binned ? 0u : kNoSlot);
        BLevel lv;
        lv.num_tasks = 1;
        lv.node_base = 0;
        lv.slots = binned ? 1u : 0u;
        lv.split_ticket = 0;
        lv.part_ticket = 0;
        lv.pad[0] = lv.pad[1] = lv.pad[2] = 0;
        levels[0] = lv;
    }
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

ResultCode gpu_build_binned(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                            size_t n64, size_t max_leaf, hipStream_t stream, DeviceBvh* out) {
    if (n64 == 0) return NoPrimitives;
    if (n64 > (1u << 30)) {
        set_error("prim_count %zu too large", n64);
        return InvalidArgument;
    }
    const uint32_t n = uint32_t(n64);
    const uint32_t stride = uint32_t(stride_floats);
    const uint32_t leaf =
        max_leaf == 0 ? 1u : uint32_t(max_leaf > 0xFFFFFFFFu ? 0xFFFFFFFFu : max_leaf);
    hipError_t e;
    DeviceBvh res;
    const uint32_t cap_nodes = 2 * n - 1;
    if ((e = hipMalloc(&res.nodes, size_t(cap_nodes) * sizeof(DevAabb))) != hipSuccess)
        return hip_fail(e, "hipMalloc(nodes)");
    if ((e = hipMalloc(&res.indices, size_t(n) * 4)) != hipSuccess) {
        (void)hipFree(res.nodes);
        return hip_fail(e, "hipMalloc(indices)");
    }
    res.prim_count = n;
    auto fail = [&](hipError_t err, const char* what) {
        (void)hipFree(res.nodes);
        (void)hipFree(res.indices);
        return hip_fail(err, what);
    };
    // One workspace.  A level holds at most n tasks (disjoint non-empty
    // ranges) and at most n / 65 of them have bins; the children tile words
    // of all levels fit (2n - 1) / kTaskTile + levels + 1 (levels use
    // consecutive node ids), the partition words one row per level.
    const uint32_t ptiles = (n + kPrimTile - 1) / kPrimTile;
    const size_t slot_cap = size_t(n) / (kSmallTask + 1) + 1;
    const size_t split_words = size_t(2) * n / kTaskTile + kMaxLevels + 2;
    const size_t sz_pts = aabbs ? 0 : align256(size_t(n) * sizeof(DevAabb));
    const size_t sz_idx = align256(size_t(n) * 4);
    const size_t sz_tasks = align256(size_t(n) * sizeof(BTask));
    const size_t sz_parts = align256(size_t(n) * sizeof(BPart));
    const size_t sz_decs = align256(size_t(n) * sizeof(SDec));
    const size_t sz_root = align256(size_t(kRootParts) * 6 * 4);
    // cleared together: level records, tile words, bins
    const size_t sz_levels = align256(size_t(kMaxLevels + 1) * sizeof(BLevel));
    const size_t sz_words = align256((split_words + size_t(kMaxLevels) * ptiles) * 8);
    const size_t sz_bins = align256(slot_cap * 3 * kBins * sizeof(BinAcc));
    const size_t sz_hot = align256(size_t(kHotSlots) * kCopies * 3 * kBins * sizeof(BinAcc));
    const size_t total = sz_pts + 3 * sz_idx + 2 * sz_tasks + sz_parts + sz_decs + sz_root +
                         sz_levels + sz_words + sz_bins + sz_hot;
    Workspace wsh;
    if ((e = workspace_acquire(kWsBinned, total, &wsh)) != hipSuccess)
        return fail(e, "binned workspace");
    // released on every return, after the stream's work that used it
    struct WsRelease {
        Workspace* w;
        hipStream_t s;
        ~WsRelease() { workspace_release(w, s, true); }
    } wsr{&wsh, stream};
    char* ws = static_cast<char*>(wsh.ptr);
    BLevel* host_level = nullptr;
    if ((e = hipHostMalloc(&host_level, sizeof(BLevel))) != hipSuccess)
        return fail(e, "hipHostMalloc");
    struct HostFree {
        BLevel* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{host_level};
    char* cur = ws;
    auto carve = [&](size_t bytes) {
        char* p = cur;
        cur += bytes;
        return p;
    };
    DevAabb* pts = aabbs ? nullptr : reinterpret_cast<DevAabb*>(carve(sz_pts));
    uint32_t* idx1 = reinterpret_cast<uint32_t*>(carve(sz_idx));
    int32_t* pt0 = reinterpret_cast<int32_t*>(carve(sz_idx));
    int32_t* pt1 = reinterpret_cast<int32_t*>(carve(sz_idx));
    BTask* tasks0 = reinterpret_cast<BTask*>(carve(sz_tasks));
    BTask* tasks1 = reinterpret_cast<BTask*>(carve(sz_tasks));
    BPart* parts = reinterpret_cast<BPart*>(carve(sz_parts));
    SDec* decs = reinterpret_cast<SDec*>(carve(sz_decs));
    float* root_parts = reinterpret_cast<float*>(carve(sz_root));
    char* cleared = cur;
    BLevel* levels = reinterpret_cast<BLevel*>(carve(sz_levels));
    uint64_t* split_w = reinterpret_cast<uint64_t*>(carve(sz_words));
    uint64_t* part_w = split_w + split_words;
    BinAcc* bins = reinterpret_cast<BinAcc*>(carve(sz_bins));
    BinAcc* hot = reinterpret_cast<BinAcc*>(carve(sz_hot));
    if ((e = hipMemsetAsync(cleared, 0, sz_levels + sz_words + sz_bins + sz_hot, stream)) !=
        hipSuccess)
        return fail(e, "binned clear");

    const DevAabb* in_aabbs = reinterpret_cast<const DevAabb*>(aabbs);
    const uint32_t g = (n + kBlock - 1) / kBlock;
    if (!aabbs) {
        hipLaunchKernelGGL(point_aabbs_kernel, dim3(g), dim3(kBlock), 0, stream, centers, stride,
                           n, pts);
        in_aabbs = pts;
    }
    uint32_t* idx[2] = {res.indices, idx1};
    int32_t* ptk[2] = {pt0, pt1};
    BTask* tasks[2] = {tasks0, tasks1};
    hipLaunchKernelGGL(iota_kernel, dim3(g), dim3(kBlock), 0, stream, idx[0], ptk[0], n);
    hipLaunchKernelGGL(root_partial_kernel, dim3(kRootParts), dim3(kBlock), 0, stream, in_aabbs, n,
                       root_parts);
    hipLaunchKernelGGL(root_bounds_kernel, dim3(1), dim3(kBlock), 0, stream, root_parts, n,
                       tasks[0], levels);
    const uint32_t child_cap = std::min<uint32_t>((n + kTaskTile - 1) / kTaskTile, kChildGrid);
    const uint32_t decide_cap = std::min<uint32_t>((n + kBins - 1) / kBins, kDecideGrid);
    // bin_kernel: runs of whole chunks, about kBinGrid of them
    const uint32_t run = ((n + kBinGrid - 1) / kBinGrid + kBlock - 1) / kBlock * kBlock;
    const uint32_t bgrid = (n + run - 1) / run;
    // The first batch covers a tree twice as deep as a balanced one (levels
    // past the last cost four launches that return at once); then batches
    // of 8 until a level has no task.
    uint32_t lg = 0;
    while (lg < 31 && (1u << lg) < n) ++lg;
    uint32_t batch = std::min<uint32_t>(kMaxLevels, 2 * lg + 2);
    uint32_t level = 0;
    uint64_t known = 1;  // tasks at `level`
    for (;;) {
        const uint32_t end = std::min(kMaxLevels, level + batch);
        for (uint32_t k = level; k < end; ++k) {
            // a level holds at most twice the tasks of the one before
            const uint32_t sh = std::min<uint32_t>(k - level, 32);
            const uint64_t bound = std::min<uint64_t>(n, known << sh);
            const uint32_t cgrid = uint32_t(
                std::min<uint64_t>(child_cap, (bound + kTaskTile - 1) / kTaskTile));
            const uint32_t dgrid = uint32_t(std::min<uint64_t>(
                decide_cap, (bound + kBlock / kBins - 1) / (kBlock / kBins)));
            const uint32_t a = k & 1u, b = a ^ 1u;
            hipLaunchKernelGGL(bin_kernel, dim3(bgrid), dim3(kBlock), 0, stream, levels, k, n, run,
                               ptk[a], idx[a], tasks[a], centers, stride, in_aabbs, bins, hot);
            hipLaunchKernelGGL(decide_kernel, dim3(dgrid), dim3(kBlock), 0, stream, levels, k,
                               tasks[a], bins, hot, leaf, idx[a], centers, stride, in_aabbs,
                               decs);
            hipLaunchKernelGGL(children_kernel, dim3(cgrid), dim3(kTaskTile), 0, stream, levels, k,
                               tasks[a], decs, split_w, res.nodes, tasks[b], parts);
            // one tile per block: the grid is the tile count
            hipLaunchKernelGGL(partition_kernel, dim3(ptiles), dim3(kPartBlock), 0, stream, levels,
                               k, n, ptk[a], idx[a], parts, centers, stride, part_w, idx[b],
                               ptk[b]);
        }
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "binned level");
        level = end;
        if ((e = hipMemcpyAsync(host_level, levels + level, sizeof(BLevel), hipMemcpyDeviceToHost,
                                stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
            return fail(e, "binned levels");
        if (host_level->num_tasks == 0) break;
        if (level >= kMaxLevels) {  // a task at depth 64 is a leaf
            set_error("binned SAH: tasks left past depth %u", kMaxDepth);
            (void)hipFree(res.nodes);
            (void)hipFree(res.indices);
            return Error;
        }
        known = host_level->num_tasks;
        batch = 8;
    }
    // every primitive was written to both index buffers when its leaf was made
    res.node_count = host_level->node_base;
    *out = res;
    return Ok;
}

}  // namespace rtbvh
