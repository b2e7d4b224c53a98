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
#include <algorithm>

#include "internal.hpp"

namespace rtbvh {

hipError_t scan_exclusive_u32(uint32_t* data, uint32_t count, uint32_t* total, void* scratch,
                              size_t scratch_bytes, hipStream_t stream);
size_t scan_scratch_bytes(uint32_t count);

namespace {

constexpr int kBins = 16;          // binned_sah.rs:314
constexpr uint32_t kMaxDepth = 64;  // binned_sah.rs:313
constexpr int kBlock = 256;
// Tasks of at most kSmallTask primitives are binned by their split group
// (one lane per primitive) instead of by bin_kernel's atomics.
constexpr uint32_t kSmallTask = 16;

struct BTask {
    float mn[3];
    uint32_t node;
    float mx[3];
    uint32_t begin;
    uint32_t end;
    uint32_t depth;
    uint32_t pad[2];
};

struct BinAcc {
    uint32_t cnt;
    uint32_t mn[3];  // order-preserving encodings of floats
    uint32_t mx[3];
    uint32_t pad;
};

struct BDecision {
    float mn[3];      // padded node box
    uint32_t split;   // 1 = split, 0 = leaf
    float mx[3];
    uint32_t axis;
    float ctb[3];
    uint32_t split_index;
    float bo[3];
    uint32_t left_count;
    float lmn[3], lmx[3], rmn[3], rmx[3];
};

__device__ __forceinline__ uint32_t fenc(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t e) {
    return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
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

// Per task: padded box, bin parameters, cleared bins.
__global__ void level_init_kernel(const BTask* tasks, uint32_t num_tasks, BDecision* dec,
                                  BinAcc* bins) {
    const uint32_t t = blockIdx.x;
    if (t >= num_tasks) return;
    const BTask tk = tasks[t];
    if (threadIdx.x == 0) {
        BDecision d;
        for (int k = 0; k < 3; ++k) {
            d.mn[k] = tk.mn[k] - 0.0001f;  // binned_sah.rs:133
            d.mx[k] = tk.mx[k] + 0.0001f;
            d.ctb[k] = (1.0f / (d.mx[k] - d.mn[k])) * float(kBins);  // :146
            d.bo[k] = (-d.mn[k]) * d.ctb[k];                        // :147
        }
        d.split = 0;
        d.axis = 0;
        d.split_index = 0;
        d.left_count = 0;
        dec[t] = d;
    }
    if (threadIdx.x < 3 * kBins && tk.end - tk.begin > kSmallTask) {
        BinAcc a;
        a.cnt = 0;
        for (int k = 0; k < 3; ++k) {
            a.mn[k] = fenc(1e34f);  // Aabb::empty
            a.mx[k] = fenc(-1e34f);
        }
        a.pad = 0;
        bins[size_t(t) * 3 * kBins + threadIdx.x] = a;
    }
}

// Bin every primitive of every task with more than kSmallTask primitives
// (smaller tasks with two or more are binned in split_kernel).  The live task
// ids along the index array never decrease (the scatter places each split
// task's children, left then right, inside the parent's range, and tasks are
// numbered in range order); finished primitives carry -1 and sit between
// them.  A block whose live primitives span at most kLdsTasks tasks
// accumulates in LDS and merges one set of atomics per task and bin.
constexpr int kLdsTasks = 16;  // 4: 41 ms spatial, 22 ms binned at 1M; 16: 36, 19.5; 32: no better

__global__ void __launch_bounds__(kBlock)
    bin_kernel(uint32_t n, const int32_t* ptask, const uint32_t* idx, const BTask* tasks,
               const BDecision* dec, const float* centers, uint32_t stride,
               const DevAabb* aabbs, BinAcc* bins) {
    constexpr int kSlots = kLdsTasks * 3 * kBins;
    __shared__ uint32_t s_cnt[kSlots];
    __shared__ uint32_t s_mn[kSlots][3];
    __shared__ uint32_t s_mx[kSlots][3];
    __shared__ int32_t s_lo, s_hi;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (threadIdx.x == 0) {
        s_lo = 0x7fffffff;
        s_hi = -1;
    }
    for (uint32_t k = threadIdx.x; k < uint32_t(kSlots); k += kBlock) {
        s_cnt[k] = 0;
        for (int q = 0; q < 3; ++q) {
            s_mn[k][q] = fenc(1e34f);
            s_mx[k][q] = fenc(-1e34f);
        }
    }
    __syncthreads();
    const int32_t t = i < n ? ptask[i] : -1;
    if (t >= 0) {
        atomicMin(&s_lo, t);
        atomicMax(&s_hi, t);
    }
    __syncthreads();
    const int32_t lo = s_lo, hi = s_hi;
    const bool local = hi >= 0 && hi - lo < kLdsTasks;
    if (t >= 0) {
        const BTask tk = tasks[t];
        if (tk.end - tk.begin > kSmallTask && tk.depth < kMaxDepth) {
            const BDecision& d = dec[t];
            const uint32_t p = idx[i];
            const float* c = centers + size_t(p) * stride;
            const DevAabb a = aabbs[p];
            for (int ax = 0; ax < 3; ++ax) {
                const uint32_t b = bin_index(c[ax], d.ctb[ax], d.bo[ax]);
                const uint32_t slot = ax * kBins + b;
                if (local) {
                    const uint32_t ls = uint32_t(t - lo) * 3 * kBins + slot;
                    atomicAdd(&s_cnt[ls], 1u);
                    for (int k = 0; k < 3; ++k) {
                        if (a.min[k] == a.min[k]) atomicMin(&s_mn[ls][k], fenc(a.min[k]));
                        if (a.max[k] == a.max[k]) atomicMax(&s_mx[ls][k], fenc(a.max[k]));
                    }
                } else {
                    BinAcc* g = bins + size_t(t) * 3 * kBins + slot;
                    atomicAdd(&g->cnt, 1u);
                    for (int k = 0; k < 3; ++k) {
                        if (a.min[k] == a.min[k]) atomicMin(&g->mn[k], fenc(a.min[k]));
                        if (a.max[k] == a.max[k]) atomicMax(&g->mx[k], fenc(a.max[k]));
                    }
                }
            }
        }
    }
    __syncthreads();
    if (local)
        for (uint32_t k = threadIdx.x; k < uint32_t(hi - lo + 1) * 3 * kBins; k += kBlock) {
            if (s_cnt[k] == 0) continue;
            BinAcc* g = bins + size_t(lo) * 3 * kBins + k;
            atomicAdd(&g->cnt, s_cnt[k]);
            for (int q = 0; q < 3; ++q) {
                atomicMin(&g->mn[q], s_mn[k][q]);
                atomicMax(&g->mx[q], s_mx[k][q]);
            }
        }
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

__device__ __forceinline__ GBox gbox_shfl_up(const GBox& b, int off) {
    GBox r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = __shfl_up(b.mn[k], off, kBins);
        r.mx[k] = __shfl_up(b.mx[k], off, kBins);
    }
    r.cnt = __shfl_up(b.cnt, off, kBins);
    return r;
}

__device__ __forceinline__ GBox gbox_shfl_down(const GBox& b, int off) {
    GBox r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = __shfl_down(b.mn[k], off, kBins);
        r.mx[k] = __shfl_down(b.mx[k], off, kBins);
    }
    r.cnt = __shfl_down(b.cnt, off, kBins);
    return r;
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

// Per task: the best SAH plane per axis (find_split, binned_sah.rs:80-114),
// the axis, the fallback and the child boxes (:150-230).  One group of 16
// lanes per task, lane i holding bin i: inclusive prefix (bins 0..i) and
// suffix (bins i..15) runs by shuffle scans, the cost of the plane after bin
// i as the one-thread sweep computes it, half_area(left) * left count +
// half_area(right) * right count, and the first lowest cost below FLT_MAX
// by a reduction.  The bins' bounds hold no NaN (the binning skips NaN
// bounds), so every bound equals the sequential fold's, up to the sign of a
// zero in the right runs of the sweep, which no comparison tells apart; the
// child boxes come from increasing-order runs, as the sweep builds them.
__global__ void __launch_bounds__(kBlock)
    split_kernel(const BTask* tasks, uint32_t num_tasks, BDecision* dec, const BinAcc* bins,
                 uint32_t max_leaf, uint32_t* split_flag, const uint32_t* idx,
                 const float* centers, uint32_t stride, const DevAabb* aabbs) {
    const uint32_t t = (blockIdx.x * kBlock + threadIdx.x) / kBins;
    const int lane = int(threadIdx.x % kBins);
    if (t >= num_tasks) return;  // whole groups: kBlock is a multiple of kBins
    const BTask tk = tasks[t];
    BDecision d = dec[t];
    const uint32_t work = tk.end - tk.begin;
    d.split = 0;
    if (work > 1 && tk.depth < kMaxDepth) {
        const BinAcc* tb = bins + size_t(t) * 3 * kBins;
        BinAcc sb[3];  // this lane's bin of each axis, for a small task
        if (work <= kSmallTask) {
            // bin_kernel's accumulation without atomics: lane q holds
            // primitive q, lane j gathers bin j; the minimum and maximum of
            // the same ordered encodings, NaN bounds skipped
            uint32_t pb[3] = {0, 0, 0}, pmn[3], pmx[3];
            for (int k = 0; k < 3; ++k) {
                pmn[k] = fenc(1e34f);
                pmx[k] = fenc(-1e34f);
            }
            if (uint32_t(lane) < work) {
                const uint32_t p = idx[tk.begin + lane];
                const float* c = centers + size_t(p) * stride;
                const DevAabb a = aabbs[p];
                for (int k = 0; k < 3; ++k) {
                    pb[k] = bin_index(c[k], d.ctb[k], d.bo[k]);
                    if (a.min[k] == a.min[k]) pmn[k] = fenc(a.min[k]);
                    if (a.max[k] == a.max[k]) pmx[k] = fenc(a.max[k]);
                }
            }
            for (int ax = 0; ax < 3; ++ax) {
                sb[ax].cnt = 0;
                for (int k = 0; k < 3; ++k) {
                    sb[ax].mn[k] = fenc(1e34f);
                    sb[ax].mx[k] = fenc(-1e34f);
                }
            }
            for (uint32_t q = 0; q < work; ++q) {
                uint32_t qb[3], qmn[3], qmx[3];
                for (int k = 0; k < 3; ++k) {
                    qb[k] = __shfl(pb[k], int(q), kBins);
                    qmn[k] = __shfl(pmn[k], int(q), kBins);
                    qmx[k] = __shfl(pmx[k], int(q), kBins);
                }
                for (int ax = 0; ax < 3; ++ax)
                    if (qb[ax] == uint32_t(lane)) {
                        ++sb[ax].cnt;
                        for (int k = 0; k < 3; ++k) {
                            sb[ax].mn[k] = min(sb[ax].mn[k], qmn[k]);
                            sb[ax].mx[k] = max(sb[ax].mx[k], qmx[k]);
                        }
                    }
            }
        }
        GBox pre[3], suf[3];
        float cost[3];
        uint32_t cnt[3];
        for (int ax = 0; ax < 3; ++ax) {
            const BinAcc a = work <= kSmallTask ? sb[ax] : tb[ax * kBins + lane];
            GBox own;
            for (int k = 0; k < 3; ++k) {
                own.mn[k] = fdec(a.mn[k]);
                own.mx[k] = fdec(a.mx[k]);
            }
            own.cnt = a.cnt;
            GBox p = own, q = own;
            for (int off = 1; off < kBins; off <<= 1) {
                GBox lower = gbox_shfl_up(p, off);
                const GBox higher = gbox_shfl_down(q, off);
                if (lane >= off) {
                    gbox_join(lower, p);
                    p = lower;
                }
                if (lane + off < kBins) gbox_join(q, higher);
            }
            pre[ax] = p;
            suf[ax] = q;
            const GBox right = gbox_shfl_down(q, 1);
            float key = __builtin_inff();
            if (lane < kBins - 1) {
                const float c = gbox_half_area(p) * float(p.cnt) +
                                gbox_half_area(right) * float(right.cnt);
                if (c < 3.40282347e38f) key = c;
            }
            int idx = lane;
            for (int off = kBins / 2; off > 0; off >>= 1) {
                const float ok = __shfl_xor(key, off, kBins);
                const int oi = __shfl_xor(idx, off, kBins);
                if (ok < key || (ok == key && oi < idx)) {
                    key = ok;
                    idx = oi;
                }
            }
            if (key == __builtin_inff()) {
                cost[ax] = 3.40282347e38f;
                cnt[ax] = kBins;
            } else {
                cost[ax] = key;
                cnt[ax] = uint32_t(idx + 1);
            }
        }
        uint32_t axis = 0;
        if (cost[0] > cost[1]) axis = 1;
        if (cost[axis] > cost[2]) axis = 2;
        uint32_t split_index = cnt[axis];
        const float max_split_cost = half_area6(d.mn, d.mx) * (float(work) - 1.0f);
        bool make_leaf = false;
        if (cnt[axis] == uint32_t(kBins) || cost[axis] >= max_split_cost) {
            if (work > max_leaf) {
                // Aabb::longest_axis of the padded box, aabb.rs:354-363
                axis = 0;
                if (d.mx[1] - d.mn[1] > d.mx[0] - d.mn[0]) axis = 1;
                if (d.mx[2] - d.mn[2] > d.mx[axis] - d.mn[axis]) axis = 2;
                // the first bin whose prefix count reaches 2/5 of the work
                const bool reach =
                    lane < kBins - 1 && uint64_t(pre[axis].cnt) >= uint64_t(work * 2 / 5 + 1);
                const uint32_t base = threadIdx.x & 63u & ~uint32_t(kBins - 1);  // group's lane 0
                const uint32_t m = uint32_t(__ballot(reach) >> base) & 0xFFFFu;
                if (m) split_index = uint32_t(__builtin_ctz(m) + 1);
            } else {
                make_leaf = true;
            }
        }
        if (!make_leaf) {
            const uint32_t lim = split_index < uint32_t(kBins) ? split_index : uint32_t(kBins);
            const GBox lb = gbox_shfl(pre[axis], lim > 0 ? int(lim) - 1 : 0);
            GBox rb = gbox_shfl(suf[axis], lim < uint32_t(kBins) ? int(lim) : 0);
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
        }
    }
    if (lane == 0) {
        dec[t] = d;
        split_flag[t] = d.split;
    }
}

// Writes each task's node (leaf or interior) and the next level's tasks.
__global__ void children_kernel(const BTask* tasks, uint32_t num_tasks, const BDecision* dec,
                                const uint32_t* split_excl, uint32_t counter, DevAabb* nodes,
                                BTask* next) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tasks) return;
    const BTask tk = tasks[t];
    const BDecision& d = dec[t];
    DevAabb nd;
    for (int k = 0; k < 3; ++k) {
        nd.min[k] = d.mn[k];
        nd.max[k] = d.mx[k];
    }
    if (d.split) {
        const uint32_t child = counter + 2 * split_excl[t];
        nd.extra1 = -1;
        nd.extra2 = int32_t(child);
        BTask l, r;
        for (int k = 0; k < 3; ++k) {
            l.mn[k] = d.lmn[k];
            l.mx[k] = d.lmx[k];
            r.mn[k] = d.rmn[k];
            r.mx[k] = d.rmx[k];
        }
        l.node = child;
        r.node = child + 1;
        l.begin = tk.begin;
        l.end = tk.begin + d.left_count;
        r.begin = l.end;
        r.end = tk.end;
        l.depth = r.depth = tk.depth + 1;
        l.pad[0] = l.pad[1] = r.pad[0] = r.pad[1] = 0;
        next[2 * split_excl[t]] = l;
        next[2 * split_excl[t] + 1] = r;
    } else {
        nd.extra1 = int32_t(tk.end - tk.begin);
        nd.extra2 = int32_t(tk.begin);
    }
    nodes[tk.node] = nd;
}

// Left-side flags, written twice: kept in flag, scanned in place in lscan.
__global__ void flag_kernel(uint32_t n, const int32_t* ptask, const uint32_t* idx,
                            const BDecision* dec, const float* centers, uint32_t stride,
                            uint32_t* flag, uint32_t* lscan) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t t = ptask[i];
    uint32_t f = 0;
    if (t >= 0) {
        const BDecision& d = dec[t];
        if (d.split) {
            const float c = centers[size_t(idx[i]) * stride + d.axis];
            f = bin_index(c, d.ctb[d.axis], d.bo[d.axis]) < d.split_index ? 1u : 0u;
        }
    }
    flag[i] = f;
    lscan[i] = f;
}

__global__ void scatter_kernel(uint32_t n, const int32_t* ptask, const uint32_t* idx,
                               const BTask* tasks, const BDecision* dec,
                               const uint32_t* split_excl, const uint32_t* flag,
                               const uint32_t* lscan, uint32_t* idx_out, int32_t* ptask_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t t = ptask[i];
    if (t < 0 || !dec[t].split) {
        idx_out[i] = idx[i];
        ptask_out[i] = -1;
        return;
    }
    const BTask tk = tasks[t];
    const uint32_t lefts_before = lscan[i] - lscan[tk.begin];
    uint32_t pos;
    int32_t nt;
    if (flag[i]) {
        pos = tk.begin + lefts_before;
        nt = int32_t(2 * split_excl[t]);
    } else {
        pos = tk.begin + dec[t].left_count + (i - tk.begin - lefts_before);
        nt = int32_t(2 * split_excl[t] + 1);
    }
    idx_out[pos] = idx[i];
    ptask_out[pos] = nt;
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

__global__ void __launch_bounds__(kBlock)
    root_bounds_kernel(const float* parts, uint32_t n, BTask* task) {
    __shared__ float sh[kBlock][6];
    float acc[6] = {1e34f, 1e34f, 1e34f, -1e34f, -1e34f, -1e34f};
    for (uint32_t p = threadIdx.x; p < kRootParts; p += kBlock)
        for (int k = 0; k < 3; ++k) {
            acc[k] = rs_min(acc[k], parts[p * 6 + k]);
            acc[3 + k] = rs_max(acc[3 + k], parts[p * 6 + 3 + k]);
        }
    block_box(acc, sh);
    if (threadIdx.x == 0) {
        BTask tk;
        for (int k = 0; k < 3; ++k) {
            tk.mn[k] = sh[0][k] - 0.0001f;
            tk.mx[k] = sh[0][3 + k] + 0.0001f;
        }
        tk.node = 0;
        tk.begin = 0;
        tk.end = n;
        tk.depth = 0;
        tk.pad[0] = tk.pad[1] = 0;
        *task = tk;
    }
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t ensure(size_t count) {
        if (count <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, count * sizeof(T) + 16);
        if (e == hipSuccess) cap = count;
        return e;
    }
};

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
    const uint32_t leaf = max_leaf == 0 ? 1u : uint32_t(max_leaf > 0xFFFFFFFFu ? 0xFFFFFFFFu : max_leaf);
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
    DevBuf<DevAabb> pts;
    const DevAabb* in_aabbs = reinterpret_cast<const DevAabb*>(aabbs);
    const uint32_t g = (n + kBlock - 1) / kBlock;
    if (!aabbs) {
        if ((e = pts.ensure(n)) != hipSuccess) return fail(e, "hipMalloc(aabbs)");
        hipLaunchKernelGGL(point_aabbs_kernel, dim3(g), dim3(kBlock), 0, stream, centers, stride,
                           n, pts.p);
        in_aabbs = pts.p;
    }
    DevBuf<uint32_t> idx1, flag, lscan, split_flag, scan_scratch;
    DevBuf<int32_t> pt0, pt1;
    DevBuf<BTask> tasks0, tasks1;
    DevBuf<BDecision> dec;
    DevBuf<BinAcc> bins;
    DevBuf<float> root_parts;
    uint32_t* host_total = nullptr;
    if ((e = hipHostMalloc(&host_total, 8)) != hipSuccess) return fail(e, "hipHostMalloc");
    struct HostFree {
        uint32_t* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{host_total};
    if ((e = idx1.ensure(n)) != hipSuccess || (e = flag.ensure(n)) != hipSuccess ||
        (e = lscan.ensure(n)) != hipSuccess || (e = pt0.ensure(n)) != hipSuccess ||
        (e = pt1.ensure(n)) != hipSuccess || (e = tasks0.ensure(1)) != hipSuccess ||
        (e = scan_scratch.ensure(scan_scratch_bytes(n) / 4 + 1)) != hipSuccess ||
        (e = root_parts.ensure(size_t(kRootParts) * 6)) != hipSuccess)
        return fail(e, "binned workspace");
    uint32_t* idx_cur = res.indices;
    uint32_t* idx_nxt = idx1.p;
    int32_t* pt_cur = pt0.p;
    int32_t* pt_nxt = pt1.p;
    hipLaunchKernelGGL(iota_kernel, dim3(g), dim3(kBlock), 0, stream, idx_cur, pt_cur, n);
    hipLaunchKernelGGL(root_partial_kernel, dim3(kRootParts), dim3(kBlock), 0, stream, in_aabbs, n,
                       root_parts.p);
    hipLaunchKernelGGL(root_bounds_kernel, dim3(1), dim3(kBlock), 0, stream, root_parts.p, n,
                       tasks0.p);
    BTask* t_cur = tasks0.p;
    uint32_t num_tasks = 1;
    uint32_t counter = 1;  // AtomicNodeStack starts at 1 (builders/mod.rs:51-56)
    int flip = 0;
    // Per-task buffers hold tcap tasks: at least 65536 (at most n, as every
    // task has a primitive), doubled when a level needs more, so the levels
    // near the root, whose task count doubles, do not free and allocate
    // every buffer each time (a hipFree and a hipMalloc cost more than the
    // level's kernels there).
    size_t tcap = 0;
    while (num_tasks > 0) {
        if (num_tasks > tcap)
            tcap = std::max<size_t>(
                num_tasks, std::min<size_t>(n, std::max<size_t>(65536, 2 * size_t(num_tasks))));
        if ((e = dec.ensure(tcap)) != hipSuccess ||
            (e = bins.ensure(tcap * 3 * kBins)) != hipSuccess ||
            (e = split_flag.ensure(tcap)) != hipSuccess)
            return fail(e, "binned level workspace");
        DevBuf<BTask>& nxt_buf = flip ? tasks0 : tasks1;
        if ((e = nxt_buf.ensure(2 * tcap)) != hipSuccess)
            return fail(e, "binned tasks");
        BTask* t_nxt = nxt_buf.p;
        const uint32_t tg = (num_tasks + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(level_init_kernel, dim3(num_tasks), dim3(64), 0, stream, t_cur,
                           num_tasks, dec.p, bins.p);
        hipLaunchKernelGGL(bin_kernel, dim3(g), dim3(kBlock), 0, stream, n, pt_cur, idx_cur, t_cur,
                           dec.p, centers, stride, in_aabbs, bins.p);
        hipLaunchKernelGGL(split_kernel, dim3((num_tasks + kBlock / kBins - 1) / (kBlock / kBins)),
                           dim3(kBlock), 0, stream, t_cur, num_tasks, dec.p, bins.p, leaf,
                           split_flag.p, idx_cur, centers, stride, in_aabbs);
        if ((e = scan_exclusive_u32(split_flag.p, num_tasks, host_total, scan_scratch.p,
                                    scan_scratch.cap * 4, stream)) != hipSuccess)
            return fail(e, "binned task scan");
        hipLaunchKernelGGL(children_kernel, dim3(tg), dim3(kBlock), 0, stream, t_cur, num_tasks,
                           dec.p, split_flag.p, counter, res.nodes, t_nxt);
        hipLaunchKernelGGL(flag_kernel, dim3(g), dim3(kBlock), 0, stream, n, pt_cur, idx_cur, dec.p,
                           centers, stride, flag.p, lscan.p);
        if ((e = scan_exclusive_u32(lscan.p, n, nullptr, scan_scratch.p, scan_scratch.cap * 4,
                                    stream)) != hipSuccess)
            return fail(e, "binned partition scan");
        hipLaunchKernelGGL(scatter_kernel, dim3(g), dim3(kBlock), 0, stream, n, pt_cur, idx_cur,
                           t_cur, dec.p, split_flag.p, flag.p, lscan.p, idx_nxt, pt_nxt);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "binned level");
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e, "binned level");
        const uint32_t splits = *host_total;
        counter += 2 * splits;
        num_tasks = 2 * splits;
        std::swap(idx_cur, idx_nxt);
        std::swap(pt_cur, pt_nxt);
        t_cur = t_nxt;
        flip ^= 1;
    }
    if (idx_cur != res.indices) {
        if ((e = hipMemcpyAsync(res.indices, idx_cur, size_t(n) * 4, hipMemcpyDeviceToDevice,
                                stream)) != hipSuccess)
            return fail(e, "binned indices");
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e, "binned indices");
    }
    res.node_count = counter;
    *out = res;
    return Ok;
}

}  // namespace rtbvh
