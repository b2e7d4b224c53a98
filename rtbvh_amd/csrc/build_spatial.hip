// GPU spatial-split SAH builder (SBVH) for gfx950, breadth-first.
//
// The capability of SpatialSahBuilder (src/builders/spatial_sah.rs:247-1034,
// defaults :940-956: 16 bins, depth 64, leaf size from primitives_per_leaf,
// traversal cost 1, alpha 1e-5, split factor 0.75): a node is split either by
// its references' centroids (object split) or by a plane that cuts the
// triangles straddling it into two references, one per side, each bounded by
// the clipped triangle (spatial split).  The design is this build's own, one
// tree level per step with every node of the level in parallel:
//   node box      union of the node's reference boxes (padded by 1e-4, as
//                 the reference pads a parent, :352)
//   object split  16 centroid bins per axis over the node box, SAH sweep
//                 (the reference sorts and sweeps every position, :318-363)
//   spatial split only where the best object split's children overlap by
//                 more than alpha times the root's surface (:768-775): 16
//                 bins per axis, each reference clipped into every bin it
//                 spans (chopped binning, :500-560; one pass instead of two),
//                 entry and exit counts, SAH sweep over the 15 planes
//   budget        the references may grow by split_factor * n (:1016); a
//                 spatial split is taken when it is cheaper and all of its
//                 duplicates still fit, deciding tasks in node order (the
//                 reference falls back per reference when space runs out)
//   leaf          one reference, depth 64, or no split cheaper than
//                 A * (n - traversal cost) with at most max_leaf references;
//                 otherwise an approximate median object split on the
//                 longest axis, and an index median when every centroid
//                 falls into one bin (:785-810)
// Deliberate differences, each a reference defect: triangles are clipped
// with a + t (b - a) (the reference's split_edge computes a * t * (b - a),
// :91-94, so its clipped boxes can miss the triangle); exit counts go to the
// last bin a reference spans (the reference adds both counts to the first,
// :536-537); the straddle fallback keeps the node's own range (:636-645
// index from 0).  A clipped box is widened on the two other axes by 2^-20
// of the vertices' magnitude (the rounding of the crossings) and kept inside
// the reference's box, so the references of a triangle always cover it and
// every hit is found.
//
// The level loop runs on the device: each level's counts (tasks, references,
// node counter, leaf base, duplicates so far) sit in a 64-byte record that
// level_end_kernel writes from the level before's scan totals, every kernel
// reads its counts there and returns at once past them, and the host
// enqueues batches of levels with grids from bounds, reading one record per
// batch.  A level: node boxes (ref_bounds_kernel, task_setup_kernel),
// object bins and split (obj_bin_kernel, obj_split_kernel), a scan of the
// spatial candidates, their bins and splits (cand_clear_kernel,
// sp_bin_kernel, sp_split_kernel), a scan of the duplicates, the decisions
// (decide_kernel), the side flags (side_kernel), one scan of the task
// counts and side flags, then nodes and child tasks (children_kernel), the
// references' scatter (scatter_kernel, which also clears the next level's
// boxes and bins) and the next record.
#include <algorithm>

#include "internal.hpp"

namespace rtbvh {

namespace {

constexpr int kBins = 16;
constexpr uint32_t kMaxDepth = 64;
constexpr float kSplitFactor = 0.75f;
constexpr float kAlpha = 1e-5f;
constexpr int kBlock = 256;
constexpr uint32_t kMaxCand = 65536;  // spatial candidates evaluated per level
// Tasks of at most kSmallTask references get their object bins from their
// split group (one lane per reference) instead of obj_bin_kernel's atomics.
constexpr uint32_t kSmallTask = 16;

// A reference: the part of triangle `prim` inside [mn, mx]; `task` is the
// index of the current level's task that holds it.
struct SRef {
    float mn[3];
    uint32_t prim;
    float mx[3];
    int32_t task;
};
static_assert(sizeof(SRef) == 32, "reference record is 32 bytes");

struct STask {
    uint32_t node, begin, end, depth;
    uint32_t slot;  // object bins in obins (more than kSmallTask references), else kNoSlot
    uint32_t pad[3];
};
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Device record of one level, written by the level before (the root's by
// root_task_kernel; levels past the last carry the final counts): the
// level's counts, then its scan totals.
struct SLevel {
    uint32_t num_tasks, nrefs, counter, leaf_base;
    uint32_t dups;  // duplicate references made by the levels before
    uint32_t pad0[3];
    uint32_t cand_total, dup_total, splits, next_refs;
    uint32_t leaf_refs, slots;
    uint32_t fail;  // 1: out of reference, node or task space; 2: references lost
    uint32_t pad1;
};
static_assert(sizeof(SLevel) == 64, "level record is 64 bytes");

// Box accumulator with order-preserving float encodings; c0/c1 are counts
// (object bins: c0 references; spatial bins: c0 entries, c1 exits).
struct Acc {
    uint32_t c0;
    uint32_t mn[3];
    uint32_t mx[3];
    uint32_t c1;
};

enum : uint32_t { kLeaf = 0, kObject = 1, kSpatial = 2, kIndex = 3 };

struct SDec {
    float mn[3], mx[3];   // padded node box
    float org[3], inv[3];  // bins: b = (x - org) * inv, 16 per axis
    float width[3];       // bin width
    float area;           // half area of the unpadded node box
    uint32_t kind, axis, k, left, right;  // split bin boundary k; child sizes
    float cost;           // best cost so far
    uint32_t cand;        // 1: spatial split evaluated
    uint32_t s_axis, s_k, s_left, s_right, dup;  // best spatial split
    float s_cost;
};

__device__ __forceinline__ uint32_t fenc(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t e) {
    return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}

__device__ __forceinline__ void acc_clear(Acc& a) {
    a.c0 = a.c1 = 0;
    for (int k = 0; k < 3; ++k) {
        a.mn[k] = fenc(1e34f);
        a.mx[k] = fenc(-1e34f);
    }
}

__device__ __forceinline__ float half_area6(const float* mn, const float* mx) {
    const float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return (d0 + d1) * d2 + d0 * d1;
}

__device__ __forceinline__ uint32_t bin_of(float x, float org, float inv) {
    const float b = (x - org) * inv;
    if (!(b > 0.0f)) return 0u;  // also NaN
    if (b >= float(kBins - 1)) return uint32_t(kBins - 1);
    return uint32_t(b);
}

struct Box {
    float mn[3], mx[3];
};

__device__ __forceinline__ void box_empty(Box& b) {
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = 1e34f;
        b.mx[k] = -1e34f;
    }
}

__device__ __forceinline__ void box_grow_acc(Box& b, const Acc& a) {
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = fminf(b.mn[k], fdec(a.mn[k]));
        b.mx[k] = fmaxf(b.mx[k], fdec(a.mx[k]));
    }
}

__device__ __forceinline__ bool box_valid(const Box& b) {
    return b.mn[0] <= b.mx[0] && b.mn[1] <= b.mx[1] && b.mn[2] <= b.mx[2];
}

// v[ax] by selects: a runtime index into a register array would place the
// array in scratch memory.
__device__ __forceinline__ float sel3(const float* v, int ax) {
    return ax == 0 ? v[0] : (ax == 1 ? v[1] : v[2]);
}

// Bounding box of the part of triangle v (9 floats) with lo <= x[ax] <= hi,
// widened by rounding slack and kept inside [rmn, rmx]: the vertices inside
// the slab and the edges' crossings of its two planes bound the clipped
// polygon.  False when nothing is left.
__device__ __forceinline__ bool clip_box(const float* v, int ax, float lo, float hi,
                                         const float* rmn, const float* rmx, float* omn,
                                         float* omx) {
    for (int k = 0; k < 3; ++k) {
        omn[k] = 3.0e38f;
        omx[k] = -3.0e38f;
    }
    float p[3][3], pa[3];
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) p[i][k] = v[i * 3 + k];
        pa[i] = sel3(p[i], ax);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (pa[i] >= lo && pa[i] <= hi)
            for (int k = 0; k < 3; ++k) {
                omn[k] = fminf(omn[k], p[i][k]);
                omx[k] = fmaxf(omx[k], p[i][k]);
            }
        const float* a = p[i];
        const float* b = p[(i + 1) % 3];
        const float aa = pa[i], ba = pa[(i + 1) % 3];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const float plane = s ? hi : lo;
            if (!((aa < plane && ba > plane) || (aa > plane && ba < plane))) continue;
            const float t = (plane - aa) / (ba - aa);
            for (int k = 0; k < 3; ++k) {
                const float q = k == ax ? plane : a[k] + t * (b[k] - a[k]);
                omn[k] = fminf(omn[k], q);
                omx[k] = fmaxf(omx[k], q);
            }
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (k != ax) {
            // a crossing's other coordinates carry the rounding of t and of
            // the edge, a few ulps of the vertices' magnitude
            const float m = fmaxf(fmaxf(fabsf(p[0][k]), fabsf(p[1][k])), fabsf(p[2][k]));
            const float e = m * 0x1p-20f + 1e-30f;
            omn[k] -= e;
            omx[k] += e;
        } else {
            omn[k] = fmaxf(omn[k], lo);
            omx[k] = fminf(omx[k], hi);
        }
        omn[k] = fmaxf(omn[k], rmn[k]);
        omx[k] = fminf(omx[k], rmx[k]);
    }
    return omn[0] <= omx[0] && omn[1] <= omx[1] && omn[2] <= omx[2];
}

// ------------------------------------------------------------- kernels --

__global__ void init_refs_kernel(const DevAabb* aabbs, uint32_t n, SRef* refs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevAabb a = aabbs[i];
    SRef r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = a.min[k];
        r.mx[k] = a.max[k];
    }
    r.prim = i;
    r.task = 0;
    refs[i] = r;
}

__global__ void point_refs_kernel(const float* centers, uint32_t stride, uint32_t n,
                                  SRef* refs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SRef r;
    for (int k = 0; k < 3; ++k) r.mn[k] = r.mx[k] = centers[size_t(i) * stride + k];
    r.prim = i;
    r.task = 0;
    refs[i] = r;
}

// Per task: node box and object bins cleared (block t, 64 threads; a task of
// at most kSmallTask references has no stored object bins: obj_split_kernel
// bins it in registers).
// The spatial bins of the level's evaluated candidates (the first
// kMaxCand) cleared, the list of their tasks (list[cand_excl[t]] = t), and
// dup[t] = 0 for every other task.
__global__ void cand_clear_kernel(const SLevel* L, const uint32_t* cand_excl, uint32_t* list,
                                  uint32_t* dup, Acc* sbins) {
    const uint32_t num_tasks = L->num_tasks, total = L->cand_total;
    if (num_tasks == 0) return;
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < num_tasks; t += step) {
        const uint32_t e = cand_excl[t];
        const uint32_t next = t + 1 < num_tasks ? cand_excl[t + 1] : total;
        if (next > e && e < kMaxCand)
            list[e] = t;
        else
            dup[t] = 0;
    }
    const uint32_t count = min(total, kMaxCand) * 3 * kBins;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += step) {
        Acc a;
        acc_clear(a);
        sbins[i] = a;
    }
}

// A level's references are ordered by task (the scatter writes each split
// task's children, left then right, at the exclusive scan of the counts), so
// the run of references [first, end) of a block belongs to the tasks
// refs[first].task .. refs[end - 1].task.  When those are at most kLdsTasks,
// the block reduces in LDS and merges one value per task and slot; returns
// the first task, or -1 for the global-atomic path.  A reference whose task
// lies outside the run (never, by the ordering) takes the global path as
// well (in_run).  A block's run is `run` references (a multiple of kBlock):
// longer runs near the root, where a few tasks hold every reference, divide
// the blocks that merge into each task's bins.
constexpr uint32_t kSplitGrid = 8192;  // obj_split_kernel blocks at most (grid stride)
constexpr uint32_t kRunLong = 8 * kBlock, kRunMid = 4 * kBlock, kRunMidTasks = 256;
constexpr int kLdsTasks = 16;  // 4: 41 ms spatial, 22 ms binned at 1M; 16: 36, 19.5; 32: no better

__device__ __forceinline__ int32_t block_tasks(const SRef* refs, uint32_t first, uint32_t end,
                                               int32_t* count) {
    const int32_t a = refs[first].task, b = refs[end - 1].task;
    *count = b - a + 1;
    return (a >= 0 && b >= a && b - a < kLdsTasks) ? a : -1;
}

__device__ __forceinline__ bool in_run(int32_t t, int32_t t0, int32_t nt) {
    return t0 >= 0 && uint32_t(t - t0) < uint32_t(nt);
}

// Node boxes: the union of each task's reference boxes (tasks of at most
// kSmallTask references are left to task_setup_kernel).
__global__ void __launch_bounds__(kBlock)
    ref_bounds_kernel(const SLevel* L, const SRef* refs, const STask* tasks, Acc* tbox,
                      uint32_t run) {
    __shared__ uint32_t s_mn[kLdsTasks][3], s_mx[kLdsTasks][3];
    __shared__ int32_t s_t0, s_nt;
    const uint32_t nrefs = L->nrefs;
    const uint32_t first = blockIdx.x * run;
    if (first >= nrefs) return;  // whole block
    const uint32_t end = first + run < nrefs ? first + run : nrefs;
    if (threadIdx.x == 0) {
        int32_t nt = 0;
        s_t0 = block_tasks(refs, first, end, &nt);
        s_nt = nt;
    }
    if (threadIdx.x < kLdsTasks * 3) {
        s_mn[threadIdx.x / 3][threadIdx.x % 3] = fenc(1e34f);
        s_mx[threadIdx.x / 3][threadIdx.x % 3] = fenc(-1e34f);
    }
    __syncthreads();
    const int32_t t0 = s_t0;
    for (uint32_t i = first + threadIdx.x; i < end; i += kBlock) {
        const SRef r = refs[i];
        const bool lds = in_run(r.task, t0, s_nt);
        if (r.task >= 0 && tasks[r.task].end - tasks[r.task].begin > kSmallTask)
            for (int k = 0; k < 3; ++k) {
                if (!(r.mn[k] <= r.mx[k])) continue;  // empty or NaN
                if (lds) {
                    atomicMin(&s_mn[r.task - t0][k], fenc(r.mn[k]));
                    atomicMax(&s_mx[r.task - t0][k], fenc(r.mx[k]));
                } else {
                    atomicMin(&tbox[r.task].mn[k], fenc(r.mn[k]));
                    atomicMax(&tbox[r.task].mx[k], fenc(r.mx[k]));
                }
            }
    }
    __syncthreads();
    if (t0 >= 0 && threadIdx.x < uint32_t(s_nt) * 3) {
        const int tt = threadIdx.x / 3, k = threadIdx.x % 3;
        atomicMin(&tbox[t0 + tt].mn[k], s_mn[tt][k]);
        atomicMax(&tbox[t0 + tt].mx[k], s_mx[tt][k]);
    }
}

// Padded node box and bin parameters; the root's surface at level 0.
__global__ void task_setup_kernel(const SLevel* L, const STask* tasks, const SRef* refs,
                                  const Acc* tbox, SDec* dec, float* root_half_area,
                                  uint32_t level) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L->num_tasks) return;
    SDec d;
    const STask tk = tasks[t];
    Acc b;
    if (tk.end - tk.begin <= kSmallTask) {
        // ref_bounds_kernel's union for a small task: the same ordered
        // encodings, empty and NaN bounds skipped
        acc_clear(b);
        for (uint32_t i = tk.begin; i < tk.end; ++i) {
            const SRef r = refs[i];
            for (int k = 0; k < 3; ++k) {
                if (!(r.mn[k] <= r.mx[k])) continue;
                b.mn[k] = min(b.mn[k], fenc(r.mn[k]));
                b.mx[k] = max(b.mx[k], fenc(r.mx[k]));
            }
        }
    } else {
        b = tbox[t];
    }
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = fdec(b.mn[k]);
        mx[k] = fdec(b.mx[k]);
        if (!(mn[k] <= mx[k])) mn[k] = mx[k] = 0.0f;  // no valid reference box
        d.mn[k] = mn[k] - 0.0001f;
        d.mx[k] = mx[k] + 0.0001f;
        d.org[k] = d.mn[k];
        const float ext = d.mx[k] - d.mn[k];
        d.width[k] = ext / float(kBins);
        d.inv[k] = ext > 0.0f ? float(kBins) / ext : 0.0f;
    }
    d.kind = kLeaf;
    d.axis = d.k = d.left = d.right = 0;
    d.cost = 3.0e38f;
    d.cand = 0;
    d.s_axis = d.s_k = d.s_left = d.s_right = d.dup = 0;
    d.s_cost = 3.0e38f;
    d.area = half_area6(mn, mx);
    dec[t] = d;
    if (level == 0 && t == 0) *root_half_area = d.area;
}

__global__ void __launch_bounds__(kBlock)
    obj_bin_kernel(const SLevel* L, const SRef* refs, const STask* tasks, const SDec* dec,
                   Acc* obins, uint32_t run) {
    __shared__ Acc s_b[kLdsTasks][3 * kBins];
    __shared__ int32_t s_t0, s_nt;
    const uint32_t nrefs = L->nrefs;
    const uint32_t first = blockIdx.x * run;
    if (first >= nrefs) return;  // whole block
    const uint32_t end = first + run < nrefs ? first + run : nrefs;
    if (threadIdx.x == 0) {
        int32_t nt = 0;
        s_t0 = block_tasks(refs, first, end, &nt);
        s_nt = nt;
    }
    for (uint32_t k = threadIdx.x; k < uint32_t(kLdsTasks * 3 * kBins); k += kBlock)
        acc_clear(s_b[k / (3 * kBins)][k % (3 * kBins)]);
    __syncthreads();
    const int32_t t0 = s_t0;
    for (uint32_t i = first + threadIdx.x; i < end; i += kBlock) {
        const SRef r = refs[i];
        const int32_t t = r.task;
        if (t >= 0) {
            const STask tk = tasks[t];
            if (tk.end - tk.begin > kSmallTask && tk.depth < kMaxDepth) {
                const SDec& d = dec[t];
                for (int ax = 0; ax < 3; ++ax) {
                    const float c = 0.5f * (r.mn[ax] + r.mx[ax]);
                    const uint32_t slot = ax * kBins + bin_of(c, d.org[ax], d.inv[ax]);
                    Acc* g = in_run(t, t0, s_nt) ? &s_b[t - t0][slot]
                                                 : &obins[size_t(tk.slot) * 3 * kBins + slot];
                    atomicAdd(&g->c0, 1u);
                    for (int k = 0; k < 3; ++k) {
                        if (!(r.mn[k] <= r.mx[k])) continue;
                        atomicMin(&g->mn[k], fenc(r.mn[k]));
                        atomicMax(&g->mx[k], fenc(r.mx[k]));
                    }
                }
            }
        }
    }
    __syncthreads();
    if (t0 >= 0)
        for (uint32_t k = threadIdx.x; k < uint32_t(s_nt * 3 * kBins); k += kBlock) {
            const uint32_t tt = k / (3 * kBins), slot = k % (3 * kBins);
            const Acc& a = s_b[tt][slot];
            if (a.c0 == 0) continue;
            Acc* g = &obins[size_t(tasks[t0 + tt].slot) * 3 * kBins + slot];
            atomicAdd(&g->c0, a.c0);
            for (int q = 0; q < 3; ++q) {
                atomicMin(&g->mn[q], a.mn[q]);
                atomicMax(&g->mx[q], a.mx[q]);
            }
        }
}

// A run of bins: its box and its reference count.
struct GRun {
    float mn[3], mx[3];
    uint32_t cnt;
};

__device__ __forceinline__ void grun_join(GRun& a, const GRun& b) {
    for (int k = 0; k < 3; ++k) {
        a.mn[k] = fminf(a.mn[k], b.mn[k]);
        a.mx[k] = fmaxf(a.mx[k], b.mx[k]);
    }
    a.cnt += b.cnt;
}

__device__ __forceinline__ GRun grun_shfl(const GRun& b, int src) {
    GRun r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = __shfl(b.mn[k], src, kBins);
        r.mx[k] = __shfl(b.mx[k], src, kBins);
    }
    r.cnt = __shfl(b.cnt, src, kBins);
    return r;
}

__device__ __forceinline__ GRun grun_shfl_up(const GRun& b, int off) {
    GRun r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = __shfl_up(b.mn[k], off, kBins);
        r.mx[k] = __shfl_up(b.mx[k], off, kBins);
    }
    r.cnt = __shfl_up(b.cnt, off, kBins);
    return r;
}

__device__ __forceinline__ GRun grun_shfl_down(const GRun& b, int off) {
    GRun r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = __shfl_down(b.mn[k], off, kBins);
        r.mx[k] = __shfl_down(b.mx[k], off, kBins);
    }
    r.cnt = __shfl_down(b.cnt, off, kBins);
    return r;
}

// The run of axis ax, member by member: three named values and selects (a
// runtime index into an array of runs would place it in scratch memory).
__device__ __forceinline__ GRun grun_pick(uint32_t ax, const GRun& r0, const GRun& r1,
                                          const GRun& r2) {
    GRun g;
    for (int k = 0; k < 3; ++k) {
        g.mn[k] = ax == 0 ? r0.mn[k] : (ax == 1 ? r1.mn[k] : r2.mn[k]);
        g.mx[k] = ax == 0 ? r0.mx[k] : (ax == 1 ? r1.mx[k] : r2.mx[k]);
    }
    g.cnt = ax == 0 ? r0.cnt : (ax == 1 ? r1.cnt : r2.cnt);
    return g;
}

// Best object split and whether a spatial split is worth evaluating.  One
// group of 16 lanes per task, lane i holding object bin i of each axis:
// prefix runs (bins 0..i) and suffix runs (i..15) by shuffle scans, the cost
// of the plane after bin i as a sweep computes it, half_area(left) * left
// count + half_area(right) * right count, skipping planes with an empty
// side, and the lowest cost below the node's current one, the first in
// (axis, plane) order, by a reduction.  Bin bounds hold no NaN, so the runs
// equal a sweep's up to the sign of a zero, which no comparison tells apart.
// A task of at most kSmallTask references is binned here (lane q takes
// reference q, lane j gathers bin j with obj_bin_kernel's ordered-encoding
// minimum and maximum); decide_kernel counts its references itself.
__global__ void __launch_bounds__(kBlock)
    obj_split_kernel(const SLevel* L, const STask* tasks, SDec* dec, const Acc* obins,
                     const SRef* refs, const float* root_half_area, uint32_t* cand_flag) {
    const uint32_t num_tasks = L->num_tasks;
    const int lane = int(threadIdx.x % kBins);
    // whole groups (kBlock is a multiple of kBins), grid-stride over tasks
    for (uint32_t t = (blockIdx.x * kBlock + threadIdx.x) / kBins; t < num_tasks;
         t += gridDim.x * (kBlock / kBins)) {
        const STask tk = tasks[t];
        SDec d = dec[t];
        const uint32_t work = tk.end - tk.begin;
        uint32_t cand = 0;
        if (work > 1 && tk.depth < kMaxDepth) {
            Acc own[3];
            if (work <= kSmallTask) {
                uint32_t pb[3] = {0, 0, 0}, pmn[3], pmx[3];
                for (int k = 0; k < 3; ++k) {
                    pmn[k] = fenc(1e34f);
                    pmx[k] = fenc(-1e34f);
                }
                if (uint32_t(lane) < work) {
                    const SRef r = refs[tk.begin + lane];
                    for (int k = 0; k < 3; ++k) {
                        pb[k] = bin_of(0.5f * (r.mn[k] + r.mx[k]), d.org[k], d.inv[k]);
                        if (r.mn[k] <= r.mx[k]) {
                            pmn[k] = fenc(r.mn[k]);
                            pmx[k] = fenc(r.mx[k]);
                        }
                    }
                }
                for (int ax = 0; ax < 3; ++ax) acc_clear(own[ax]);
                for (uint32_t q = 0; q < work; ++q) {
                    uint32_t qb[3], qmn[3], qmx[3];
                    for (int k = 0; k < 3; ++k) {
                        qb[k] = __shfl(pb[k], int(q), kBins);
                        qmn[k] = __shfl(pmn[k], int(q), kBins);
                        qmx[k] = __shfl(pmx[k], int(q), kBins);
                    }
                    for (int ax = 0; ax < 3; ++ax)
                        if (qb[ax] == uint32_t(lane)) {
                            ++own[ax].c0;
                            for (int k = 0; k < 3; ++k) {
                                own[ax].mn[k] = min(own[ax].mn[k], qmn[k]);
                                own[ax].mx[k] = max(own[ax].mx[k], qmx[k]);
                            }
                        }
                }
            } else {
                const Acc* tb = obins + size_t(tk.slot) * 3 * kBins;
                for (int ax = 0; ax < 3; ++ax) own[ax] = tb[ax * kBins + lane];
            }
            GRun pre0, pre1, pre2, suf0, suf1, suf2;
            float key = __builtin_inff();
            int idx = 3 * kBins;
    #pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                GRun p;
                for (int k = 0; k < 3; ++k) {
                    p.mn[k] = fdec(own[ax].mn[k]);
                    p.mx[k] = fdec(own[ax].mx[k]);
                }
                p.cnt = own[ax].c0;
                GRun q = p;
                for (int off = 1; off < kBins; off <<= 1) {
                    GRun lower = grun_shfl_up(p, off);
                    const GRun higher = grun_shfl_down(q, off);
                    if (lane >= off) {
                        grun_join(lower, p);
                        p = lower;
                    }
                    if (lane + off < kBins) grun_join(q, higher);
                }
                if (ax == 0) {
                    pre0 = p;
                    suf0 = q;
                } else if (ax == 1) {
                    pre1 = p;
                    suf1 = q;
                } else {
                    pre2 = p;
                    suf2 = q;
                }
                const GRun right = grun_shfl_down(q, 1);
                if (lane < kBins - 1 && p.cnt != 0 && p.cnt != work) {
                    const float c = half_area6(p.mn, p.mx) * float(p.cnt) +
                                    half_area6(right.mn, right.mx) * float(right.cnt);
                    if (c < d.cost && c < key) {  // first axis wins a tie
                        key = c;
                        idx = ax * kBins + lane;
                    }
                }
            }
            for (int off = kBins / 2; off > 0; off >>= 1) {
                const float ok = __shfl_xor(key, off, kBins);
                const int oi = __shfl_xor(idx, off, kBins);
                if (ok < key || (ok == key && oi < idx)) {
                    key = ok;
                    idx = oi;
                }
            }
            if (idx < 3 * kBins) {
                const uint32_t ax = uint32_t(idx / kBins);
                const int i = idx % kBins;
                const GRun lbest = grun_shfl(grun_pick(ax, pre0, pre1, pre2), i);
                const GRun rbest = grun_shfl(grun_pick(ax, suf0, suf1, suf2), i + 1);
                d.cost = key;
                d.kind = kObject;
                d.axis = ax;
                d.k = uint32_t(i + 1);
                d.left = lbest.cnt;
                d.right = work - lbest.cnt;
                Box ov;
                for (int k = 0; k < 3; ++k) {
                    ov.mn[k] = fmaxf(lbest.mn[k], rbest.mn[k]);
                    ov.mx[k] = fminf(lbest.mx[k], rbest.mx[k]);
                }
                if (box_valid(ov) && half_area6(ov.mn, ov.mx) > kAlpha * *root_half_area) cand = 1;
            } else {
                cand = 1;  // every centroid in one bin: a plane may still separate them
            }
        }
        if (lane == 0) {
            d.cand = cand;
            dec[t] = d;
            cand_flag[t] = cand;
        }
    }
}

// Chopped spatial binning of the references of candidate tasks.  A
// reference counts one entry (c0) in its first bin and one exit (c1) in its
// last, and grows the box of every bin it spans with the triangle clipped to
// that bin, so a bin can hold a box without counts.  Blocks spanning at most
// kLdsTasks tasks bin in LDS and merge every touched bin.
__device__ __forceinline__ bool sp_evaluated(const SDec* dec, const uint32_t* cand_excl,
                                             int32_t t) {
    return dec[t].cand && cand_excl[t] < kMaxCand;
}

__global__ void __launch_bounds__(kBlock)
    sp_bin_kernel(const SLevel* L, const SRef* refs, const SDec* dec, const uint32_t* cand_excl,
                  const float* verts, uint32_t prim_count, Acc* sbins, uint32_t run) {
    __shared__ Acc s_b[kLdsTasks][3 * kBins];
    __shared__ int32_t s_t0, s_nt;
    const uint32_t nrefs = L->nrefs;
    const uint32_t first = blockIdx.x * run;
    if (first >= nrefs || L->cand_total == 0) return;  // whole block
    const uint32_t end = first + run < nrefs ? first + run : nrefs;
    if (threadIdx.x == 0) {
        int32_t nt = 0;
        s_t0 = block_tasks(refs, first, end, &nt);
        s_nt = nt;
    }
    for (uint32_t k = threadIdx.x; k < uint32_t(kLdsTasks * 3 * kBins); k += kBlock)
        acc_clear(s_b[k / (3 * kBins)][k % (3 * kBins)]);
    __syncthreads();
    const int32_t t0 = s_t0;
    for (uint32_t i = first + threadIdx.x; i < end; i += kBlock) {
        const SRef r = refs[i];
        const int32_t t = r.task;
        if (t >= 0 && sp_evaluated(dec, cand_excl, t) && r.prim < prim_count) {
            const SDec& d = dec[t];
            Acc* base = in_run(t, t0, s_nt) ? s_b[t - t0]
                                            : sbins + size_t(cand_excl[t]) * 3 * kBins;
            const float* v = verts + size_t(r.prim) * 9;
            for (int ax = 0; ax < 3; ++ax) {
                const float org = sel3(d.org, ax), inv = sel3(d.inv, ax);
                const uint32_t b0 = bin_of(sel3(r.mn, ax), org, inv);
                const uint32_t b1 = max(b0, bin_of(sel3(r.mx, ax), org, inv));
                atomicAdd(&base[ax * kBins + b0].c0, 1u);
                atomicAdd(&base[ax * kBins + b1].c1, 1u);
                for (uint32_t b = b0; b <= b1 && b < uint32_t(kBins); ++b) {
                    float mn[3], mx[3];
                    if (b0 == b1) {
                        for (int k = 0; k < 3; ++k) {
                            mn[k] = r.mn[k];
                            mx[k] = r.mx[k];
                        }
                    } else {
                        const float lo = b == b0 ? -3.0e38f : d.org[ax] + float(b) * d.width[ax];
                        const float hi =
                            b == b1 ? 3.0e38f : d.org[ax] + float(b + 1) * d.width[ax];
                        if (!clip_box(v, ax, lo, hi, r.mn, r.mx, mn, mx)) continue;
                    }
                    Acc* g = &base[ax * kBins + b];
                    for (int k = 0; k < 3; ++k) {
                        if (!(mn[k] <= mx[k])) continue;
                        atomicMin(&g->mn[k], fenc(mn[k]));
                        atomicMax(&g->mx[k], fenc(mx[k]));
                    }
                }
            }
        }
    }
    __syncthreads();
    if (t0 >= 0)
        for (uint32_t k = threadIdx.x; k < uint32_t(s_nt * 3 * kBins); k += kBlock) {
            const uint32_t tt = k / (3 * kBins), slot = k % (3 * kBins);
            const int32_t t = t0 + int32_t(tt);
            if (!sp_evaluated(dec, cand_excl, t)) continue;
            const Acc& a = s_b[tt][slot];
            const uint32_t empty = fenc(1e34f);
            if ((a.c0 | a.c1) == 0 && a.mn[0] == empty && a.mn[1] == empty && a.mn[2] == empty)
                continue;
            Acc* g = &sbins[size_t(cand_excl[t]) * 3 * kBins + slot];
            if (a.c0) atomicAdd(&g->c0, a.c0);
            if (a.c1) atomicAdd(&g->c1, a.c1);
            for (int q = 0; q < 3; ++q) {
                atomicMin(&g->mn[q], a.mn[q]);
                atomicMax(&g->mx[q], a.mx[q]);
            }
        }
}

// Best spatial split of each evaluated candidate (list: their tasks);
// dup[t] = its duplicates when it beats the object split, else 0.  One group
// of 16 lanes per candidate, lane i holding spatial bin i of each axis: the
// boxes and entry counts of bins 0..i and the boxes and exit counts of bins
// i+1..15 by shuffle scans, the cost of the plane after bin i as a sweep
// computes it (left half area * entries + right half area * exits, planes
// with an empty side or fewer references than the task skipped), and the
// lowest cost below the task's, the first in (axis, plane) order.  Bin
// bounds hold no NaN, so the unions equal a sweep's up to the sign of a
// zero, which no comparison tells apart.
__global__ void __launch_bounds__(kBlock)
    sp_split_kernel(const SLevel* L, const uint32_t* list, const STask* tasks, SDec* dec,
                    const Acc* sbins, uint32_t* dup) {
    const uint32_t c = (blockIdx.x * kBlock + threadIdx.x) / kBins;
    const int lane = int(threadIdx.x % kBins);
    if (c >= min(L->cand_total, kMaxCand)) return;  // whole groups
    const uint32_t t = list[c];
    const STask tk = tasks[t];
    SDec d = dec[t];
    const uint32_t work = tk.end - tk.begin;
    const Acc* tb = sbins + size_t(c) * 3 * kBins;
    float key = d.s_cost;
    int idx = 3 * kBins;
    uint32_t best_l = 0, best_r = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const Acc a = tb[ax * kBins + lane];
        GRun p, q;
        for (int k = 0; k < 3; ++k) {
            p.mn[k] = q.mn[k] = fdec(a.mn[k]);
            p.mx[k] = q.mx[k] = fdec(a.mx[k]);
        }
        p.cnt = a.c0;  // entries, bins 0..lane
        q.cnt = a.c1;  // exits, bins lane..15
        for (int off = 1; off < kBins; off <<= 1) {
            GRun lower = grun_shfl_up(p, off);
            const GRun higher = grun_shfl_down(q, off);
            if (lane >= off) {
                grun_join(lower, p);
                p = lower;
            }
            if (lane + off < kBins) grun_join(q, higher);
        }
        const GRun right = grun_shfl_down(q, 1);
        const uint32_t l = p.cnt, r = right.cnt;
        if (lane < kBins - 1 && l != 0 && r != 0 && l + r >= work) {
            const float cost =
                half_area6(p.mn, p.mx) * float(l) + half_area6(right.mn, right.mx) * float(r);
            if (cost < key) {  // the first axis wins a tie
                key = cost;
                idx = ax * kBins + lane;
                best_l = l;
                best_r = r;
            }
        }
    }
    for (int off = kBins / 2; off > 0; off >>= 1) {
        const float ok = __shfl_xor(key, off, kBins);
        const int oi = __shfl_xor(idx, off, kBins);
        if (ok < key || (ok == key && oi < idx)) {
            key = ok;
            idx = oi;
        }
    }
    // the winner's counts, from the lane that holds them
    const int src = idx < 3 * kBins ? idx % kBins : 0;
    const uint32_t wl = __shfl(best_l, src, kBins), wr = __shfl(best_r, src, kBins);
    if (lane != 0) return;
    uint32_t want = 0;
    if (idx < 3 * kBins) {
        d.s_cost = key;
        d.s_axis = uint32_t(idx / kBins);
        d.s_k = uint32_t(idx % kBins + 1);
        d.s_left = wl;
        d.s_right = wr;
        if (d.s_cost < d.cost) {
            d.dup = wl + wr - work;
            want = 1;
        }
        dec[t] = d;
    }
    dup[t] = want ? d.dup : 0u;
}

// Final decision per task; counts for the scans (big_cnt: children of more
// than kSmallTask references, which get object-bin slots).  A spatial split
// is taken while the duplicates of the wanted spatial splits up to and
// including its own fit what the budget has left after the levels before.
__global__ void decide_kernel(const SLevel* L, const STask* tasks, SDec* dec,
                              const uint32_t* dup_excl, uint32_t budget, uint32_t max_leaf,
                              const Acc* obins, const SRef* refs, uint32_t* split_flag,
                              uint32_t* out_cnt, uint32_t* leaf_cnt, uint32_t* big_cnt) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L->num_tasks) return;
    const uint32_t remaining = L->dups < budget ? budget - L->dups : 0u;
    const STask tk = tasks[t];
    SDec d = dec[t];
    const uint32_t work = tk.end - tk.begin;
    if (work > 1 && tk.depth < kMaxDepth) {
        float best = d.cost;
        if (d.s_cost < d.cost && d.dup <= remaining && dup_excl[t] <= remaining - d.dup) {
            d.kind = kSpatial;
            d.axis = d.s_axis;
            d.k = d.s_k;
            d.left = d.s_left;
            d.right = d.s_right;
            best = d.s_cost;
        }
        // against the unpadded box, as the reference compares before it pads
        // a parent (spatial_sah.rs:781-782, :352): the padding would make
        // every split of a small node look cheap
        const float max_split_cost = d.area * (float(work) - 1.0f);
        if (d.kind == kLeaf || !(best < max_split_cost)) {
            if (work > max_leaf) {
                // approximate median on the longest axis; an index median
                // when the centroids share one bin
                const float e0 = d.mx[0] - d.mn[0], e1 = d.mx[1] - d.mn[1],
                            e2 = d.mx[2] - d.mn[2];
                uint32_t ax = e1 > e0 ? 1u : 0u;
                if (e2 > (ax ? e1 : e0)) ax = 2;
                const float org = sel3(d.org, int(ax)), inv = sel3(d.inv, int(ax));
                const Acc* b = work > kSmallTask
                                   ? obins + size_t(tk.slot) * 3 * kBins + ax * kBins
                                   : nullptr;
                // a small task's bins were never stored: count its
                // references' centroid bins (obj_bin_kernel's count) in one
                // pass, 8-bit counters packed eight to a word
                uint64_t h_lo = 0, h_hi = 0;
                if (work <= kSmallTask)
                    for (uint32_t q = tk.begin; q < tk.end; ++q) {
                        const SRef& r = refs[q];
                        const uint32_t bq = bin_of(0.5f * (r.mn[ax] + r.mx[ax]), org, inv);
                        if (bq < 8)
                            h_lo += 1ull << (8 * bq);
                        else
                            h_hi += 1ull << (8 * (bq - 8));
                    }
                uint32_t c = 0, k = 0;
                for (int i = 0; i < kBins - 1; ++i) {
                    if (work > kSmallTask)
                        c += b[i].c0;
                    else
                        c += uint32_t((i < 8 ? h_lo >> (8 * i) : h_hi >> (8 * (i - 8))) & 0xFFu);
                    if (c >= work * 2 / 5 + 1) {
                        k = uint32_t(i + 1);
                        break;
                    }
                }
                if (k > 0 && c > 0 && c < work) {
                    d.kind = kObject;
                    d.axis = ax;
                    d.k = k;
                    d.left = c;
                    d.right = work - c;
                } else {
                    d.kind = kIndex;
                    d.left = work / 2;
                    d.right = work - work / 2;
                }
            } else {
                d.kind = kLeaf;
            }
        }
    } else {
        d.kind = kLeaf;
    }
    dec[t] = d;
    const bool split = d.kind != kLeaf;
    split_flag[t] = split ? 1u : 0u;
    out_cnt[t] = split ? d.left + d.right : 0u;
    leaf_cnt[t] = split ? 0u : work;
    big_cnt[t] = split ? (d.left > kSmallTask ? 1u : 0u) + (d.right > kSmallTask ? 1u : 0u) : 0u;
}

// Node records and the next level's tasks (the children of split task t
// are tasks 2 split_excl[t] and the one after, nodes counter + the same).
__global__ void children_kernel(const SLevel* L, const STask* tasks, const SDec* dec,
                                const uint32_t* split_excl, const uint32_t* out_excl,
                                const uint32_t* leaf_excl, const uint32_t* big_excl,
                                uint32_t node_cap, uint32_t task_cap, DevAabb* nodes, STask* next,
                                uint32_t* fail) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L->num_tasks) return;
    const uint32_t counter = L->counter, leaf_base = L->leaf_base;
    const STask tk = tasks[t];
    const SDec& d = dec[t];
    DevAabb nd;
    for (int k = 0; k < 3; ++k) {
        nd.min[k] = d.mn[k];
        nd.max[k] = d.mx[k];
    }
    if (d.kind != kLeaf) {
        const uint32_t child = counter + 2 * split_excl[t];
        if (child + 1 >= node_cap || 2 * split_excl[t] + 1 >= task_cap) {
            atomicOr(fail, 1u);
            return;
        }
        nd.extra1 = -1;
        nd.extra2 = int32_t(child);
        STask l, r;
        l.node = child;
        r.node = child + 1;
        l.begin = out_excl[t];
        l.end = l.begin + d.left;
        r.begin = l.end;
        r.end = r.begin + d.right;
        l.depth = r.depth = tk.depth + 1;
        l.slot = d.left > kSmallTask ? big_excl[t] : kNoSlot;
        r.slot = d.right > kSmallTask ? big_excl[t] + (d.left > kSmallTask ? 1u : 0u) : kNoSlot;
        l.pad[0] = l.pad[1] = l.pad[2] = r.pad[0] = r.pad[1] = r.pad[2] = 0;
        next[2 * split_excl[t]] = l;
        next[2 * split_excl[t] + 1] = r;
    } else {
        nd.extra1 = int32_t(tk.end - tk.begin);
        nd.extra2 = int32_t(leaf_base + leaf_excl[t]);
    }
    if (tk.node < node_cap) nodes[tk.node] = nd;
}

// Which child (or both) each reference goes to.
// Side flags of each reference, written twice: kept in fl/fr, scanned in
// place in fle/fre.
__global__ void side_kernel(const SLevel* L, const SRef* refs, const STask* tasks,
                            const SDec* dec, uint32_t* fl, uint32_t* fr, uint32_t* fle,
                            uint32_t* fre) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L->nrefs) return;
    const SRef r = refs[i];
    uint32_t l = 0, rr = 0;
    if (r.task >= 0) {
        const SDec& d = dec[r.task];
        const uint32_t ax = d.axis;
        if (d.kind == kObject) {
            l = bin_of(0.5f * (r.mn[ax] + r.mx[ax]), d.org[ax], d.inv[ax]) < d.k ? 1u : 0u;
            rr = 1u - l;
        } else if (d.kind == kSpatial) {  // the spans of the spatial binning
            const uint32_t b0 = bin_of(r.mn[ax], d.org[ax], d.inv[ax]);
            const uint32_t b1 = max(b0, bin_of(r.mx[ax], d.org[ax], d.inv[ax]));
            l = b0 < d.k ? 1u : 0u;
            rr = b1 >= d.k ? 1u : 0u;
        } else if (d.kind == kIndex) {
            l = i - tasks[r.task].begin < d.left ? 1u : 0u;
            rr = 1u - l;
        }
    }
    fl[i] = l;
    fr[i] = rr;
    fle[i] = l;
    fre[i] = rr;
}

__global__ void scatter_kernel(const SLevel* L, const SRef* refs, const STask* tasks,
                               const SDec* dec, const uint32_t* split_excl,
                               const uint32_t* out_excl, const uint32_t* leaf_excl,
                               const uint32_t* fl, const uint32_t* fr, const uint32_t* fl_excl,
                               const uint32_t* fr_excl, const float* verts, uint32_t prim_count,
                               uint32_t cap, SRef* out, uint32_t* indices, uint32_t* fail,
                               Acc* tbox, Acc* obins, uint32_t slot_cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // the next level's node boxes and object bins: this level's readers
    // (task_setup_kernel, obj_split_kernel, decide_kernel) are done
    if (L->num_tasks > 0) {
        const uint32_t step = gridDim.x * blockDim.x;
        // (bounded by the buffers: counts past them have failed the build)
        const uint32_t next_tasks = min(2 * L->splits, cap);
        const uint32_t bins = min(L->slots, slot_cap) * 3 * kBins;
        for (uint32_t t = i; t < next_tasks; t += step) acc_clear(tbox[t]);
        for (uint32_t q = i; q < bins; q += step) acc_clear(obins[q]);
    }
    if (i >= L->nrefs) return;
    const uint32_t leaf_base = L->leaf_base;
    const SRef r = refs[i];
    if (r.task < 0) return;
    const STask tk = tasks[r.task];
    const SDec& d = dec[r.task];
    if (d.kind == kLeaf) {
        const uint32_t pos = leaf_base + leaf_excl[r.task] + (i - tk.begin);
        if (pos < cap)
            indices[pos] = r.prim;
        else
            atomicOr(fail, 1u);
        return;
    }
    const bool l = fl[i] != 0, rr = fr[i] != 0;
    SRef a = r, b = r;
    if (l && rr) {  // a straddling reference: one clipped part per side
        const int ax = int(d.axis);
        const float plane = d.org[ax] + float(d.k) * d.width[ax];
        bool ok_l = false, ok_r = false;
        if (r.prim < prim_count) {
            const float* v = verts + size_t(r.prim) * 9;
            ok_l = clip_box(v, ax, -3.0e38f, plane, r.mn, r.mx, a.mn, a.mx);
            ok_r = clip_box(v, ax, plane, 3.0e38f, r.mn, r.mx, b.mn, b.mx);
        }
        // keep coverage when clipping leaves nothing on a side: the
        // reference's own box cut at the plane
        if (!ok_l)
            for (int k = 0; k < 3; ++k) {
                a.mn[k] = r.mn[k];
                a.mx[k] = k == ax ? fmaxf(fminf(r.mx[k], plane), r.mn[k]) : r.mx[k];
            }
        if (!ok_r)
            for (int k = 0; k < 3; ++k) {
                b.mn[k] = k == ax ? fminf(fmaxf(r.mn[k], plane), r.mx[k]) : r.mn[k];
                b.mx[k] = r.mx[k];
            }
    }
    if (l) {
        const uint32_t pos = out_excl[r.task] + (fl_excl[i] - fl_excl[tk.begin]);
        a.task = int32_t(2 * split_excl[r.task]);
        if (pos < cap)
            out[pos] = a;
        else
            atomicOr(fail, 1u);
    }
    if (rr) {
        const uint32_t pos = out_excl[r.task] + d.left + (fr_excl[i] - fr_excl[tk.begin]);
        b.task = int32_t(2 * split_excl[r.task] + 1);
        if (pos < cap)
            out[pos] = b;
        else
            atomicOr(fail, 1u);
    }
}

// The root task, level 0's record (the records are zero before) and the
// root's node box and object bins cleared (scatter_kernel clears them for
// the levels after).
__global__ void root_task_kernel(STask* task, SLevel* levels, uint32_t n, Acc* tbox,
                                 Acc* obins) {
    if (blockIdx.x != 0) return;
    if (threadIdx.x < uint32_t(3 * kBins)) acc_clear(obins[threadIdx.x]);
    if (threadIdx.x == 0) {
        acc_clear(tbox[0]);
        STask t;
        t.node = 0;
        t.begin = 0;
        t.end = n;
        t.depth = 0;
        t.slot = n > kSmallTask ? 0u : kNoSlot;
        t.pad[0] = t.pad[1] = t.pad[2] = 0;
        *task = t;
        levels[0].num_tasks = 1;
        levels[0].nrefs = n;
        levels[0].counter = 1;
    }
}

// levels[level + 1] from the counts and scan totals of levels[level]; a
// level without tasks, or one that failed, leaves none to the next, so
// every later level returns at once and carries the final counts.
__global__ void level_end_kernel(SLevel* levels, uint32_t level, uint32_t cap,
                                 const uint32_t* fail_word) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const SLevel L = levels[level];
    SLevel nx;
    nx.num_tasks = 0;
    nx.nrefs = 0;
    nx.counter = L.counter;
    nx.leaf_base = L.leaf_base;
    nx.dups = L.dups;
    nx.pad0[0] = nx.pad0[1] = nx.pad0[2] = 0;
    nx.cand_total = nx.dup_total = nx.splits = nx.next_refs = 0;
    nx.leaf_refs = nx.slots = 0;
    nx.fail = L.fail | *fail_word;
    nx.pad1 = 0;
    if (L.num_tasks > 0 && nx.fail == 0) {
        if (L.next_refs + L.leaf_refs < L.nrefs || L.next_refs > cap ||
            L.leaf_base + L.leaf_refs > cap) {
            nx.fail = 2;
        } else {
            nx.num_tasks = 2 * L.splits;
            nx.nrefs = L.next_refs;
            nx.counter = L.counter + 2 * L.splits;
            nx.leaf_base = L.leaf_base + L.leaf_refs;
            nx.dups = L.dups + (L.next_refs + L.leaf_refs - L.nrefs);
        }
    }
    levels[level + 1] = nx;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

ResultCode gpu_build_spatial(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                             const float* verts, size_t n64, size_t max_leaf, hipStream_t stream,
                             DeviceBvh* out) {
    if (n64 == 0) return NoPrimitives;
    if (n64 > (1u << 29)) {
        set_error("prim_count %zu too large for the spatial builder", n64);
        return InvalidArgument;
    }
    const uint32_t n = uint32_t(n64);
    const uint32_t cap = n + uint32_t(float(n) * kSplitFactor);  // references at most
    const uint32_t node_cap = 2 * cap;
    const uint32_t leaf = max_leaf == 0 ? 1u : uint32_t(std::min<size_t>(max_leaf, 0xFFFFFFFFu));
    const uint32_t stride = uint32_t(stride_floats);
    hipError_t e;
    DeviceBvh res;
    if ((e = hipMalloc(&res.nodes, size_t(node_cap) * sizeof(DevAabb))) != hipSuccess)
        return hip_fail(e, "hipMalloc(nodes)");
    if ((e = hipMalloc(&res.indices, size_t(cap) * 4)) != hipSuccess) {
        (void)hipFree(res.nodes);
        return hip_fail(e, "hipMalloc(indices)");
    }
    auto fail = [&](hipError_t err, const char* what) {
        (void)hipFree(res.nodes);
        (void)hipFree(res.indices);
        return hip_fail(err, what);
    };
    auto fail_msg = [&]() {  // the message is already set
        (void)hipFree(res.nodes);
        (void)hipFree(res.indices);
        return Error;
    };
    // One workspace.  A level holds at most cap references and at most cap
    // tasks (each has a reference); at most cap / (kSmallTask + 1) tasks
    // have object-bin slots.
    constexpr uint32_t kLevels = kMaxDepth + 3;  // records of levels 0 .. kMaxDepth + 2
    const size_t slot_cap = size_t(cap) / (kSmallTask + 1) + 1;
    const size_t sz_refs = align256(size_t(cap) * sizeof(SRef));
    const size_t sz_tasks = align256(size_t(cap) * sizeof(STask));
    const size_t sz_dec = align256(size_t(cap) * sizeof(SDec));
    const size_t sz_tbox = align256(size_t(cap) * sizeof(Acc));
    const size_t sz_obins = align256(slot_cap * 3 * kBins * sizeof(Acc));
    const size_t sz_sbins = align256(size_t(std::min(cap, kMaxCand)) * 3 * kBins * sizeof(Acc));
    const size_t sz_u32 = align256(size_t(cap) * 4);
    const size_t sz_list = align256(size_t(std::min(cap, kMaxCand)) * 4);
    const size_t sz_scan = align256(scan_dev_scratch_bytes(cap, 6));
    const size_t sz_levels = align256(size_t(kLevels) * sizeof(SLevel));
    const size_t sz_words = 256;
    const size_t total = 2 * sz_refs + 2 * sz_tasks + sz_dec + sz_tbox + sz_obins + sz_sbins +
                         10 * sz_u32 + sz_list + sz_scan + sz_levels + sz_words;
    Workspace wsh;
    if ((e = workspace_acquire(kWsSpatial, total, &wsh)) != hipSuccess)
        return fail(e, "spatial builder workspace");
    // released on every return, after the stream's work that used it
    struct WsRelease {
        Workspace* w;
        hipStream_t s;
        ~WsRelease() { workspace_release(w, s, true); }
    } wsr{&wsh, stream};
    SLevel* host_level = nullptr;
    if ((e = hipHostMalloc(&host_level, sizeof(SLevel))) != hipSuccess)
        return fail(e, "hipHostMalloc");
    struct HostFree {
        SLevel* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{host_level};
    char* cur = static_cast<char*>(wsh.ptr);
    auto carve = [&](size_t bytes) {
        char* p = cur;
        cur += bytes;
        return p;
    };
    SRef* refs[2] = {reinterpret_cast<SRef*>(carve(sz_refs)),
                     reinterpret_cast<SRef*>(carve(sz_refs))};
    STask* tasks[2] = {reinterpret_cast<STask*>(carve(sz_tasks)),
                       reinterpret_cast<STask*>(carve(sz_tasks))};
    SDec* dec = reinterpret_cast<SDec*>(carve(sz_dec));
    Acc* tbox = reinterpret_cast<Acc*>(carve(sz_tbox));
    Acc* obins = reinterpret_cast<Acc*>(carve(sz_obins));
    Acc* sbins = reinterpret_cast<Acc*>(carve(sz_sbins));
    uint32_t* u32[10];
    for (int k = 0; k < 10; ++k) u32[k] = reinterpret_cast<uint32_t*>(carve(sz_u32));
    uint32_t* cand_list = reinterpret_cast<uint32_t*>(carve(sz_list));
    uint32_t *split_flag = u32[0], *out_cnt = u32[1], *leaf_cnt = u32[2], *big_cnt = u32[3];
    uint32_t *cand = u32[4], *dup = u32[5], *fl = u32[6], *fr = u32[7], *fle = u32[8],
             *fre = u32[9];
    void* scan_scratch = carve(sz_scan);
    char* cleared = cur;
    SLevel* levels = reinterpret_cast<SLevel*>(carve(sz_levels));
    uint32_t* words = reinterpret_cast<uint32_t*>(carve(sz_words));
    float* root_half_area = reinterpret_cast<float*>(words);
    uint32_t* fail_word = words + 1;
    if ((e = hipMemsetAsync(cleared, 0, sz_levels + sz_words, stream)) != hipSuccess)
        return fail(e, "spatial builder state");
    const uint32_t g0 = (n + kBlock - 1) / kBlock;
    if (aabbs)
        hipLaunchKernelGGL(init_refs_kernel, dim3(g0), dim3(kBlock), 0, stream,
                           reinterpret_cast<const DevAabb*>(aabbs), n, refs[0]);
    else
        hipLaunchKernelGGL(point_refs_kernel, dim3(g0), dim3(kBlock), 0, stream, centers, stride,
                           n, refs[0]);
    hipLaunchKernelGGL(root_task_kernel, dim3(1), dim3(64), 0, stream, tasks[0], levels, n, tbox,
                       obins);
    const uint32_t budget = cap - n;
    // Levels are enqueued in batches without waiting; the host reads one
    // level record per batch.  Grids come from bounds on the counts of the
    // record read last (level r: T tasks, N references, D duplicates): a
    // later level holds at most N + (budget - D) references and at most
    // T << (levels since r) tasks.  The first batch covers a tree twice as
    // deep as a balanced one, then batches of 8 until a level has no task.
    uint32_t lg = 0;
    while (lg < 31 && (1u << lg) < n) ++lg;
    uint32_t batch = std::min<uint32_t>(kMaxDepth + 2, 2 * lg + 2);
    uint32_t level = 0;
    uint64_t known_tasks = 1, known_refs = n, known_dups = 0;
    for (;;) {
        const uint32_t end = std::min<uint32_t>(kMaxDepth + 2, level + batch);
        const uint32_t ref_bound =
            uint32_t(std::min<uint64_t>(cap, known_refs + (budget - known_dups)));
        const uint32_t rg = (ref_bound + kBlock - 1) / kBlock;
        for (uint32_t k = level; k < end; ++k) {
            const uint32_t sh = std::min<uint32_t>(k - level, 32);
            const uint32_t tb = uint32_t(std::min<uint64_t>(ref_bound, known_tasks << sh));
            // binning runs: kRunLong references per block while the level has
            // at most kLdsTasks tasks, kRunMid up to kRunMidTasks, else kBlock
            const uint32_t run = tb <= uint32_t(kLdsTasks) ? kRunLong
                                 : tb <= kRunMidTasks      ? kRunMid
                                                           : uint32_t(kBlock);
            const uint32_t bg = (ref_bound + run - 1) / run;
            const uint32_t tg = (tb + kBlock - 1) / kBlock;
            const uint32_t gg = std::min<uint32_t>((tb + kBlock / kBins - 1) / (kBlock / kBins),
                                                   kSplitGrid);  // task groups
            const uint32_t ccl = std::min(tb, kMaxCand) * 3 * kBins;
            const uint32_t cg = std::min<uint32_t>(
                std::max<uint32_t>((ccl + kBlock - 1) / kBlock, (tb + kBlock - 1) / kBlock), 4096);
            const uint32_t sg = (std::min(tb, kMaxCand) * kBins + kBlock - 1) / kBlock;
            const SLevel* L = levels + k;
            SLevel* Lw = levels + k;
            const uint32_t a = k & 1u, b = a ^ 1u;
            const SRef* rc = refs[a];
            const STask* tc = tasks[a];
            hipLaunchKernelGGL(ref_bounds_kernel, dim3(bg), dim3(kBlock), 0, stream, L, rc, tc,
                               tbox, run);
            hipLaunchKernelGGL(task_setup_kernel, dim3(tg), dim3(kBlock), 0, stream, L, tc, rc,
                               tbox, dec, root_half_area, k);
            hipLaunchKernelGGL(obj_bin_kernel, dim3(bg), dim3(kBlock), 0, stream, L, rc, tc, dec,
                               obins, run);
            hipLaunchKernelGGL(obj_split_kernel, dim3(gg), dim3(kBlock), 0, stream, L, tc, dec,
                               obins, rc, root_half_area, cand);
            if ((e = scan_exclusive_dev(ScanArrays{{cand}}, 1, ScanCounts{{&L->num_tasks}}, tb,
                                        ScanTotals{{&Lw->cand_total}}, scan_scratch, sz_scan,
                                        stream)) != hipSuccess)
                return fail(e, "spatial candidate scan");
            hipLaunchKernelGGL(cand_clear_kernel, dim3(cg), dim3(kBlock), 0, stream, L, cand,
                               cand_list, dup, sbins);
            hipLaunchKernelGGL(sp_bin_kernel, dim3(bg), dim3(kBlock), 0, stream, L, rc, dec, cand,
                               verts, n, sbins, run);
            hipLaunchKernelGGL(sp_split_kernel, dim3(sg), dim3(kBlock), 0, stream, L, cand_list,
                               tc, dec, sbins, dup);
            if ((e = scan_exclusive_dev(ScanArrays{{dup}}, 1, ScanCounts{{&L->num_tasks}}, tb,
                                        ScanTotals{{&Lw->dup_total}}, scan_scratch, sz_scan,
                                        stream)) != hipSuccess)
                return fail(e, "spatial budget scan");
            hipLaunchKernelGGL(decide_kernel, dim3(tg), dim3(kBlock), 0, stream, L, tc, dec, dup,
                               budget, leaf, obins, rc, split_flag, out_cnt, leaf_cnt, big_cnt);
            hipLaunchKernelGGL(side_kernel, dim3(rg), dim3(kBlock), 0, stream, L, rc, tc, dec, fl,
                               fr, fle, fre);
            // the task counts (tb of them at most) and the side flags (ref_bound)
            const ScanCounts cnt6{{&L->num_tasks, &L->num_tasks, &L->num_tasks, &L->num_tasks,
                                   &L->nrefs, &L->nrefs}};
            if ((e = scan_exclusive_dev(
                     ScanArrays{{split_flag, out_cnt, leaf_cnt, big_cnt, fle, fre}}, 6, cnt6,
                     std::max(tb, ref_bound),
                     ScanTotals{{&Lw->splits, &Lw->next_refs, &Lw->leaf_refs, &Lw->slots}},
                     scan_scratch, sz_scan, stream)) != hipSuccess)
                return fail(e, "spatial task and partition scans");
            hipLaunchKernelGGL(children_kernel, dim3(tg), dim3(kBlock), 0, stream, L, tc, dec,
                               split_flag, out_cnt, leaf_cnt, big_cnt, node_cap, cap, res.nodes,
                               tasks[b], fail_word);
            hipLaunchKernelGGL(scatter_kernel, dim3(rg), dim3(kBlock), 0, stream, L, rc, tc, dec,
                               split_flag, out_cnt, leaf_cnt, fl, fr, fle, fre, verts, n, cap,
                               refs[b], res.indices, fail_word, tbox, obins,
                               uint32_t(slot_cap));
            hipLaunchKernelGGL(level_end_kernel, dim3(1), dim3(64), 0, stream, levels, k, cap,
                               fail_word);
        }
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "spatial level");
        level = end;
        if ((e = hipMemcpyAsync(host_level, levels + level, sizeof(SLevel), hipMemcpyDeviceToHost,
                                stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
            return fail(e, "spatial levels");
        if (host_level->fail & 1u) {
            set_error("spatial builder ran out of reference or node space");
            return fail_msg();
        }
        if (host_level->fail) {
            set_error("spatial builder lost references");
            return fail_msg();
        }
        if (host_level->num_tasks == 0) break;
        if (level >= kMaxDepth + 2) {  // a task at depth kMaxDepth is a leaf
            set_error("spatial builder did not finish in %u levels", kMaxDepth + 1);
            return fail_msg();
        }
        known_tasks = host_level->num_tasks;
        known_refs = host_level->nrefs;
        known_dups = host_level->dups;
        batch = 8;
    }
    res.node_count = host_level->counter;
    res.prim_count = host_level->leaf_base;  // index entries: primitives and duplicates
    res.source_prims = n;
    res.root_count = 1;
    *out = res;
    return Ok;
}

}  // namespace rtbvh
