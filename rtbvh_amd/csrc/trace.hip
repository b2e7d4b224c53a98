// Closest-hit and any-hit traversal kernels for gfx950.
//
// One ray per lane in a persistent grid: each lane walks its own sequence of
// rays (grid-stride) and every loop iteration performs one node visit, so a
// lane whose ray finishes picks up its next ray on the following iteration
// instead of idling until the whole wave is done.  The traversal stack keeps
// its top kLdsStack entries in LDS (lane-interleaved, so entries of different
// lanes never share a bank) and the rest in a per-lane private array; a ray
// that would exceed RTBVH_MAX_STACK entries is terminated with t = NaN and
// the device status word records it.
//
// Visit order, box tests and the triangle test restate the reference:
//   MBVH: MbvhIterator::next (src/iter.rs:308-354), MbvhNode::intersect
//         (src/mbvh_node.rs:177-240);
//   BVH:  BvhIterator::next (src/iter.rs:88-127), Aabb::intersect
//         (src/aabb.rs:146-181), BvhNode::sort_nodes (src/bvh_node.rs:150-177);
//   triangle: SpatialTriangle::intersect (src/builders/spatial_sah.rs:132-163).
// REF = true reproduces the reference's visit order and first-accepted tie
// rule; REF = false returns the lexicographic minimum of (t, prim) and culls
// boxes that cannot hold such a hit.
#include "internal.hpp"

namespace rtbvh {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsStack = 24;
constexpr int kDeepStack = RTBVH_MAX_STACK - kLdsStack;

// The stack pointer is a separate local so that it stays in a register; the
// private overflow array is dynamically indexed and lives in scratch.
struct DeepStack {
    int32_t v[kDeepStack];
};

__device__ __forceinline__ bool stack_push(int32_t* lds, DeepStack& deep, int& sp, int32_t v) {
    if (sp < kLdsStack) {
        lds[sp * kBlock] = v;  // lane-interleaved: entry e of this lane at lds[e * kBlock]
    } else if (sp < RTBVH_MAX_STACK) {
        deep.v[sp - kLdsStack] = v;
    } else {
        return false;
    }
    ++sp;
    return true;
}

__device__ __forceinline__ int32_t stack_pop(const int32_t* lds, const DeepStack& deep, int& sp) {
    --sp;
    return sp < kLdsStack ? lds[sp * kBlock] : deep.v[sp - kLdsStack];
}

struct RayState {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz, t_min;
    float bt, bu, bv;
    uint32_t bp;
};

__device__ __forceinline__ bool load_ray(const DevRay* rays, uint64_t rid, RayState& s) {
    const float4* r4 = reinterpret_cast<const float4*>(rays + rid);
    const float4 a = r4[0];
    const float4 b = r4[1];
    s.ox = a.x;
    s.oy = a.y;
    s.oz = a.z;
    s.t_min = a.w;
    s.dx = b.x;
    s.dy = b.y;
    s.dz = b.z;
    s.bt = b.w;
    // Ray::new, src/ray.rs:166-182: inv = 1 / d (IEEE), t from the caller.
    s.ix = 1.0f / s.dx;
    s.iy = 1.0f / s.dy;
    s.iz = 1.0f / s.dz;
    s.bu = 0.f;
    s.bv = 0.f;
    s.bp = RTBVH_MISS;
    const bool nan = (s.ox != s.ox) || (s.oy != s.oy) || (s.oz != s.oz) || (s.dx != s.dx) ||
                     (s.dy != s.dy) || (s.dz != s.dz);
    return !nan;
}

// Tests the primitives indices[first .. first+count) against the ray.
// Returns true (any-hit) as soon as one is accepted.
template <bool REF, bool ANY>
__device__ __forceinline__ bool test_leaf(const TraceParams& p, int32_t first, int32_t count,
                                          RayState& s) {
    for (int32_t j = 0; j < count; ++j) {
        const uint32_t pos = uint32_t(first) + uint32_t(j);
        if (pos >= p.index_count) break;
        uint32_t prim;
        MtResult m;
        if (p.tris) {
            m = mt_prepared(p.tris, pos, s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, s.t_min, &prim);
            if (prim == RTBVH_MISS) continue;
        } else {
            prim = p.indices[pos];
            if (prim >= p.prim_count) continue;  // iter.rs:337 skips missing primitives
            const float* v = p.verts + size_t(prim) * 9;
            m = mt_candidate(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], s.ox, s.oy,
                             s.oz, s.dx, s.dy, s.dz, s.t_min);
        }
        if (!m.ok) continue;
        bool take;
        if (REF || ANY)
            take = m.t < s.bt;
        else
            take = lowest_first(m.t, prim, s.bt, s.bp);
        if (take) {
            s.bt = m.t;
            s.bu = m.u;
            s.bv = m.v;
            s.bp = prim;
            if (ANY) return true;
        }
    }
    return false;
}

template <bool ANY>
__device__ __forceinline__ void write_result(const TraceParams& p, uint64_t rid,
                                             const RayState& s, bool overflow) {
    if (ANY) {
        p.occluded[rid] = (s.bp != RTBVH_MISS) ? 1 : 0;
    } else {
        float4 h;
        h.x = overflow ? __builtin_nanf("") : s.bt;
        h.y = s.bu;
        h.z = s.bv;
        h.w = __uint_as_float(overflow ? RTBVH_MISS : s.bp);
        reinterpret_cast<float4*>(p.hits)[rid] = h;
    }
}

#define RTBVH_CSWAP(a, b)                   \
    if (tn[a] > tn[b]) {                    \
        float t_ = tn[a];                   \
        tn[a] = tn[b];                      \
        tn[b] = t_;                         \
        int32_t c_ = ch[a];                 \
        ch[a] = ch[b];                      \
        ch[b] = c_;                         \
        c_ = ct[a];                         \
        ct[a] = ct[b];                      \
        ct[b] = c_;                         \
        bool h_ = hit[a];                   \
        hit[a] = hit[b];                    \
        hit[b] = h_;                        \
    }

template <bool REF, bool ANY, bool STATS = false>
__global__ void __launch_bounds__(kBlock) mbvh_trace_kernel(TraceParams p) {
    __shared__ int32_t lds_stack[kLdsStack * kBlock];
    uint64_t n_nodes = 0, n_tris = 0;  // STATS only
    int32_t* lds = lds_stack + threadIdx.x;
    DeepStack deep;
    int sp = 0;
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    uint64_t rid = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    RayState s;
    int32_t cur = 0;
    bool active = false;
    for (;;) {
        if (!active) {
            if (rid >= p.ray_count) break;
            const bool ok = load_ray(p.rays, rid, s);
            if (!ok || p.node_count == 0) {
                write_result<ANY>(p, rid, s, false);
                rid += stride;
                continue;
            }
            active = true;
            cur = 0;
            sp = 0;
        }
        const DevMbvhNode nd = p.mnodes[cur];
        if (STATS) ++n_nodes;
        float tn[4], tf[4];
        int32_t ch[4], ct[4];
        bool hit[4];
        {
            const float mnx[4] = {nd.min_x.x, nd.min_x.y, nd.min_x.z, nd.min_x.w};
            const float mxx[4] = {nd.max_x.x, nd.max_x.y, nd.max_x.z, nd.max_x.w};
            const float mny[4] = {nd.min_y.x, nd.min_y.y, nd.min_y.z, nd.min_y.w};
            const float mxy[4] = {nd.max_y.x, nd.max_y.y, nd.max_y.z, nd.max_y.w};
            const float mnz[4] = {nd.min_z.x, nd.min_z.y, nd.min_z.z, nd.min_z.w};
            const float mxz[4] = {nd.max_z.x, nd.max_z.y, nd.max_z.z, nd.max_z.w};
            ch[0] = nd.children.x;
            ch[1] = nd.children.y;
            ch[2] = nd.children.z;
            ch[3] = nd.children.w;
            ct[0] = nd.counts.x;
            ct[1] = nd.counts.y;
            ct[2] = nd.counts.z;
            ct[3] = nd.counts.w;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float tx0 = (mnx[i] - s.ox) * s.ix;
                const float tx1 = (mxx[i] - s.ox) * s.ix;
                const float ty0 = (mny[i] - s.oy) * s.iy;
                const float ty1 = (mxy[i] - s.oy) * s.iy;
                const float tz0 = (mnz[i] - s.oz) * s.iz;
                const float tz1 = (mxz[i] - s.oz) * s.iz;
                const float a = ps_max(ps_min(tx0, tx1), ps_max(ps_min(ty0, ty1), ps_min(tz0, tz1)));
                const float b = ps_min(ps_max(tx0, tx1), ps_min(ps_max(ty0, ty1), ps_max(tz0, tz1)));
                tn[i] = a;
                tf[i] = b;
                if (REF)
                    hit[i] = (b >= a) && (a < s.bt);  // mbvh_node.rs:205
                else
                    hit[i] = (b >= a) && (a <= s.bt) && (b >= s.t_min);
            }
        }
        bool found = false;
        bool overflow = false;
        if (REF) {
            // mbvh_node.rs:219-237: the fifth comparator is (2,3), not (1,2).
            RTBVH_CSWAP(0, 1);
            RTBVH_CSWAP(2, 3);
            RTBVH_CSWAP(0, 2);
            RTBVH_CSWAP(1, 3);
            RTBVH_CSWAP(2, 3);
            // iter.rs:326: slots visited ids[3], ids[2], ids[1], ids[0]
#pragma unroll
            for (int k = 3; k >= 0; --k) {
                if (!hit[k] || found || overflow) continue;
                if (ct[k] > -1) {
                    if (STATS) n_tris += uint64_t(ct[k]);
                    found = test_leaf<REF, ANY>(p, ch[k], ct[k], s);
                } else if (ch[k] > -1) {
                    overflow = !stack_push(lds, deep, sp, ch[k]);
                }
            }
        } else {
            RTBVH_CSWAP(0, 1);
            RTBVH_CSWAP(2, 3);
            RTBVH_CSWAP(0, 2);
            RTBVH_CSWAP(1, 3);
            RTBVH_CSWAP(1, 2);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!hit[k] || found || ct[k] < 0 || !(tn[k] <= s.bt)) continue;
                if (STATS) n_tris += uint64_t(ct[k]);
                found = test_leaf<REF, ANY>(p, ch[k], ct[k], s);
            }
#pragma unroll
            for (int k = 3; k >= 0; --k) {
                if (!hit[k] || found || overflow || ct[k] >= 0 || ch[k] < 0) continue;
                overflow = !stack_push(lds, deep, sp, ch[k]);
            }
        }
        bool done = found || overflow;
        if (!done) {
            if (sp == 0) {
                done = true;
            } else {
                cur = stack_pop(lds, deep, sp);
                if (uint32_t(cur) >= p.node_count) done = true;  // malformed tree
            }
        }
        if (overflow) atomicOr(p.status, kStatusStackOverflow);
        if (done) {
            write_result<ANY>(p, rid, s, overflow);
            active = false;
            rid += stride;
        }
    }
    if (STATS) {
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats), (unsigned long long)n_nodes);
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats) + 1, (unsigned long long)n_tris);
    }
}

// Aabb::intersect, src/aabb.rs:146-181 (returns t_far; `near` receives the
// entry distance including the z slab, used only for culling in the
// lowest-id mode).
__device__ __forceinline__ bool bvh_slab(const DevAabb& b, const RayState& s, float* far_out,
                                         float* near_out) {
    const float lox = s.dx < 0.0f ? b.max[0] : b.min[0];
    const float hix = s.dx < 0.0f ? b.min[0] : b.max[0];
    const float loy = s.dy < 0.0f ? b.max[1] : b.min[1];
    const float hiy = s.dy < 0.0f ? b.min[1] : b.max[1];
    const float loz = s.dz < 0.0f ? b.max[2] : b.min[2];
    const float hiz = s.dz < 0.0f ? b.min[2] : b.max[2];
    float ray_min = (lox - s.ox) * s.ix;
    float ray_max = (hix - s.ox) * s.ix;
    const float y_min = (loy - s.oy) * s.iy;
    const float y_max = (hiy - s.oy) * s.iy;
    if ((ray_min > y_max) || (y_min > ray_max)) return false;
    if (y_min > ray_min) ray_min = y_min;
    if (y_max < ray_max) ray_max = y_max;
    const float z_min = (loz - s.oz) * s.iz;
    const float z_max = (hiz - s.oz) * s.iz;
    if ((ray_min > z_max) || (z_min > ray_max)) return false;
    if (z_max < ray_max) ray_max = z_max;
    *far_out = ray_max;
    *near_out = z_min > ray_min ? z_min : ray_min;
    return ray_max > s.t_min;
}

template <bool REF, bool ANY>
__global__ void __launch_bounds__(kBlock) bvh_trace_kernel(TraceParams p) {
    __shared__ int32_t lds_stack[kLdsStack * kBlock];
    int32_t* lds = lds_stack + threadIdx.x;
    DeepStack deep;
    int sp = 0;
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    uint64_t rid = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    RayState s;
    int32_t cur = 0;
    bool active = false;
    for (;;) {
        if (!active) {
            if (rid >= p.ray_count) break;
            const bool ok = load_ray(p.rays, rid, s);
            if (!ok || p.node_count == 0) {
                write_result<ANY>(p, rid, s, false);
                rid += stride;
                continue;
            }
            active = true;
            cur = 0;  // the root is visited without a box test (iter.rs:36-44)
            sp = 0;
        }
        const DevAabb nd = p.bnodes[cur];
        bool found = false;
        bool overflow = false;
        if (nd.extra1 > -1) {
            found = test_leaf<REF, ANY>(p, nd.extra2, nd.extra1, s);
        } else if (nd.extra2 > -1 && uint32_t(nd.extra2) + 1 < p.node_count) {
            const int32_t lf = nd.extra2;
            const DevAabb l = p.bnodes[lf];
            const DevAabb r = p.bnodes[lf + 1];
            float tl = 0.f, tr = 0.f, nl = 0.f, nr = 0.f;
            bool hl = bvh_slab(l, s, &tl, &nl);
            bool hr = bvh_slab(r, s, &tr, &nr);
            if (!REF) {
                hl = hl && nl <= s.bt;
                hr = hr && nr <= s.bt;
            }
            // sort_nodes, bvh_node.rs:150-177
            if (hl && hr) {
                if (tl < tr) {
                    overflow = !stack_push(lds, deep, sp, lf) || !stack_push(lds, deep, sp, lf + 1);
                } else {
                    overflow = !stack_push(lds, deep, sp, lf + 1) || !stack_push(lds, deep, sp, lf);
                }
            } else if (hl) {
                overflow = !stack_push(lds, deep, sp, lf);
            } else if (hr) {
                overflow = !stack_push(lds, deep, sp, lf + 1);
            }
        }
        bool done = found || overflow;
        if (!done) {
            if (sp == 0) {
                done = true;
            } else {
                cur = stack_pop(lds, deep, sp);
                if (uint32_t(cur) >= p.node_count) done = true;
            }
        }
        if (overflow) atomicOr(p.status, kStatusStackOverflow);
        if (done) {
            write_result<ANY>(p, rid, s, overflow);
            active = false;
            rid += stride;
        }
    }
}

template <typename K>
int persistent_grid(K kernel, uint64_t work_items) {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int& blocks = cached[dev];
    if (blocks == 0) {
        int cus = 0;
        int per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        blocks = cus * per_cu;
    }
    const uint64_t need = (work_items + kBlock - 1) / kBlock;
    return int(need < uint64_t(blocks) ? need : uint64_t(blocks));
}

template <bool REF, bool ANY, bool STATS = false>
hipError_t launch_mbvh(const TraceParams& p, hipStream_t stream) {
    auto k = mbvh_trace_kernel<REF, ANY, STATS>;
    const int grid = persistent_grid(k, p.ray_count);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, stream, p);
    return hipGetLastError();
}

template <bool REF, bool ANY>
hipError_t launch_bvh(const TraceParams& p, hipStream_t stream) {
    auto k = bvh_trace_kernel<REF, ANY>;
    const int grid = persistent_grid(k, p.ray_count);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, stream, p);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_trace(const TraceParams& p, bool mbvh, bool ref_order, bool any_hit,
                        hipStream_t stream) {
    if (mbvh && !ref_order) return launch_trace_fast(p, any_hit, stream);
    if (p.stats) {
        if (!mbvh || any_hit) return hipErrorInvalidValue;
        return ref_order ? launch_mbvh<true, false, true>(p, stream)
                         : launch_mbvh<false, false, true>(p, stream);
    }
    if (mbvh) {
        if (ref_order)
            return any_hit ? launch_mbvh<true, true>(p, stream) : launch_mbvh<true, false>(p, stream);
        return any_hit ? launch_mbvh<false, true>(p, stream) : launch_mbvh<false, false>(p, stream);
    }
    if (ref_order)
        return any_hit ? launch_bvh<true, true>(p, stream) : launch_bvh<true, false>(p, stream);
    return any_hit ? launch_bvh<false, true>(p, stream) : launch_bvh<false, false>(p, stream);
}

}  // namespace rtbvh
