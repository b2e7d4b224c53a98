// Closest-hit / any-hit traversal for the default tie rule (the
// lexicographic minimum of (t, prim id) — identical to brute force), tuned
// for gfx950: the 4-wide kernels (reference-layout MBVH and 4-wide
// traversal tree) and, further down, the compressed 8-wide tree's kernel.
// The reference-order kernel in trace.hip reproduces MbvhIterator exactly;
// these keep its triangle test but are free to choose the visit order,
// because the lowest (t, prim) over all triangles does not depend on it.
// The 4-wide kernels keep its box arithmetic:
//
//   * slab test with IEEE min/max (v_min3/v_max3); along a zero direction
//     component the planes are measured from the floats next to the origin
//     (slab_origin), so that a ray inside a face plane touches the box;
//   * a slot is visited when t_entry <= t_best and t_exit >= t_min;
//   * slots are ordered by a 32-bit key (entry distance bits with the slot
//     id in the two low bits), sorted with five integer min/max pairs;
//   * leaves are tested near to far, each re-checked against the current t;
//   * the nearest interior child is visited next without a stack round trip;
//     the others are pushed far to near together with their entry distance,
//     and a popped entry whose distance now exceeds t_best is dropped without
//     fetching its node.
// STATS variants count node visits and triangle tests (the bytes-per-ray
// figure of the benchmark).
#include "internal.hpp"

namespace rtbvh {

namespace {

// A hit record (t, u, v, primitive bits) written with a non-temporal store:
// each record is written once and never read by the launch, so it should not
// take L2 lines from the nodes and triangles (0.7 % faster bounce traces at
// 1M in an A/B on one box).
typedef float HitVec __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_hit(DevHit* hits, uint64_t rid, float t, float u, float v,
                                          uint32_t prim) {
    HitVec h;
    h.x = t;
    h.y = u;
    h.z = v;
    h.w = __uint_as_float(prim);
    __builtin_nontemporal_store(h, reinterpret_cast<HitVec*>(hits) + rid);
}

constexpr int kMaxBlock = 256;

// The stack's first `depth` entries live in LDS (entry k of lane l at
// lds[k * stride + l], stride = block size: consecutive lanes, consecutive
// banks), the rest in per-lane private arrays.  The stack pointer is kept
// apart from those arrays so that it stays in a register (a struct holding
// both would be placed in scratch as a whole).
struct DeepStack {
    int32_t i[RTBVH_MAX_STACK];
    float t[RTBVH_MAX_STACK];
};

struct LdsStack {
    int32_t* i;
    float* t;
    int depth;
    int stride;
};

__device__ __forceinline__ bool stack_push(const LdsStack& ls, DeepStack& deep, int& sp,
                                           int32_t node, float t) {
    if (sp < ls.depth) {
        ls.i[sp * ls.stride] = node;
        ls.t[sp * ls.stride] = t;
    } else if (sp < RTBVH_MAX_STACK) {
        deep.i[sp - ls.depth] = node;
        deep.t[sp - ls.depth] = t;
    } else {
        return false;
    }
    ++sp;
    return true;
}

__device__ __forceinline__ void stack_pop(const LdsStack& ls, const DeepStack& deep, int& sp,
                                          int32_t& node, float& t) {
    --sp;
    if (sp < ls.depth) {
        node = ls.i[sp * ls.stride];
        t = ls.t[sp * ls.stride];
    } else {
        node = deep.i[sp - ls.depth];
        t = deep.t[sp - ls.depth];
    }
}

// Origins of the slab tests of the lowest-id 4-wide kernels, per axis: the
// ray's own, except where 1/d is infinite (a zero or tiny direction
// component), where the low planes are measured from the next float above it
// and the high planes from the next float below.  A ray that runs inside a box's face
// plane then gets -inf / +inf for that slab (it touches the box) instead of
// 0 * inf = NaN, which min/max would turn into a closed slab and miss hits on
// that face that brute force finds; a ray one float or more outside still
// gets an empty slab.  Nonzero components keep the exact arithmetic of
// MbvhNode::intersect (src/mbvh_node.rs:177-240).
__device__ __forceinline__ void slab_origin(float o, float inv, float& lo, float& hi) {
    const bool flat = fabsf(inv) > 3.4028235e38f;
    lo = flat ? nextafterf(o, 3.4028235e38f) : o;
    hi = flat ? nextafterf(o, -3.4028235e38f) : o;
}

template <typename T>
__device__ __forceinline__ T sel4(uint32_t k, T a, T b, T c, T d) {
    return k == 0 ? a : (k == 1 ? b : (k == 2 ? c : d));
}

__device__ __forceinline__ void csort(uint32_t& a, uint32_t& b) {
    const uint32_t lo = a < b ? a : b;
    const uint32_t hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Scene of ray rid in a batch: the s with off[s] <= rid < off[s+1].
__device__ __forceinline__ uint32_t scene_of_ray(const uint32_t* off, uint32_t scenes,
                                                 uint32_t rid) {
    uint32_t lo = 0, hi = scenes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= rid)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// BATCH: the MBVH is a forest whose node s is the root of scene s, and each
// ray starts at the root of its scene.
template <bool ANY, bool STATS, bool PREP, bool BATCH>
__global__ void __launch_bounds__(kMaxBlock) mbvh_fast_kernel(TraceParams p) {
    if (p.gate && *p.gate) return;  // the tree this launch would read is incomplete
    extern __shared__ int32_t s_stack[];  // p.lds_depth entries per lane: ids, then distances
    const LdsStack ls{s_stack + threadIdx.x,
                      reinterpret_cast<float*>(s_stack + p.lds_depth * blockDim.x) + threadIdx.x,
                      int(p.lds_depth), int(blockDim.x)};
    DeepStack deep;
    int sp = 0;
    uint64_t n_nodes = 0, n_tris = 0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    uint64_t rid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float ox = 0.f, oy = 0.f, oz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, ix = 0.f, iy = 0.f,
          iz = 0.f, tmin = 0.f;
    float lox = 0.f, loy = 0.f, loz = 0.f, hox = 0.f, hoy = 0.f, hoz = 0.f;  // slab origins
    float bt = 0.f, bu = 0.f, bv = 0.f;
    uint32_t bp = RTBVH_MISS;
    int32_t cur = -1;
    bool active = false;
    for (;;) {
        if (!active) {
            if (rid >= p.ray_count) break;
            const float4* r4 = reinterpret_cast<const float4*>(p.rays + rid);
            const float4 ra = r4[0];
            const float4 rb = r4[1];
            ox = ra.x;
            oy = ra.y;
            oz = ra.z;
            tmin = ra.w;
            dx = rb.x;
            dy = rb.y;
            dz = rb.z;
            bt = rb.w;
            ix = 1.0f / dx;  // Ray::new, src/ray.rs:166-182
            iy = 1.0f / dy;
            iz = 1.0f / dz;
            slab_origin(ox, ix, lox, hox);
            slab_origin(oy, iy, loy, hoy);
            slab_origin(oz, iz, loz, hoz);
            bu = 0.f;
            bv = 0.f;
            bp = RTBVH_MISS;
            const bool nan = (ox != ox) || (oy != oy) || (oz != oz) || (dx != dx) ||
                             (dy != dy) || (dz != dz);
            if (nan || p.node_count == 0) {
                if (ANY) {
                    p.occluded[rid] = 0;
                } else {
                    store_hit(p.hits, rid, bt, 0.f, 0.f, RTBVH_MISS);
                }
                rid += stride;
                continue;
            }
            active = true;
            cur = BATCH ? int32_t(scene_of_ray(p.ray_offsets, p.scene_count, uint32_t(rid))) : 0;
            sp = 0;
        }
        bool done = false;
        bool overflow = false;
        bool found = false;
        if (cur < 0) {
            // pop until an entry can still hold a hit at or before bt
            while (sp > 0) {
                int32_t c;
                float t;
                stack_pop(ls, deep, sp, c, t);
                if (t <= bt) {
                    cur = c;
                    break;
                }
            }
            if (cur < 0 || uint32_t(cur) >= p.node_count) done = true;
        }
        if (!done) {
            const DevMbvhNode nd = p.mnodes[cur];
            if (STATS) ++n_nodes;
            const float mnx[4] = {nd.min_x.x, nd.min_x.y, nd.min_x.z, nd.min_x.w};
            const float mxx[4] = {nd.max_x.x, nd.max_x.y, nd.max_x.z, nd.max_x.w};
            const float mny[4] = {nd.min_y.x, nd.min_y.y, nd.min_y.z, nd.min_y.w};
            const float mxy[4] = {nd.max_y.x, nd.max_y.y, nd.max_y.z, nd.max_y.w};
            const float mnz[4] = {nd.min_z.x, nd.min_z.y, nd.min_z.z, nd.min_z.w};
            const float mxz[4] = {nd.max_z.x, nd.max_z.y, nd.max_z.z, nd.max_z.w};
            float tn[4];
            uint32_t key[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float tx0 = (mnx[i] - lox) * ix, tx1 = (mxx[i] - hox) * ix;
                const float ty0 = (mny[i] - loy) * iy, ty1 = (mxy[i] - hoy) * iy;
                const float tz0 = (mnz[i] - loz) * iz, tz1 = (mxz[i] - hoz) * iz;
                const float a = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
                const float b = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
                tn[i] = a;
                const bool h = (b >= a) && (a <= bt) && (b >= tmin);
                key[i] = h ? ((__float_as_uint(a > 0.0f ? a : 0.0f) & ~3u) | uint32_t(i))
                           : 0xFFFFFFFFu;
            }
            csort(key[0], key[1]);
            csort(key[2], key[3]);
            csort(key[0], key[2]);
            csort(key[1], key[3]);
            csort(key[1], key[2]);
            const int32_t c0 = nd.children.x, c1 = nd.children.y, c2 = nd.children.z,
                          c3 = nd.children.w;
            const int32_t n0 = nd.counts.x, n1 = nd.counts.y, n2 = nd.counts.z, n3 = nd.counts.w;
            // leaves, near to far
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (key[k] == 0xFFFFFFFFu || found) continue;
                const uint32_t s = key[k] & 3u;
                const int32_t cnt = sel4(s, n0, n1, n2, n3);
                if (cnt < 0) continue;
                if (!(sel4(s, tn[0], tn[1], tn[2], tn[3]) <= bt)) continue;
                const int32_t first = sel4(s, c0, c1, c2, c3);
                if (STATS) n_tris += uint64_t(cnt);
                for (int32_t j = 0; j < cnt; ++j) {
                    const uint32_t pos = uint32_t(first) + uint32_t(j);
                    if (pos >= p.index_count) break;
                    uint32_t prim;
                    MtResult m;
                    if (PREP) {
                        m = mt_prepared(p.tris, pos, ox, oy, oz, dx, dy, dz, tmin, &prim);
                        if (prim == RTBVH_MISS) continue;
                    } else {
                        prim = p.indices[pos];
                        if (prim >= p.prim_count) continue;
                        const float* v = p.verts + size_t(prim) * 9;
                        m = mt_candidate(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], ox,
                                         oy, oz, dx, dy, dz, tmin);
                    }
                    if (!m.ok) continue;
                    if (ANY) {
                        if (m.t < bt) {
                            bp = prim;
                            found = true;
                            break;
                        }
                    } else if (lowest_first(m.t, prim, bt, bp)) {
                        bt = m.t;
                        bu = m.u;
                        bv = m.v;
                        bp = prim;
                    }
                }
            }
            // interior children, far to near; the nearest becomes the next node
            int32_t next = -1;
            float next_t = 0.f;
#pragma unroll
            for (int k = 3; k >= 0; --k) {
                if (key[k] == 0xFFFFFFFFu || found || overflow) continue;
                const uint32_t s = key[k] & 3u;
                if (sel4(s, n0, n1, n2, n3) >= 0) continue;
                const int32_t c = sel4(s, c0, c1, c2, c3);
                const float t = sel4(s, tn[0], tn[1], tn[2], tn[3]);
                if (c < 0 || !(t <= bt)) continue;
                if (next >= 0) overflow = !stack_push(ls, deep, sp, next, next_t);
                next = c;
                next_t = t;
            }
            cur = next;
            if (found || overflow) done = true;
        }
        if (overflow) atomicOr(p.status, kStatusStackOverflow);
        if (done) {
            if (ANY) {
                p.occluded[rid] = (bp != RTBVH_MISS) ? 1 : 0;
            } else {
                store_hit(p.hits, rid, overflow ? __builtin_nanf("") : bt, bu, bv,
                          overflow ? RTBVH_MISS : bp);
            }
            active = false;
            cur = -1;
            rid += stride;
        }
    }
    if (STATS) {
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats), (unsigned long long)n_nodes);
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats) + 1, (unsigned long long)n_tris);
    }
}

// ---------------------------------------------------------------------------
// One item per lane per iteration.  The kernel above tests a node's leaf
// slots inside the node step, so every iteration of a wave pays for up to
// four triangle tests and their loops whenever any lane has a leaf there
// (most of its issue slots then run with few lanes active).  Here a lane
// holds one item — a node, or a leaf's remaining triangles — and each
// iteration does at most one node visit and one triangle test per lane:
// the node step makes the nearest child slot (node or leaf) the item in
// hand and pushes the others, far to near, with their entry distance; the
// triangle step tests one triangle of the leaf in hand.  Leaves travel on
// the stack as tagged entries: bit 31 set, bits 27..30 count - 1, bits 0..26
// the first triangle, so this kernel needs fewer than 2^27 index entries and
// leaves of at most 16 triangles (the launcher checks both).
//
// Lanes whose ray is done take new rays together, once at least p.refill of
// them are idle (or no lane has work left): loading a ray and its three IEEE
// reciprocals then runs for many lanes at once instead of on almost every
// iteration for one.  Culling and the lowest-(t, id) rule are those of the
// kernel above.
constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kLeafFirstMask = (1u << 27) - 1u;

__device__ __forceinline__ int32_t sel4i(uint32_t s, int32_t a, int32_t b, int32_t c, int32_t d) {
    const int32_t lo = (s & 1u) ? b : a;
    const int32_t hi = (s & 1u) ? d : c;
    return (s & 2u) ? hi : lo;
}

__device__ __forceinline__ float sel4f(uint32_t s, float a, float b, float c, float d) {
    const float lo = (s & 1u) ? b : a;
    const float hi = (s & 1u) ? d : c;
    return (s & 2u) ? hi : lo;
}

// Stack entry of slot s: the child node, or the tagged leaf.
__device__ __forceinline__ int32_t slot_entry(int32_t c, int32_t n) {
    return n < 0 ? c : int32_t(kLeafFlag | (uint32_t(n - 1) << 27) | uint32_t(c));
}

// Ray hand-out of the persistent kernels.  Static (schedule 0): lane k of
// the grid takes rays k, k + stride, ...  Wave ranges (schedule 1): wave w
// owns the contiguous rays [w L, (w + 1) L), L = ceil(rays / waves), and its
// idle lanes draw the next ones in lane order as they free up, so a wave's
// lanes stay busy until its range is used up (no atomics: the cursor is
// wave-uniform).
// A value known to be equal in every lane of the wave, read from the first
// active lane so that the compiler keeps it in scalar registers.
__device__ __forceinline__ uint64_t wave_uniform(uint64_t v) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v))));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32))));
    return (uint64_t(hi) << 32) | lo;
}

struct RaySchedule {
    uint64_t rid;
    uint64_t stride;
    uint64_t cursor;  // wave ranges: next ray of the wave's range (wave-uniform)
    uint64_t end;
    bool ranges;

    __device__ void init(const TraceParams& p) {
        stride = uint64_t(gridDim.x) * blockDim.x;
        const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
        rid = tid;
        ranges = p.wave_ranges != 0;
        const uint64_t waves = (stride + 63) >> 6;
        const uint64_t per = (p.ray_count + waves - 1) / waves;
        const uint64_t b = (tid >> 6) * per;
        // the same for every lane of the wave (blocks are whole waves): held
        // in scalar registers
        cursor = wave_uniform(b < p.ray_count ? b : p.ray_count);
        end = wave_uniform(b + per < p.ray_count ? b + per : p.ray_count);
    }

    // This lane is idle and rays may be left for it.
    __device__ bool can_take(const TraceParams& p, bool active) const {
        return !active && (ranges ? cursor < end : rid < p.ray_count);
    }

    // Called by the whole wave when it hands out rays (takers = ballot of
    // can_take); true when this lane got ray `rid`.
    __device__ bool take(bool me, uint64_t takers) {
        if (!ranges) return me;
        // takers in the lanes below this one (v_mbcnt)
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
            uint32_t(takers >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(takers), 0u));
        const uint64_t r = cursor + below;
        cursor += uint64_t(__popcll(takers));
        if (cursor > end) cursor = end;
        if (!me || r >= end) return false;
        rid = r;
        return true;
    }

    // The lane's ray is finished (or skipped).
    __device__ void next() {
        if (!ranges) rid += stride;
    }
};

template <bool ANY, bool STATS, bool BATCH>
__global__ void __launch_bounds__(kMaxBlock) mbvh_item_kernel(TraceParams p) {
    if (p.gate && *p.gate) return;  // the tree this launch would read is incomplete
    extern __shared__ int32_t s_stack[];
    const LdsStack ls{s_stack + threadIdx.x,
                      reinterpret_cast<float*>(s_stack + p.lds_depth * blockDim.x) + threadIdx.x,
                      int(p.lds_depth), int(blockDim.x)};
    DeepStack deep;
    int sp = 0;
    uint64_t n_nodes = 0, n_tris = 0;
    RaySchedule sched;
    sched.init(p);
    const uint32_t refill = p.refill ? p.refill : 1u;
    float ox = 0.f, oy = 0.f, oz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, ix = 0.f, iy = 0.f,
          iz = 0.f, tmin = 0.f;
    float lox = 0.f, loy = 0.f, loz = 0.f, hox = 0.f, hoy = 0.f, hoz = 0.f;  // slab origins
    float bt = 0.f, bu = 0.f, bv = 0.f;
    uint32_t bp = RTBVH_MISS;
    int32_t node = -1;     // node in hand, or -1
    uint32_t lfirst = 0;   // triangles in hand: [lfirst, lfirst + lcount)
    int32_t lcount = 0;
    bool active = false;
    for (;;) {
        // ---- refill: idle lanes that still have rays take them together
        const bool can_take = sched.can_take(p, active);
        const uint64_t takers = __ballot(can_take);
        const uint64_t busy = __ballot(active);
        if (takers == 0 && busy == 0) break;  // wave-uniform: no ray left for any lane
        bool take = false;
        if (takers != 0 && (uint32_t(__popcll(takers)) >= refill || busy == 0))
            take = sched.take(can_take, takers);
        const uint64_t rid = sched.rid;
        if (take) {
            const float4* r4 = reinterpret_cast<const float4*>(p.rays + rid);
            const float4 ra = r4[0];
            const float4 rb = r4[1];
            ox = ra.x;
            oy = ra.y;
            oz = ra.z;
            tmin = ra.w;
            dx = rb.x;
            dy = rb.y;
            dz = rb.z;
            bt = rb.w;
            ix = 1.0f / dx;  // Ray::new, src/ray.rs:166-182
            iy = 1.0f / dy;
            iz = 1.0f / dz;
            slab_origin(ox, ix, lox, hox);
            slab_origin(oy, iy, loy, hoy);
            slab_origin(oz, iz, loz, hoz);
            bu = 0.f;
            bv = 0.f;
            bp = RTBVH_MISS;
            const bool nan = (ox != ox) || (oy != oy) || (oz != oz) || (dx != dx) ||
                             (dy != dy) || (dz != dz);
            if (nan || p.node_count == 0) {
                if (ANY) {
                    p.occluded[rid] = 0;
                } else {
                    store_hit(p.hits, rid, bt, 0.f, 0.f, RTBVH_MISS);
                }
                sched.next();
            } else {
                active = true;
                node = BATCH ? int32_t(scene_of_ray(p.ray_offsets, p.scene_count,
                                                    uint32_t(rid)))
                             : 0;
                lcount = 0;
                sp = 0;
            }
        }
        if (!active) continue;
        bool overflow = false;
        bool found = false;
        // ---- nothing in hand: the next entry that can still hold a hit at or before bt
        if (node < 0 && lcount == 0) {
            while (sp > 0) {
                int32_t e;
                float t;
                stack_pop(ls, deep, sp, e, t);
                if (!(t <= bt)) continue;
                if (e >= 0) {
                    node = e;
                } else {
                    lfirst = uint32_t(e) & kLeafFirstMask;
                    lcount = int32_t((uint32_t(e) >> 27) & 15u) + 1;
                }
                break;
            }
        }
        // ---- node step
        if (node >= 0) {
            const uint32_t cur = uint32_t(node);
            node = -1;
            if (cur < p.node_count) {
                const DevMbvhNode nd = p.mnodes[cur];
                if (STATS) ++n_nodes;
                const float mnx[4] = {nd.min_x.x, nd.min_x.y, nd.min_x.z, nd.min_x.w};
                const float mxx[4] = {nd.max_x.x, nd.max_x.y, nd.max_x.z, nd.max_x.w};
                const float mny[4] = {nd.min_y.x, nd.min_y.y, nd.min_y.z, nd.min_y.w};
                const float mxy[4] = {nd.max_y.x, nd.max_y.y, nd.max_y.z, nd.max_y.w};
                const float mnz[4] = {nd.min_z.x, nd.min_z.y, nd.min_z.z, nd.min_z.w};
                const float mxz[4] = {nd.max_z.x, nd.max_z.y, nd.max_z.z, nd.max_z.w};
                const int32_t ch[4] = {nd.children.x, nd.children.y, nd.children.z,
                                       nd.children.w};
                const int32_t ct[4] = {nd.counts.x, nd.counts.y, nd.counts.z, nd.counts.w};
                float tn[4];
                uint32_t key[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float tx0 = (mnx[i] - lox) * ix, tx1 = (mxx[i] - hox) * ix;
                    const float ty0 = (mny[i] - loy) * iy, ty1 = (mxy[i] - hoy) * iy;
                    const float tz0 = (mnz[i] - loz) * iz, tz1 = (mxz[i] - hoz) * iz;
                    const float a =
                        fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
                    const float b =
                        fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
                    tn[i] = a;
                    // a slot holds an interior child or a non-empty leaf
                    const bool used = (ct[i] < 0 && ch[i] >= 0) || ct[i] > 0;
                    const bool h = used && (b >= a) && (a <= bt) && (b >= tmin);
                    key[i] = h ? ((__float_as_uint(a > 0.0f ? a : 0.0f) & ~3u) | uint32_t(i))
                               : 0xFFFFFFFFu;
                }
                csort(key[0], key[1]);
                csort(key[2], key[3]);
                csort(key[0], key[2]);
                csort(key[1], key[3]);
                csort(key[1], key[2]);
                // key[0] (the nearest) goes in hand; key[nv-1] .. key[1] are pushed,
                // farthest first, so the nearer ones pop first
                const int nv = (key[0] != 0xFFFFFFFFu) + (key[1] != 0xFFFFFFFFu) +
                               (key[2] != 0xFFFFFFFFu) + (key[3] != 0xFFFFFFFFu);
                int32_t ent[3];
                float ent_t[3];
#pragma unroll
                for (int j = 1; j < 4; ++j) {
                    const uint32_t sl = key[j] & 3u;
                    ent[j - 1] = slot_entry(sel4i(sl, ch[0], ch[1], ch[2], ch[3]),
                                            sel4i(sl, ct[0], ct[1], ct[2], ct[3]));
                    ent_t[j - 1] = sel4f(sl, tn[0], tn[1], tn[2], tn[3]);
                }
                const int np = nv > 1 ? nv - 1 : 0;
                if (sp + np <= ls.depth) {
                    // all pushes land in LDS: entry k of this node goes to sp + k,
                    // holding key[np - k] (farthest at the bottom)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        if (k < np) {
                            const int j = np - 1 - k;  // ent index of key[np - k]
                            const int32_t e = j == 0 ? ent[0] : (j == 1 ? ent[1] : ent[2]);
                            const float t = j == 0 ? ent_t[0] : (j == 1 ? ent_t[1] : ent_t[2]);
                            ls.i[(sp + k) * ls.stride] = e;
                            ls.t[(sp + k) * ls.stride] = t;
                        }
                    }
                    sp += np;
                } else {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        if (k < np && !overflow) {
                            const int j = np - 1 - k;
                            const int32_t e = j == 0 ? ent[0] : (j == 1 ? ent[1] : ent[2]);
                            const float t = j == 0 ? ent_t[0] : (j == 1 ? ent_t[1] : ent_t[2]);
                            overflow = !stack_push(ls, deep, sp, e, t);
                        }
                    }
                }
                if (nv > 0) {
                    const uint32_t sl = key[0] & 3u;
                    const int32_t c = sel4i(sl, ch[0], ch[1], ch[2], ch[3]);
                    const int32_t n = sel4i(sl, ct[0], ct[1], ct[2], ct[3]);
                    if (n < 0) {
                        node = c;
                    } else {
                        lfirst = uint32_t(c);
                        lcount = n;
                    }
                }
            }
        }
        // ---- triangle step: one triangle of the leaf in hand
        if (lcount > 0) {
            const uint32_t pos = lfirst;
            ++lfirst;
            --lcount;
            if (pos < p.index_count) {
                if (STATS) ++n_tris;
                uint32_t prim;
                const MtResult m = mt_prepared(p.tris, pos, ox, oy, oz, dx, dy, dz, tmin, &prim);
                if (prim != RTBVH_MISS && m.ok) {
                    if (ANY) {
                        if (m.t < bt) {
                            bp = prim;
                            found = true;
                        }
                    } else if (lowest_first(m.t, prim, bt, bp)) {
                        bt = m.t;
                        bu = m.u;
                        bv = m.v;
                        bp = prim;
                    }
                }
            }
        }
        if (overflow) atomicOr(p.status, kStatusStackOverflow);
        if (found || overflow || (node < 0 && lcount == 0 && sp == 0)) {
            if (ANY) {
                p.occluded[rid] = (bp != RTBVH_MISS) ? 1 : 0;
            } else {
                store_hit(p.hits, rid, overflow ? __builtin_nanf("") : bt, bu, bv,
                          overflow ? RTBVH_MISS : bp);
            }
            active = false;
            node = -1;
            lcount = 0;
            sched.next();
        }
    }
    if (STATS) {
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats), (unsigned long long)n_nodes);
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats) + 1, (unsigned long long)n_tris);
    }
}

// ---- compressed 8-wide tree (DevCw8Node) ----------------------------------
//
// The stack holds groups rather than single nodes:
//   node group      y >= 2^24: bits 24..31 the node's interior hits not yet
//                   visited, by key (slot ^ ray octant); bits 16..23 its
//                   interior-slot mask; bits 0..15 a lower bound on their
//                   entry distances (a group beyond t_best is dropped
//                   unvisited); x its first child;
//   triangle group  y < 2^24: bits 0..7 the triangle hits not yet tested, by
//                   slot; bits 8..15 the triangle-slot mask; x the node's
//                   first triangle.
// Each iteration visits at most one node (the lowest key of the node group
// in hand, which pushes the rest of that group) and tests at most one
// triangle; with `postpone` set, the triangle test waits until that many
// lanes of the wave have one or a lane has no node group left to visit
// (a newer triangle group then pushes the waiting one).  A slot is entered
// when max(t_near, t_min) <= min(t_far, t_best), the near planes computed
// from the grid origin's distance b lowered and the far planes from b raised
// by e = 2^-20 (|b| + 255 |s|), s the grid step over d: twice the rounding of
// s, b (1-ulp reciprocal, two products) and the plane fma, so that no
// computed near plane lies beyond the exact one and no far plane before it;
// this never rejects a box that the exact slab test would enter.  Along an
// axis whose 1/d is infinite (a zero direction component) the plane
// distances would be infinite or NaN (0 * inf), and a NaN dropping out of
// min/max would leave that axis unculled; such an axis uses the distance
// H (q - u) instead, u the ray's origin in grid steps and H = 2^118: a slot
// whose planes lie a step or more beyond the origin is out of reach
// (entry >= H, beyond any t_max up to 3e35), one that holds the origin is
// open along that axis (entry <= -H, exit >= H).  The spare grid step makes
// every decision within a step of a plane safe either way.  The offset -H u
// is clamped to +-2^126, so that H q - H u and the slack stay finite.  Which
// triangles are tested can therefore only grow, and the lowest (t, id) is
// that of brute force.
constexpr uint32_t kNodeGroup = 1u << 24;
constexpr float kBoxEps = 0x1p-20f;
constexpr float kFlatScale = 0x1p118f;

// -H u of a flat axis, clamped so that H q - H u (q <= 255) stays finite.
__device__ __forceinline__ float flat_offset(float u) {
    return fminf(fmaxf(-kFlatScale * u, -0x1p126f), 0x1p126f);
}

// bit s of an 8-bit mask moves to bit s ^ oct
__device__ __forceinline__ uint32_t oct_permute(uint32_t m, uint32_t oct) {
    const uint32_t a = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
    m = (oct & 1u) ? a : m;
    const uint32_t b = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
    m = (oct & 2u) ? b : m;
    const uint32_t c = ((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu);
    return (oct & 4u) ? c : m;
}

// The LDS part of the group stack is accessed with explicit DS instructions,
// so that the compiler cannot fold them with the private-array accesses into
// generic (flat) ones.  addr: LDS byte address, 8-byte aligned.
__device__ __forceinline__ void lds_put(uint32_t addr, uint32_t x, uint32_t y) {
    const uint64_t v = (uint64_t(y) << 32) | x;
    asm volatile("ds_write_b64 %0, %1" : : "v"(addr), "v"(v));
}

__device__ __forceinline__ uint2 lds_get(uint32_t addr) {
    uint64_t v;
    asm volatile("ds_read_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(addr));
    return make_uint2(uint32_t(v), uint32_t(v >> 32));
}

// The low 16 bits of a node group: the upper half of the nearest entry
// distance among its children (non-negative, so the truncation rounds down
// and the bound stays a lower bound).
__device__ __forceinline__ uint32_t bound_bits(float t) {
    return __float_as_uint(fmaxf(t, 0.0f)) >> 16;
}

__device__ __forceinline__ float group_bound(uint32_t y) { return __uint_as_float(y << 16); }

// 1/d for the compressed tree's box tests: the hardware reciprocal (within
// 1 ulp, which the 2^-20 widening covers).  A zero component gives an
// infinity of its sign; a component beyond 2^126 gives 0 or a denormal, so
// its plane distances are 0 and only t_min <= 0 <= t_best decides, which
// enters at least the slots the exact distances would.
__device__ __forceinline__ float box_inv(float d) { return __builtin_amdgcn_rcpf(d); }

__device__ __forceinline__ float qbyte(uint32_t w, int i) {
    return float((w >> ((i & 3) * 8)) & 0xFFu);
}

template <bool ANY, bool STATS, bool BATCH, bool IDS = false>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(8)))
cw8_kernel(TraceParams p) {
    if (p.gate && *p.gate) return;  // the tree this launch would read is incomplete
    extern __shared__ int32_t s_stack[];  // p.lds_depth groups (8 B) per lane
    // LDS byte address of this lane's first entry (the low half of a shared
    // pointer) and the step between its entries
    const uint32_t lbase = uint32_t(reinterpret_cast<uintptr_t>(s_stack)) + threadIdx.x * 8u;
    const uint32_t lstep = blockDim.x * 8u;
    const int ldepth = int(p.lds_depth);
    uint2 deep[RTBVH_MAX_STACK];
    int sp = 0;
    uint64_t n_nodes = 0, n_tris = 0;
    RaySchedule sched;
    sched.init(p);
    const uint32_t refill = p.refill ? p.refill : 1u;
    float ox = 0.f, oy = 0.f, oz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, ix = 0.f, iy = 0.f,
          iz = 0.f, tmin = 0.f;
    float bt = 0.f, bu = 0.f, bv = 0.f;
    uint32_t bp = RTBVH_MISS;
    uint32_t oct = 0;           // bits 0..2 direction signs, 3..5 axes whose 1/d is infinite
    uint32_t ngx = 0, ngy = 0;  // node group in hand
    uint32_t tgx = 0, tgy = 0;  // triangle group in hand
    bool active = false;
    uint32_t cur = 0;  // IDS: the caller's index of the ray in hand
    const auto push = [&](uint32_t x, uint32_t y) -> bool {
        if (sp < ldepth) {
            lds_put(lbase + uint32_t(sp) * lstep, x, y);
        } else if (sp < RTBVH_MAX_STACK) {
            deep[sp - ldepth] = make_uint2(x, y);
        } else {
            return false;
        }
        ++sp;
        return true;
    };
    for (;;) {
        // ---- refill: idle lanes that still have rays take them together
        const bool can_take = sched.can_take(p, active);
        const uint64_t takers = __ballot(can_take);
        const uint64_t busy = __ballot(active);
        if (takers == 0 && busy == 0) break;  // wave-uniform: no ray left for any lane
        bool take = false;
        if (takers != 0 && (uint32_t(__popcll(takers)) >= refill || busy == 0))
            take = sched.take(can_take, takers);
        const uint64_t rid = sched.rid;
        if (take) {
            if (IDS) cur = p.ray_ids[rid];
            const float4* r4 = reinterpret_cast<const float4*>(p.rays + (IDS ? cur : rid));
            const float4 ra = r4[0];
            const float4 rb = r4[1];
            ox = ra.x;
            oy = ra.y;
            oz = ra.z;
            tmin = ra.w;
            dx = rb.x;
            dy = rb.y;
            dz = rb.z;
            bt = rb.w;
            ix = box_inv(dx);
            iy = box_inv(dy);
            iz = box_inv(dz);
            bu = 0.f;
            bv = 0.f;
            bp = RTBVH_MISS;
            const bool nan = (ox != ox) || (oy != oy) || (oz != oz) || (dx != dx) ||
                             (dy != dy) || (dz != dz);
            if (nan || p.node_count == 0) {
                if (ANY) {
                    p.occluded[IDS ? cur : rid] = 0;
                } else {
                    store_hit(p.hits, IDS ? cur : rid, bt, 0.f, 0.f, RTBVH_MISS);
                }
                sched.next();
            } else {
                active = true;
                const uint32_t signs = (__float_as_uint(ix) >> 31) |
                                       ((__float_as_uint(iy) >> 31) << 1) |
                                       ((__float_as_uint(iz) >> 31) << 2);
                const uint32_t flat = (fabsf(ix) > 3.4028235e38f ? 1u : 0u) |
                                      (fabsf(iy) > 3.4028235e38f ? 2u : 0u) |
                                      (fabsf(iz) > 3.4028235e38f ? 4u : 0u);
                // a flat axis reads its low planes as near ones
                oct = (signs & ~flat) | (flat << 3);
                // the root as a node group of one interior slot (slot 0), bound 0
                ngx = BATCH ? scene_of_ray(p.ray_offsets, p.scene_count, uint32_t(rid)) : 0u;
                ngy = (1u << (24 + (oct & 7u))) | (1u << 16);
                tgy = 0;
                sp = 0;
            }
        }
        if (!active) continue;
        bool overflow = false;
        bool found = false;
        // ---- a group in hand whose children all start beyond t_best is done
        if (ngy >= kNodeGroup && group_bound(ngy) > bt) ngy = 0;
        // ---- no node group in hand: pop until one that can still hold a
        // hit at or before t_best, taking a triangle group on the way when
        // none is in hand (one that is stops the search)
        while (ngy < kNodeGroup && sp > 0) {
            uint2 g;
            if (sp - 1 < ldepth)
                g = lds_get(lbase + uint32_t(sp - 1) * lstep);
            else
                g = deep[sp - 1 - ldepth];
            if (g.y >= kNodeGroup) {
                --sp;
                if (group_bound(g.y) <= bt) {
                    ngx = g.x;
                    ngy = g.y;
                }
            } else if ((tgy & 0xFFu) == 0u) {
                tgx = g.x;
                tgy = g.y;
                --sp;
            } else {
                break;
            }
        }
        // ---- triangle tests of this iteration: when enough lanes have one
        // waiting, or some lane has no node group to visit instead
        bool tri_turn = true;
        if (p.postpone > 1u) {
            const bool waiting = (tgy & 0xFFu) != 0u;
            const uint64_t w = __ballot(waiting);
            tri_turn = uint32_t(__popcll(w)) >= p.postpone ||
                       __ballot(waiting && ngy < kNodeGroup) != 0;
        }
        // ---- node step: the nearest-keyed child of the group in hand
        if (ngy >= kNodeGroup) {
            const uint32_t key = uint32_t(__builtin_ctz(ngy >> 24));
            ngy &= ~(1u << (24 + key));
            const uint32_t slot = key ^ (oct & 7u);
            const uint32_t idx = ngx + uint32_t(__popc((ngy >> 16) & ((1u << slot) - 1u)));
            if (ngy >= kNodeGroup) overflow = !push(ngx, ngy);
            ngy = 0;
            if (idx < p.node_count) {
                if (STATS) ++n_nodes;
                const uint4* q = reinterpret_cast<const uint4*>(p.cnodes + idx);
                const uint4 h0 = q[0], h1 = q[1], qa = q[2], qb = q[3], qc = q[4];
                // grid step times 1/d, and the grid origin's distance
                const float stx = __uint_as_float((h0.w & 0xFFu) << 23);
                const float sty = __uint_as_float(((h0.w >> 8) & 0xFFu) << 23);
                const float stz = __uint_as_float(((h0.w >> 16) & 0xFFu) << 23);
                float sx = stx * ix, sy = sty * iy, sz = stz * iz;
                float bx = (__uint_as_float(h0.x) - ox) * ix;
                float by = (__uint_as_float(h0.y) - oy) * iy;
                float bz = (__uint_as_float(h0.z) - oz) * iz;
                if (oct >= 8u) {
                    // along a zero direction component: plane distance H (q - u), u the
                    // ray's origin in grid steps, so that a plane a step or more beyond
                    // the origin is out of reach and one behind it is -H or lower
                    if (oct & 8u) {
                        sx = kFlatScale;
                        bx = flat_offset((ox - __uint_as_float(h0.x)) / stx);
                    }
                    if (oct & 16u) {
                        sy = kFlatScale;
                        by = flat_offset((oy - __uint_as_float(h0.y)) / sty);
                    }
                    if (oct & 32u) {
                        sz = kFlatScale;
                        bz = flat_offset((oz - __uint_as_float(h0.z)) / stz);
                    }
                }
                // outward slack: near planes use b - e, far planes b + e, with
                // e = 2^-20 (|b| + 255 |s|), twice the rounding of s, b and the
                // plane fma
                const float ex = fmaf(fabsf(sx), 255.0f, fabsf(bx)) * kBoxEps;
                const float ey = fmaf(fabsf(sy), 255.0f, fabsf(by)) * kBoxEps;
                const float ez = fmaf(fabsf(sz), 255.0f, fabsf(bz)) * kBoxEps;
                const float bxn = bx - ex, bxf = bx + ex;
                const float byn = by - ey, byf = by + ey;
                const float bzn = bz - ez, bzf = bz + ez;
                // near and far planes: low planes first along a positive direction
                const bool nx = (oct & 1u) != 0u, ny = (oct & 2u) != 0u, nz = (oct & 4u) != 0u;
                const uint32_t nx0 = nx ? qb.z : qa.x, nx1 = nx ? qb.w : qa.y;
                const uint32_t fx0 = nx ? qa.x : qb.z, fx1 = nx ? qa.y : qb.w;
                const uint32_t ny0 = ny ? qc.x : qa.z, ny1 = ny ? qc.y : qa.w;
                const uint32_t fy0 = ny ? qa.z : qc.x, fy1 = ny ? qa.w : qc.y;
                const uint32_t nz0 = nz ? qc.z : qb.x, nz1 = nz ? qc.w : qb.y;
                const uint32_t fz0 = nz ? qb.x : qc.z, fz1 = nz ? qb.y : qc.w;
                const uint32_t imask = h0.w >> 24, lmask = h1.z & 0xFFu;
                uint32_t hm = 0;
                // nearest entry among all hit slots: a lower bound for the interior
                // hits that costs one select per slot (the exact minimum over the
                // interior slots alone took a mask test per slot and pruned no more)
                float gmin = 3.0e38f;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float tnx = fmaf(qbyte(i < 4 ? nx0 : nx1, i), sx, bxn);
                    const float tfx = fmaf(qbyte(i < 4 ? fx0 : fx1, i), sx, bxf);
                    const float tny = fmaf(qbyte(i < 4 ? ny0 : ny1, i), sy, byn);
                    const float tfy = fmaf(qbyte(i < 4 ? fy0 : fy1, i), sy, byf);
                    const float tnz = fmaf(qbyte(i < 4 ? nz0 : nz1, i), sz, bzn);
                    const float tfz = fmaf(qbyte(i < 4 ? fz0 : fz1, i), sz, bzf);
                    const float en = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
                    const bool hit = en <= fminf(fminf(fminf(tfx, tfy), tfz), bt);
                    hm |= hit ? (1u << i) : 0u;
                    gmin = fminf(gmin, hit ? en : 3.0e38f);
                }
                const uint32_t ih = hm & imask, th = hm & lmask;
                if (ih) {
                    ngx = h1.x;
                    ngy = (oct_permute(ih, oct) << 24) | (imask << 16) | bound_bits(gmin);
                }
                if (th) {
                    if ((tgy & 0xFFu) != 0u && !overflow) overflow = !push(tgx, tgy);
                    tgx = h1.y;
                    tgy = (lmask << 8) | th;
                }
            }
        }
        // ---- triangle step: one triangle of the group in hand
        if (tri_turn && (tgy & 0xFFu) != 0u) {
            const uint32_t slot = uint32_t(__builtin_ctz(tgy & 0xFFu));
            tgy &= ~(1u << slot);
            const uint32_t pos = tgx + uint32_t(__popc((tgy >> 8) & ((1u << slot) - 1u)));
            if (pos < p.index_count) {
                if (STATS) ++n_tris;
                uint32_t prim;
                const MtResult m = mt_prepared(p.tris, pos, ox, oy, oz, dx, dy, dz, tmin, &prim);
                if (prim != RTBVH_MISS && m.ok) {
                    if (ANY) {
                        if (m.t < bt) {
                            bp = prim;
                            found = true;
                        }
                    } else if (lowest_first(m.t, prim, bt, bp)) {
                        bt = m.t;
                        bu = m.u;
                        bv = m.v;
                        bp = prim;
                    }
                }
            }
        }
        if (overflow) atomicOr(p.status, kStatusStackOverflow);
        if (found || overflow || (ngy < kNodeGroup && (tgy & 0xFFu) == 0u && sp == 0)) {
            if (ANY) {
                p.occluded[IDS ? cur : rid] = (bp != RTBVH_MISS) ? 1 : 0;
            } else {
                store_hit(p.hits, IDS ? cur : rid, overflow ? __builtin_nanf("") : bt, bu, bv,
                          overflow ? RTBVH_MISS : bp);
            }
            active = false;
            ngy = 0;
            tgy = 0;
            sched.next();
        }
    }
    if (STATS) {
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats), (unsigned long long)n_nodes);
        atomicAdd(reinterpret_cast<unsigned long long*>(p.stats) + 1, (unsigned long long)n_tris);
    }
}

// Launch shape per device: block size, LDS stack depth and the share of the
// resident-block capacity used by the persistent grid (percent).
TraceConfig g_config[64];
bool g_config_set[64] = {false};

int current_dev() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    return dev;
}

template <bool ANY, bool STATS, bool PREP, bool BATCH = false, bool ITEM = false, bool CW8 = false,
          bool IDS = false>
hipError_t launch_fast(const TraceParams& p0, hipStream_t stream) {
    auto k = CW8 ? cw8_kernel<ANY, STATS, BATCH, IDS>
                 : (ITEM ? mbvh_item_kernel<ANY, STATS, BATCH>
                         : mbvh_fast_kernel<ANY, STATS, PREP, BATCH>);
    const int dev = current_dev();
    const TraceConfig cfg = g_config_set[dev] ? g_config[dev] : default_trace_config();
    // resident blocks per CU for this shape, cached per (device, shape)
    static int cache_key[64] = {0};
    static int cache_blocks[64] = {0};
    const size_t lds = size_t(cfg.lds_depth) * cfg.block * 8;
    const int key = int(cfg.block) * 1000 + int(cfg.lds_depth) + 1;
    if (cache_key[dev] != key) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, int(cfg.block), lds) !=
                hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        cache_blocks[dev] = cus * per_cu;
        cache_key[dev] = key;
    }
    uint64_t blocks = uint64_t(cache_blocks[dev]) * cfg.occupancy_pct / 100;
    if (blocks == 0) blocks = 1;
    // Wave ranges need a few rays per lane: with r rays per resident lane
    // they run with refill min(refill, r / 2) and enough waves to fill the
    // device, but no more than leave each lane max(refill, 8) rays; launches
    // with fewer than 2 rays per resident lane (a scene of a few thousand
    // triangles) run the static schedule with refill 1.
    const uint64_t resident = uint64_t(cache_blocks[dev]);
    const uint64_t rays_per_lane = p0.ray_count / (resident * cfg.block);
    const bool ranges = (ITEM || CW8) && cfg.schedule == 1 && rays_per_lane >= 2;
    uint32_t refill = cfg.schedule == 1 ? 1u : cfg.refill;
    if (ranges) {
        refill = uint32_t(rays_per_lane / 2 < cfg.refill ? rays_per_lane / 2 : cfg.refill);
        if (refill == 0) refill = 1;
        const uint64_t per_block = uint64_t(cfg.block) * (refill > 8 ? refill : 8u);
        uint64_t most = (p0.ray_count + per_block - 1) / per_block;
        if (most < resident) most = resident;
        if (most < blocks) blocks = most;
    }
    const uint64_t need = (p0.ray_count + cfg.block - 1) / cfg.block;
    const int grid = int(need < blocks ? need : blocks);
    if (grid == 0) return hipSuccess;
    TraceParams p = p0;
    p.lds_depth = cfg.lds_depth;
    p.refill = refill;
    p.postpone = cfg.postpone;
    p.wave_ranges = ranges ? 1u : 0u;
    hipLaunchKernelGGL(k, dim3(grid), dim3(cfg.block), lds, stream, p);
    return hipGetLastError();
}

}  // namespace

TraceConfig default_trace_config() { return TraceConfig{64, 8, 800, 0, 16, 1, 0}; }

bool valid_trace_config(const TraceConfig& c) {
    return (c.block == 64 || c.block == 128 || c.block == 256) && c.lds_depth <= 32 &&
           c.occupancy_pct >= 10 && c.occupancy_pct <= 3200 && c.kernel <= 1 && c.refill >= 1 &&
           c.refill <= 64 && c.schedule <= 1 && c.postpone <= 64;
}

void set_trace_config(const TraceConfig* c) {
    const int dev = current_dev();
    if (c) g_config[dev] = *c;
    g_config_set[dev] = c != nullptr;
}

TraceConfig get_trace_config() {
    const int dev = current_dev();
    return g_config_set[dev] ? g_config[dev] : default_trace_config();
}

hipError_t launch_trace_fast(const TraceParams& p, bool any_hit, hipStream_t stream) {
    // rays read through an index: compressed trees, one scene, no counters
    if (p.ray_ids && (!p.cnodes || p.ray_offsets || p.stats)) return hipErrorInvalidValue;
    if (p.cnodes) {  // compressed 8-wide tree
        if (!p.tris || (p.stats && (any_hit || p.ray_offsets))) return hipErrorInvalidValue;
        if (p.ray_offsets)
            return any_hit ? launch_fast<true, false, true, true, false, true>(p, stream)
                           : launch_fast<false, false, true, true, false, true>(p, stream);
        if (p.stats) return launch_fast<false, true, true, false, false, true>(p, stream);
        if (p.ray_ids)
            return any_hit ? launch_fast<true, false, true, false, false, true, true>(p, stream)
                           : launch_fast<false, false, true, false, false, true, true>(p, stream);
        return any_hit ? launch_fast<true, false, true, false, false, true>(p, stream)
                       : launch_fast<false, false, true, false, false, true>(p, stream);
    }
    // the one-item kernel over leaf-ordered triangles, unless the leaf tags
    // cannot address every index entry
    const bool item = p.tris && p.small_leaves && p.index_count <= kLeafFirstMask &&
                      get_trace_config().kernel == 0;
    if (p.ray_offsets) {
        if (p.stats || !p.tris) return hipErrorInvalidValue;
        if (item)
            return any_hit ? launch_fast<true, false, true, true, true>(p, stream)
                           : launch_fast<false, false, true, true, true>(p, stream);
        return any_hit ? launch_fast<true, false, true, true>(p, stream)
                       : launch_fast<false, false, true, true>(p, stream);
    }
    if (p.stats) {
        if (any_hit) return hipErrorInvalidValue;
        if (item) return launch_fast<false, true, true, false, true>(p, stream);
        return p.tris ? launch_fast<false, true, true>(p, stream)
                      : launch_fast<false, true, false>(p, stream);
    }
    if (item)
        return any_hit ? launch_fast<true, false, true, false, true>(p, stream)
                       : launch_fast<false, false, true, false, true>(p, stream);
    if (p.tris)
        return any_hit ? launch_fast<true, false, true>(p, stream)
                       : launch_fast<false, false, true>(p, stream);
    return any_hit ? launch_fast<true, false, false>(p, stream)
                   : launch_fast<false, false, false>(p, stream);
}

namespace {
// scene_off (batches, device, scenes + 1 entries): the id stored is the
// primitive's index within its scene.
__global__ void prepare_tris_kernel(const uint32_t* indices, uint32_t index_count,
                                    const float* verts, uint32_t prim_count, float4* tris,
                                    const uint32_t* scene_off, uint32_t scenes) {
    const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= index_count) return;
    const uint32_t prim = indices[pos];
    float4 a = make_float4(0.f, 0.f, 0.f, __uint_as_float(RTBVH_MISS));
    float4 e1 = make_float4(0.f, 0.f, 0.f, 0.f), e2 = e1;
    if (prim < prim_count) {
        const float* v = verts + size_t(prim) * 9;
        const uint32_t id = scene_off ? prim - scene_off[scene_of_ray(scene_off, scenes, prim)]
                                      : prim;
        a = make_float4(v[0], v[1], v[2], __uint_as_float(id));
        e1 = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], 0.f);
        e2 = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.f);
    }
    tris[size_t(pos) * kTriRecord] = a;
    tris[size_t(pos) * kTriRecord + 1] = e1;
    tris[size_t(pos) * kTriRecord + 2] = e2;
}
}  // namespace

namespace {
__global__ void max_leaf_kernel(const DevMbvhNode* nodes, uint32_t count, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int4 c = nodes[i].counts;
    const int32_t m = max(max(c.x, c.y), max(c.z, c.w));
    if (m > 0) atomicMax(out, uint32_t(m));
}
}  // namespace

// Largest leaf (triangles per slot) of an MBVH; synchronises the stream.
hipError_t mbvh_max_leaf(const DevMbvhNode* nodes, uint32_t count, uint32_t* result,
                         hipStream_t stream) {
    *result = 0;
    if (count == 0) return hipSuccess;
    uint32_t* d = nullptr;
    hipError_t e = hipMalloc(&d, 4);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(d, result, 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(max_leaf_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, nodes,
                           count, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(result, d, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(d);
    return e;
}

hipError_t launch_prepare_triangles(const uint32_t* indices, uint32_t index_count,
                                    const float* verts, uint32_t prim_count, float4* tris,
                                    hipStream_t stream, const uint32_t* scene_off,
                                    uint32_t scenes) {
    if (index_count == 0) return hipSuccess;
    hipLaunchKernelGGL(prepare_tris_kernel, dim3((index_count + 255) / 256), dim3(256), 0, stream,
                       indices, index_count, verts, prim_count, tris, scene_off, scenes);
    return hipGetLastError();
}

}  // namespace rtbvh
