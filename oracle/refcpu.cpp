// refcpu — CPU restatement of meirbon/rtbvh's hot path.
//
// TEST INFRASTRUCTURE ONLY: the oracle for the GPU kernels in rtbvh_amd/.
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
//
// All arithmetic is IEEE binary32, compiled with -ffp-contract=off and no
// fast-math (the reference is Rust, which never contracts, on SSE2).  The
// reference's float arithmetic goes through glam (unpinned >=0.14, not
// vendored): Vec3 is scalar, so Vec3::min/max are f32::min/max (NaN-ignoring,
// lowered on x86-64 to MINSS/MAXSS with a NaN select), and Vec4 is SSE2, so
// Vec4::min/max are MINPS/MAXPS (`a<b?a:b`, `a>b?a:b`, second operand on NaN).
// Those two families are written out below as rs_min/rs_max and ps_min/ps_max.

#include "refcpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- scalars --

// f32::min / f32::max as LLVM lowers llvm.minnum/maxnum on x86-64 SSE2:
// MINSS(b, a) followed by "a is NaN ? b : ...".  Ties (including +-0) return a.
inline float rs_min(float a, float b) { return (a != a) ? b : (b < a ? b : a); }
inline float rs_max(float a, float b) { return (a != a) ? b : (b > a ? b : a); }
// _mm_min_ps / _mm_max_ps lane semantics (glam Vec4 on SSE2).
inline float ps_min(float a, float b) { return a < b ? a : b; }
inline float ps_max(float a, float b) { return a > b ? a : b; }

// IEEE minNum/maxNum (a NaN operand is ignored), inlined.
inline float ieee_min(float a, float b) { return (b != b) ? a : ((a != a) ? b : (a < b ? a : b)); }
inline float ieee_max(float a, float b) { return (b != b) ? a : ((a != a) ? b : (a > b ? a : b)); }

inline bool is_nan3(const float v[3]) { return v[0] != v[0] || v[1] != v[1] || v[2] != v[2]; }

// Rust `f as i32` (saturating, NaN -> 0).
inline int32_t sat_i32(float g) {
    if (g != g) return 0;
    if (g >= 2147483648.0f) return std::numeric_limits<int32_t>::max();
    if (g <= -2147483904.0f) return std::numeric_limits<int32_t>::min();
    return (int32_t)g;
}

// Rust `f as usize` (saturating, NaN -> 0, negative -> 0).
inline uint64_t sat_usize(float g) {
    if (!(g > 0.0f)) return 0;
    if (g >= 18446744073709551616.0f) return std::numeric_limits<uint64_t>::max();
    return (uint64_t)g;
}

// -------------------------------------------------------------- Aabb ops --
// src/aabb.rs

constexpr float kEmptyMin = 1e34f;   // Aabb::new, aabb.rs:50-57
constexpr float kEmptyMax = -1e34f;

inline RefAabb aabb_new() {
    RefAabb a;
    a.min[0] = a.min[1] = a.min[2] = kEmptyMin;
    a.max[0] = a.max[1] = a.max[2] = kEmptyMax;
    a.extra1 = 0;
    a.extra2 = 0;
    return a;
}

// Aabb::union_of, aabb.rs:74-86 (extras reset to default 0).
inline RefAabb union_of(const RefAabb& s, const RefAabb& b) {
    RefAabb r;
    for (int k = 0; k < 3; ++k) {
        r.min[k] = rs_min(s.min[k], b.min[k]);
        r.max[k] = rs_max(s.max[k], b.max[k]);
    }
    r.extra1 = 0;
    r.extra2 = 0;
    return r;
}

// Aabb::grow_bb, aabb.rs:264-273 (extras untouched).
inline void grow_bb(RefAabb& s, const RefAabb& b) {
    for (int k = 0; k < 3; ++k) {
        s.min[k] = rs_min(s.min[k], b.min[k]);
        s.max[k] = rs_max(s.max[k], b.max[k]);
    }
}

// Aabb::grow (point), aabb.rs:252-261.
inline void grow_pt(RefAabb& s, const float p[3]) {
    for (int k = 0; k < 3; ++k) {
        s.min[k] = rs_min(s.min[k], p[k]);
        s.max[k] = rs_max(s.max[k], p[k]);
    }
}

// Aabb::offset_by / with_offset, aabb.rs:301-321.
inline void offset_by(RefAabb& s, float delta) {
    for (int k = 0; k < 3; ++k) {
        s.min[k] = s.min[k] - delta;
        s.max[k] = s.max[k] + delta;
    }
}

// Aabb::half_area, aabb.rs:343-346: d = max - min; (d0 + d1) * d2 + d0 * d1.
inline float half_area(const RefAabb& a) {
    const float d0 = a.max[0] - a.min[0];
    const float d1 = a.max[1] - a.min[1];
    const float d2 = a.max[2] - a.min[2];
    return (d0 + d1) * d2 + d0 * d1;
}

// Aabb::longest_axis, aabb.rs:354-363.
inline int longest_axis(const RefAabb& a) {
    int ax = 0;
    if (a.max[1] - a.min[1] > a.max[0] - a.min[0]) ax = 1;
    if (a.max[2] - a.min[2] > a.max[ax] - a.min[ax]) ax = 2;
    return ax;
}

// Aabb::union_of_list, aabb.rs:125-131: grow from empty, then +-1e-4.
inline RefAabb union_of_list(const RefAabb* aabbs, size_t n) {
    RefAabb a = aabb_new();
    for (size_t i = 0; i < n; ++i) grow_bb(a, aabbs[i]);
    offset_by(a, 0.0001f);
    return a;
}

// ---------------------------------------------------------------- Morton --
// src/morton.rs:10-25 — literal port of the mask loop (64-bit usize).
uint32_t morton_split(uint32_t v) {
    const uint32_t log_bits = 5;  // round_up_log2(32, 0), utils.rs:12-22
    uint64_t x = v;
    uint64_t mask = ~uint64_t(0);
    uint32_t i = log_bits;
    uint64_t n = uint64_t(1) << log_bits;
    while (i > 0) {
        mask = (mask | (mask << n)) & ~(mask << (n / 2));
        x = (x | (x << n)) & mask;
        n >>= 1;
        i -= 1;
    }
    return (uint32_t)x;
}

struct MortonEncoder {
    float wtg[3];
    float off[3];
};

// MortonEncoder::new, morton.rs:36-46 with grid_dim = MAX_GRID_DIM = 1024.
MortonEncoder make_encoder(const RefAabb& world) {
    MortonEncoder e;
    for (int k = 0; k < 3; ++k) {
        const float diag = world.max[k] - world.min[k];
        e.wtg[k] = 1024.0f * (1.0f / diag);
        e.off[k] = (-world.min[k]) * e.wtg[k];
    }
    return e;
}

// MortonEncoder::encode + morton_encode, morton.rs:48-61.
inline uint32_t encode(const MortonEncoder& e, const float c[3]) {
    uint32_t x[3];
    for (int k = 0; k < 3; ++k) {
        const float g = c[k] * e.wtg[k] + e.off[k];
        int32_t q = sat_i32(g);
        q = std::max(q, 0);
        q = std::min(1023, q);
        x[k] = (uint32_t)q;
    }
    return morton_split(x[0]) | (morton_split(x[1]) << 1) | (morton_split(x[2]) << 2);
}

// ----------------------------------------------------------------- Rays --
// Ray::new, src/ray.rs:166-182 with ray.t set by the caller
// (rtbvh_ffi/src/lib.rs:565-566).
struct Ray {
    float o[3];
    float t_min;
    float d[3];
    float t;
    float inv[3];
    int sign[3];
};

inline Ray make_ray(const RefRay& r) {
    Ray ray;
    for (int k = 0; k < 3; ++k) {
        ray.o[k] = r.o[k];
        ray.d[k] = r.d[k];
        ray.inv[k] = 1.0f / r.d[k];
        ray.sign[k] = r.d[k] < 0.0f ? 1 : 0;
    }
    ray.t_min = r.t_min;
    ray.t = r.t_max;
    return ray;
}

// ------------------------------------------------------ Moller-Trumbore --
// SpatialTriangle::intersect, src/builders/spatial_sah.rs:132-163.
// glam Vec3::cross(a, b) = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y),
// Vec3::dot = (a.x*b.x + a.y*b.y) + a.z*b.z, evaluated left to right.
struct MT {
    bool ok;
    float t, u, v;
};

inline MT mt_test(const float* p, const float o[3], const float d[3], float t_min, float t_cur) {
    MT r{false, 0.f, 0.f, 0.f};
    const float e1x = p[3] - p[0], e1y = p[4] - p[1], e1z = p[5] - p[2];
    const float e2x = p[6] - p[0], e2y = p[7] - p[1], e2z = p[8] - p[2];
    const float hx = d[1] * e2z - e2y * d[2];
    const float hy = d[2] * e2x - e2z * d[0];
    const float hz = d[0] * e2y - e2x * d[1];
    const float a = (e1x * hx + e1y * hy) + e1z * hz;
    if (a > -1e-5f && a < 1e-5f) return r;
    const float f = 1.0f / a;
    const float sx = o[0] - p[0], sy = o[1] - p[1], sz = o[2] - p[2];
    const float u = f * ((sx * hx + sy * hy) + sz * hz);
    if (!(0.0f <= u && u <= 1.0f)) return r;
    const float qx = sy * e1z - e1y * sz;
    const float qy = sz * e1x - e1z * sx;
    const float qz = sx * e1y - e1x * sy;
    const float v = f * ((d[0] * qx + d[1] * qy) + d[2] * qz);
    if (v < 0.0f || u + v > 1.0f) return r;
    const float t = f * ((e2x * qx + e2y * qy) + e2z * qz);
    if (t > t_min && t < t_cur) {
        r.ok = true;
        r.t = t;
        r.u = u;
        r.v = v;
    }
    return r;
}

// Closest-hit state.  Visit-order mode is the reference driver
// (examples/benchmark.rs:33-37): every accepted test lowers ray.t and the last
// accepted candidate is the hit.  Lowest-id mode keeps the lexicographic
// minimum of (t, prim).
struct HitState {
    float t;
    float u, v;
    uint32_t prim;
};

inline void try_prim(const float* verts, size_t prim_count, uint32_t prim, const Ray& ray,
                     HitState& h, int mode, uint64_t& tri_tests) {
    if (prim >= prim_count) return;  // MbvhIterator skips missing prims (iter.rs:337)
    ++tri_tests;
    const float* p = verts + size_t(prim) * 9;
    if (mode == REF_MODE_VISIT_ORDER) {
        MT m = mt_test(p, ray.o, ray.d, ray.t_min, h.t);
        if (m.ok) {
            h.t = m.t;
            h.u = m.u;
            h.v = m.v;
            h.prim = prim;
        }
    } else {
        // accept t == h.t when the id is lower: test against nextafter(h.t)
        // would change the comparison form, so evaluate with t_cur = +inf and
        // compare here.  The tie only counts against an accepted hit: a
        // candidate exactly at the ray's t_max is rejected, as the
        // reference's `t < ray.t` (src/builders/spatial_sah.rs:157) and brute
        // force do.
        MT m = mt_test(p, ray.o, ray.d, ray.t_min, std::numeric_limits<float>::infinity());
        if (m.ok && (m.t < h.t || (m.t == h.t && h.prim != 0xFFFFFFFFu && prim < h.prim))) {
            h.t = m.t;
            h.u = m.u;
            h.v = m.v;
            h.prim = prim;
        }
    }
}

// ------------------------------------------------------ MBVH node test ----
// MbvhNode::intersect, src/mbvh_node.rs:177-240.
struct MbvhHit {
    int ids[4];
    bool result[4];
    float tmin[4];
    float tmax[4];
};

inline MbvhHit mbvh_intersect(const RefMbvhNode& n, const Ray& ray, float t_cur) {
    MbvhHit h;
    float tmin[4];
    for (int i = 0; i < 4; ++i) {
        const float tx0 = (n.min_x[i] - ray.o[0]) * ray.inv[0];
        const float tx1 = (n.max_x[i] - ray.o[0]) * ray.inv[0];
        const float ty0 = (n.min_y[i] - ray.o[1]) * ray.inv[1];
        const float ty1 = (n.max_y[i] - ray.o[1]) * ray.inv[1];
        const float tz0 = (n.min_z[i] - ray.o[2]) * ray.inv[2];
        const float tz1 = (n.max_z[i] - ray.o[2]) * ray.inv[2];
        const float txmin = ps_min(tx0, tx1), txmax = ps_max(tx0, tx1);
        const float tymin = ps_min(ty0, ty1), tymax = ps_max(ty0, ty1);
        const float tzmin = ps_min(tz0, tz1), tzmax = ps_max(tz0, tz1);
        const float tn = ps_max(txmin, ps_max(tymin, tzmin));
        const float tf = ps_min(txmax, ps_min(tymax, tzmax));
        h.result[i] = (tf >= tn) && (tn < t_cur);
        tmin[i] = tn;
        h.tmin[i] = tn;
        h.tmax[i] = tf;
    }
    int ids[4] = {0, 1, 2, 3};
    // Sorting network of mbvh_node.rs:219-237 (last comparator swaps ids only).
    if (tmin[0] > tmin[1]) { std::swap(tmin[0], tmin[1]); std::swap(ids[0], ids[1]); }
    if (tmin[2] > tmin[3]) { std::swap(tmin[2], tmin[3]); std::swap(ids[2], ids[3]); }
    if (tmin[0] > tmin[2]) { std::swap(tmin[0], tmin[2]); std::swap(ids[0], ids[2]); }
    if (tmin[1] > tmin[3]) { std::swap(tmin[1], tmin[3]); std::swap(ids[1], ids[3]); }
    if (tmin[2] > tmin[3]) { std::swap(ids[2], ids[3]); }
    for (int i = 0; i < 4; ++i) h.ids[i] = ids[i];
    return h;
}

// MbvhIterator::next, src/iter.rs:308-354 driven to exhaustion with the
// closest-hit driver of examples/benchmark.rs:33-37.  The node's hit mask is
// taken when the node is popped; its leaf slots are then yielded in the order
// ids[3], ids[2], ids[1], ids[0] and interior slots pushed in that order.
void trace_mbvh_one(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                    size_t index_count, const float* verts, size_t prim_count, const RefRay& in,
                    RefHit& out, int mode, uint64_t& nodes_visited, uint64_t& tri_tests,
                    std::vector<int32_t>& stack) {
    Ray ray = make_ray(in);
    HitState h{ray.t, 0.f, 0.f, 0xFFFFFFFFu};
    if (m_count == 0 || is_nan3(ray.o) || is_nan3(ray.d)) {
        out = RefHit{h.t, h.u, h.v, h.prim};
        return;
    }
    stack.clear();
    int32_t current = 0;
    for (;;) {
        const RefMbvhNode& node = m[current];
        ++nodes_visited;
        MbvhHit hit = mbvh_intersect(node, ray, h.t);
        for (int i = 0; i < 4; ++i) {
            const int id = hit.ids[3 - i];
            if (!hit.result[id]) continue;
            const int32_t count = node.counts[id];
            const int32_t lf = node.children[id];
            if (count > -1) {
                for (int32_t j = 0; j < count; ++j) {
                    const size_t k = size_t(lf) + size_t(j);
                    if (k >= index_count) break;
                    try_prim(verts, prim_count, indices[k], ray, h, mode, tri_tests);
                }
            } else if (lf > -1) {
                stack.push_back(lf);
            }
        }
        if (stack.empty()) break;
        current = stack.back();
        stack.pop_back();
        if (size_t(current) >= m_count) break;  // malformed tree guard
    }
    out = RefHit{h.t, h.u, h.v, h.prim};
}

// Lowest-(t, id) traversal of the MBVH.  The result is the brute-force
// closest hit; the visit order is the one of the GPU kernel in
// rtbvh_amd/csrc/trace_fast.hip, so that their work counters agree:
// IEEE min/max slabs, slot keys (entry distance bits, 2 low bits = slot)
// sorted with comparators (0,1),(2,3),(0,2),(1,3),(1,2), leaves near to far,
// nearest interior child next, the rest stacked far to near with their entry
// distance and dropped at pop when beyond the current t.
void trace_mbvh_lowest_one(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                           size_t index_count, const float* verts, size_t prim_count,
                           const RefRay& in, RefHit& out, uint64_t& nodes_visited,
                           uint64_t& tri_tests, std::vector<std::pair<int32_t, float>>& stack) {
    Ray ray = make_ray(in);
    HitState h{ray.t, 0.f, 0.f, 0xFFFFFFFFu};
    if (m_count == 0 || is_nan3(ray.o) || is_nan3(ray.d)) {
        out = RefHit{h.t, h.u, h.v, h.prim};
        return;
    }
    stack.clear();
    int32_t cur = 0;
    for (;;) {
        if (cur < 0) {
            while (!stack.empty()) {
                const auto e = stack.back();
                stack.pop_back();
                if (e.second <= h.t) {
                    cur = e.first;
                    break;
                }
            }
            if (cur < 0 || size_t(cur) >= m_count) break;
        }
        const RefMbvhNode& n = m[cur];
        ++nodes_visited;
        float tn[4];
        uint32_t key[4];
        for (int i = 0; i < 4; ++i) {
            const float tx0 = (n.min_x[i] - ray.o[0]) * ray.inv[0];
            const float tx1 = (n.max_x[i] - ray.o[0]) * ray.inv[0];
            const float ty0 = (n.min_y[i] - ray.o[1]) * ray.inv[1];
            const float ty1 = (n.max_y[i] - ray.o[1]) * ray.inv[1];
            const float tz0 = (n.min_z[i] - ray.o[2]) * ray.inv[2];
            const float tz1 = (n.max_z[i] - ray.o[2]) * ray.inv[2];
            const float a = ieee_max(ieee_max(ieee_min(tx0, tx1), ieee_min(ty0, ty1)),
                                     ieee_min(tz0, tz1));
            const float b = ieee_min(ieee_min(ieee_max(tx0, tx1), ieee_max(ty0, ty1)),
                                     ieee_max(tz0, tz1));
            tn[i] = a;
            const bool hit = (b >= a) && (a <= h.t) && (b >= ray.t_min);
            float ak = a > 0.0f ? a : 0.0f;
            uint32_t bits;
            std::memcpy(&bits, &ak, 4);
            key[i] = hit ? ((bits & ~3u) | uint32_t(i)) : 0xFFFFFFFFu;
        }
        auto cs = [&](int x, int y) {
            if (key[x] > key[y]) std::swap(key[x], key[y]);
        };
        cs(0, 1);
        cs(2, 3);
        cs(0, 2);
        cs(1, 3);
        cs(1, 2);
        for (int k = 0; k < 4; ++k) {
            if (key[k] == 0xFFFFFFFFu) continue;
            const int s = int(key[k] & 3u);
            if (n.counts[s] < 0 || !(tn[s] <= h.t)) continue;
            for (int32_t j = 0; j < n.counts[s]; ++j) {
                const size_t pos = size_t(n.children[s]) + size_t(j);
                if (pos >= index_count) break;
                try_prim(verts, prim_count, indices[pos], ray, h, REF_MODE_LOWEST_ID, tri_tests);
            }
        }
        int32_t next = -1;
        float next_t = 0.f;
        for (int k = 3; k >= 0; --k) {
            if (key[k] == 0xFFFFFFFFu) continue;
            const int s = int(key[k] & 3u);
            if (n.counts[s] >= 0 || n.children[s] < 0 || !(tn[s] <= h.t)) continue;
            if (next >= 0) stack.push_back({next, next_t});
            next = n.children[s];
            next_t = tn[s];
        }
        cur = next;
    }
    out = RefHit{h.t, h.u, h.v, h.prim};
}

// Aabb::intersect (BVH2 slab test), src/aabb.rs:146-181.  Returns t_far.
inline bool bvh_slab(const RefAabb& b, const Ray& ray, float* t_out) {
    const float* lo[2] = {b.min, b.max};
    float ray_min = (lo[ray.sign[0]][0] - ray.o[0]) * ray.inv[0];
    float ray_max = (lo[1 - ray.sign[0]][0] - ray.o[0]) * ray.inv[0];
    const float y_min = (lo[ray.sign[1]][1] - ray.o[1]) * ray.inv[1];
    const float y_max = (lo[1 - ray.sign[1]][1] - ray.o[1]) * ray.inv[1];
    if ((ray_min > y_max) || (y_min > ray_max)) return false;
    if (y_min > ray_min) ray_min = y_min;
    if (y_max < ray_max) ray_max = y_max;
    const float z_min = (lo[ray.sign[2]][2] - ray.o[2]) * ray.inv[2];
    const float z_max = (lo[1 - ray.sign[2]][2] - ray.o[2]) * ray.inv[2];
    if ((ray_min > z_max) || (z_min > ray_max)) return false;
    if (z_max < ray_max) ray_max = z_max;
    if (ray_max > ray.t_min) {
        *t_out = ray_max;
        return true;
    }
    return false;
}

// BvhIterator::next, src/iter.rs:88-127 with BvhNode::sort_nodes,
// src/bvh_node.rs:150-177 (both children hit: the one with the larger t_far is
// popped first).  The root is pushed without a box test.
void trace_bvh_one(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                   size_t index_count, const float* verts, size_t prim_count, const RefRay& in,
                   RefHit& out, int mode, uint64_t& nodes_visited, uint64_t& tri_tests,
                   std::vector<int32_t>& stack) {
    Ray ray = make_ray(in);
    HitState h{ray.t, 0.f, 0.f, 0xFFFFFFFFu};
    if (node_count == 0 || is_nan3(ray.o) || is_nan3(ray.d)) {
        out = RefHit{h.t, h.u, h.v, h.prim};
        return;
    }
    stack.clear();
    stack.push_back(0);
    while (!stack.empty()) {
        const int32_t ni = stack.back();
        stack.pop_back();
        if (ni < 0 || size_t(ni) >= node_count) continue;
        ++nodes_visited;
        const RefAabb& node = nodes[ni];
        const int32_t count = node.extra1;
        const int32_t lf = node.extra2;
        if (count > -1) {
            for (int32_t j = 0; j < count; ++j) {
                const size_t k = size_t(lf) + size_t(j);
                if (k >= index_count) break;
                try_prim(verts, prim_count, indices[k], ray, h, mode, tri_tests);
            }
        } else if (lf > -1) {
            if (size_t(lf) + 1 >= node_count) continue;
            float tl = 0.f, tr = 0.f;
            const bool hl = bvh_slab(nodes[lf], ray, &tl);
            const bool hr = bvh_slab(nodes[lf + 1], ray, &tr);
            if (hl && hr) {
                if (tl < tr) {
                    stack.push_back(lf);
                    stack.push_back(lf + 1);
                } else {
                    stack.push_back(lf + 1);
                    stack.push_back(lf);
                }
            } else if (hl) {
                stack.push_back(lf);
            } else if (hr) {
                stack.push_back(lf + 1);
            }
        }
    }
    out = RefHit{h.t, h.u, h.v, h.prim};
}

bool occluded_mbvh_one(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                       size_t index_count, const float* verts, size_t prim_count,
                       const RefRay& in, std::vector<int32_t>& stack) {
    Ray ray = make_ray(in);
    if (m_count == 0 || is_nan3(ray.o) || is_nan3(ray.d)) return false;
    stack.clear();
    int32_t current = 0;
    for (;;) {
        const RefMbvhNode& node = m[current];
        MbvhHit hit = mbvh_intersect(node, ray, ray.t);
        for (int i = 0; i < 4; ++i) {
            const int id = hit.ids[3 - i];
            if (!hit.result[id]) continue;
            const int32_t count = node.counts[id];
            const int32_t lf = node.children[id];
            if (count > -1) {
                for (int32_t j = 0; j < count; ++j) {
                    const size_t k = size_t(lf) + size_t(j);
                    if (k >= index_count) break;
                    const uint32_t prim = indices[k];
                    if (prim >= prim_count) continue;
                    MT r = mt_test(verts + size_t(prim) * 9, ray.o, ray.d, ray.t_min, ray.t);
                    if (r.ok) return true;  // callback returns true => early out (lib.rs:571-576)
                }
            } else if (lf > -1) {
                stack.push_back(lf);
            }
        }
        if (stack.empty()) break;
        current = stack.back();
        stack.pop_back();
        if (size_t(current) >= m_count) break;
    }
    return false;
}

// Any hit over the BVH: BvhIndexIterator order (iter_indices.rs:69-106) with
// the FFI early-out (rtbvh_ffi/src/lib.rs:571-576).
bool occluded_bvh_one(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                      size_t index_count, const float* verts, size_t prim_count,
                      const RefRay& in, std::vector<int32_t>& stack) {
    Ray ray = make_ray(in);
    if (node_count == 0 || is_nan3(ray.o) || is_nan3(ray.d)) return false;
    stack.clear();
    stack.push_back(0);
    while (!stack.empty()) {
        const int32_t ni = stack.back();
        stack.pop_back();
        if (ni < 0 || size_t(ni) >= node_count) continue;
        const RefAabb& node = nodes[ni];
        if (node.extra1 > -1) {
            for (int32_t j = 0; j < node.extra1; ++j) {
                const size_t k = size_t(node.extra2) + size_t(j);
                if (k >= index_count) break;
                const uint32_t prim = indices[k];
                if (prim >= prim_count) continue;
                MT r = mt_test(verts + size_t(prim) * 9, ray.o, ray.d, ray.t_min, ray.t);
                if (r.ok) return true;
            }
        } else if (node.extra2 > -1) {
            const int32_t lf = node.extra2;
            if (size_t(lf) + 1 >= node_count) continue;
            float tl = 0.f, tr = 0.f;
            const bool hl = bvh_slab(nodes[lf], ray, &tl);
            const bool hr = bvh_slab(nodes[lf + 1], ray, &tr);
            if (hl && hr) {
                if (tl < tr) {
                    stack.push_back(lf);
                    stack.push_back(lf + 1);
                } else {
                    stack.push_back(lf + 1);
                    stack.push_back(lf);
                }
            } else if (hl) {
                stack.push_back(lf);
            } else if (hr) {
                stack.push_back(lf + 1);
            }
        }
    }
    return false;
}

template <class F>
void parallel_rays(size_t n, int threads, F&& fn) {
    if (threads <= 1 || n < 2000) {
        fn(size_t(0), n, 0);
        return;
    }
    std::atomic<size_t> next{0};
    const size_t chunk = 1000;  // examples/benchmark.rs:25 (par_chunks_mut(WIDTH))
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&, t]() {
            for (;;) {
                const size_t b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk), t);
            }
        });
    }
    for (auto& th : pool) th.join();
}

// --------------------------------------------------------- LOCB builder --
// LocallyOrderedClusteringBuilder, src/builders/locb.rs.
struct Locb {
    size_t radius;

    // search_range, locb.rs:36-45
    void search_range(size_t i, size_t begin, size_t end, size_t& sb, size_t& se) const {
        sb = (i > begin + radius) ? i - radius : begin;
        se = std::min(i + radius + 1, end);
    }

    // cluster, locb.rs:48-245.  The reference caches pair distances in a
    // rotating (radius+1) x radius matrix; every cached value is
    // half_area(input[lo].union_of(input[hi])) with lo < hi, so it is
    // evaluated directly here (same operands, same order, same value).
    void cluster(const std::vector<RefAabb>& input, std::vector<RefAabb>& output,
                 std::vector<uint32_t>& neighbours, std::vector<uint32_t>& merged,
                 size_t begin, size_t end, size_t previous_end, size_t& next_begin,
                 size_t& next_end) const {
        // nearest-neighbour search, locb.rs:92-142: ascending j, strict '<'
        for (size_t i = begin; i < end; ++i) {
            size_t sb, se;
            search_range(i, begin, end, sb, se);
            float best = std::numeric_limits<float>::max();
            int64_t best_j = -1;
            for (size_t j = sb; j < se; ++j) {
                if (j == i) continue;
                const size_t lo = std::min(i, j), hi = std::max(i, j);
                const float dist = half_area(union_of(input[lo], input[hi]));
                if (dist < best) {
                    best = dist;
                    best_j = int64_t(j);
                }
            }
            neighbours[i] = uint32_t(best_j);  // -1 as u32 when no candidate
        }
        // mutual pairs marked at the lower index, locb.rs:159-167
        for (size_t i = begin; i < end; ++i) {
            const size_t j = neighbours[i];
            const bool mergeable = (j < neighbours.size()) && size_t(neighbours[j]) == i;
            merged[i] = (i < j && mergeable) ? 1u : 0u;
        }
        // inclusive prefix sum, locb.rs:170-176 (utils.rs:42-57)
        {
            uint32_t sum = 0;
            for (size_t i = begin; i < end; ++i) {
                sum += merged[i];
                merged[i] = sum;
            }
        }
        const size_t merged_count = merged[end - 1];
        const size_t unmerged_count = end - begin - merged_count;
        const size_t children_count = merged_count * 2;
        const size_t children_begin = end - children_count;
        const size_t unmerged_begin = end - (children_count + unmerged_count);
        next_begin = unmerged_begin;
        next_end = children_begin;
        // merge / compaction, locb.rs:193-215
        for (size_t i = begin; i < end; ++i) {
            const size_t j = neighbours[i];
            if (j < neighbours.size() && size_t(neighbours[j]) == i) {
                if (i < j) {
                    RefAabb& parent = output[unmerged_begin + j - begin - merged[j]];
                    const size_t first_child = children_begin + (size_t(merged[i]) - 1) * 2;
                    parent = union_of(input[j], input[i]);
                    parent.extra1 = -1;                      // set_count(None)
                    parent.extra2 = int32_t(first_child);    // set_left_first
                    output[first_child] = input[i];
                    output[first_child + 1] = input[j];
                }
            } else {
                output[unmerged_begin + i - begin - merged[i]] = input[i];
            }
        }
        // copy the older levels forward, locb.rs:242
        for (size_t k = end; k < previous_end; ++k) output[k] = input[k];
    }
};

size_t build_locb(const RefAabb* aabbs, const float* centers, size_t stride, size_t n,
                  size_t radius, RefAabb* nodes_out, uint32_t* indices_out,
                  uint32_t* iterations) {
    if (iterations) *iterations = 0;
    if (n == 0) return 0;  // Bvh::default(), locb.rs:250-255
    const RefAabb world = union_of_list(aabbs, n);  // locb.rs:19
    if (n <= 2) {
        // single root leaf, locb.rs:258-269
        RefAabb root = world;
        root.extra2 = 0;
        root.extra1 = int32_t(n);
        nodes_out[0] = root;
        for (size_t i = 0; i < n; ++i) indices_out[i] = uint32_t(i);
        return 1;
    }
    // MortonEncoder::get_sorted_indices, morton.rs:63-102 (stable sort)
    const MortonEncoder enc = make_encoder(world);
    std::vector<uint32_t> codes(n);
    for (size_t i = 0; i < n; ++i) codes[i] = encode(enc, centers + i * stride);
    std::vector<uint32_t> prim_indices(n);
    for (size_t i = 0; i < n; ++i) prim_indices[i] = uint32_t(i);
    std::stable_sort(prim_indices.begin(), prim_indices.end(),
                     [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });

    const size_t node_count = 2 * n - 1;
    RefAabb init{};
    init.min[0] = init.min[1] = init.min[2] = 0.f;
    init.max[0] = init.max[1] = init.max[2] = 0.f;
    init.extra1 = -1;
    init.extra2 = 0;
    std::vector<RefAabb> nodes(node_count, init), nodes_copy(node_count, init);
    // locb.rs:283,296-304 slices one buffer as neighbours = aux[0..3N) and
    // merged = aux[N..3N); cluster() only touches neighbours[begin..end) and
    // merged[begin..end) with end <= N, i.e. aux[0..N) and aux[N..2N), which
    // are disjoint, so two separate buffers hold the same values.
    std::vector<uint32_t> neighbours(node_count, uint32_t(node_count));
    std::vector<uint32_t> merged(node_count, uint32_t(node_count));

    size_t begin = node_count - n, end = node_count, previous_end = end;
    for (size_t i = 0; i < n; ++i) {
        RefAabb leaf = aabbs[prim_indices[i]];
        offset_by(leaf, 0.0001f);
        leaf.extra1 = 1;
        leaf.extra2 = int32_t(i);
        nodes[begin + i] = leaf;  // locb.rs:289-294
    }
    Locb b{radius};
    uint32_t iters = 0;
    while (end - begin > 1) {
        size_t nb, ne;
        b.cluster(nodes, nodes_copy, neighbours, merged, begin, end, previous_end, nb, ne);
        std::swap(nodes, nodes_copy);
        previous_end = end;
        begin = nb;
        end = ne;
        ++iters;
        if (end <= begin) break;  // only if no mutual pair exists (NaN bounds)
    }
    if (iterations) *iterations = iters;
    std::memcpy(nodes_out, nodes.data(), node_count * sizeof(RefAabb));
    std::memcpy(indices_out, prim_indices.data(), n * sizeof(uint32_t));
    return node_count;
}

// ---------------------------------------------------- binned SAH builder --
// src/builders/binned_sah.rs

struct SahBin {
    RefAabb aabb;
    uint64_t prim_count;
    float right_cost;
};

struct SahSplit {
    float cost;
    uint32_t count;
};

struct BinnedBuilder {
    const RefAabb* aabbs;
    const float* centers;
    size_t stride;
    size_t max_depth = 64;     // binned_sah.rs:313
    size_t bin_count = 16;     // binned_sah.rs:314
    size_t max_leaf_size = 1;  // binned_sah.rs:315
    float traversal_cost = 1.0f;  // binned_sah.rs:316
    // contained = true fixes two reference defects so that every child box
    // lies inside its parent: the stale left-box count (left box from best[best_axis].count after
    // the fallback replaced best_axis, binned_sah.rs:188,196,232-235) and the
    // double leaf pad (leaves padded a second time, binned_sah.rs:133,135).
    bool contained = false;
    std::vector<RefAabb>* nodes;
    uint32_t* prim_indices;
    size_t counter = 1;  // AtomicNodeStack::new, builders/mod.rs:51-56

    struct Task {
        size_t node, begin, end, depth;
    };

    // compute_bin_index, binned_sah.rs:119-128
    size_t bin_index(const float* c, const float* bo, const float* ctb, int axis) const {
        const float b = c[axis] * ctb[axis] + bo[axis];
        const uint64_t u = sat_usize(rs_max(b, 0.0f));
        return size_t(std::min<uint64_t>(bin_count - 1, u));
    }

    // find_split, binned_sah.rs:80-114
    SahSplit find_split(std::vector<SahBin>& bins) const {
        RefAabb cur = aabb_new();
        uint64_t cur_count = 0;
        for (size_t i = bin_count - 1; i > 0; --i) {
            grow_bb(cur, bins[i].aabb);
            cur_count += bins[i].prim_count;
            bins[i].right_cost = half_area(cur) * float(cur_count);
        }
        cur = aabb_new();
        cur_count = 0;
        SahSplit best{std::numeric_limits<float>::max(), uint32_t(bin_count)};
        for (size_t i = 0; i + 1 < bin_count; ++i) {
            grow_bb(cur, bins[i].aabb);
            cur_count += bins[i].prim_count;
            const float cost = half_area(cur) * float(cur_count) + bins[i + 1].right_cost;
            if (cost < best.cost) best = SahSplit{cost, uint32_t(i + 1)};
        }
        return best;
    }

    void make_leaf(RefAabb& node, size_t begin, size_t end) const {
        if (!contained) offset_by(node, 0.0001f);
        node.extra2 = int32_t(begin);
        node.extra1 = int32_t(end - begin);
    }

    // BinnedSahBuildTask::run, binned_sah.rs:132-282
    bool run(const Task& t, Task& left, Task& right) {
        RefAabb& node = (*nodes)[t.node];
        offset_by(node, 0.0001f);
        const size_t work = t.end - t.begin;
        if (work <= 1 || t.depth >= max_depth) {
            make_leaf(node, t.begin, t.end);
            return false;
        }
        float ctb[3], bo[3];
        for (int k = 0; k < 3; ++k) {
            ctb[k] = (1.0f / (node.max[k] - node.min[k])) * float(bin_count);
            bo[k] = (-node.min[k]) * ctb[k];
        }
        std::vector<SahBin> bins[3];
        for (int a = 0; a < 3; ++a)
            bins[a].assign(bin_count, SahBin{aabb_new(), 0, std::numeric_limits<float>::max()});
        uint32_t* idx = prim_indices + t.begin;
        for (size_t i = 0; i < work; ++i) {
            const uint32_t p = idx[i];
            const float* c = centers + size_t(p) * stride;
            for (int a = 0; a < 3; ++a) {
                const size_t bi = bin_index(c, bo, ctb, a);
                bins[a][bi].prim_count += 1;
                grow_bb(bins[a][bi].aabb, aabbs[p]);
            }
        }
        SahSplit best[3];
        for (int a = 0; a < 3; ++a) best[a] = find_split(bins[a]);
        int best_axis = 0;
        if (best[0].cost > best[1].cost) best_axis = 1;
        if (best[best_axis].cost > best[2].cost) best_axis = 2;

        size_t split_index = best[best_axis].count;
        const float max_split_cost = half_area(node) * (float(work) - traversal_cost);
        if (best[best_axis].count == bin_count || best[best_axis].cost >= max_split_cost) {
            if (work > max_leaf_size) {
                // approximate median on the longest axis, binned_sah.rs:194-205
                best_axis = longest_axis(node);
                uint64_t count = 0;
                for (size_t i = 0; i + 1 < bin_count; ++i) {
                    count += bins[best_axis][i].prim_count;
                    if (count >= uint64_t(work * 2 / 5 + 1)) {
                        split_index = i + 1;
                        break;
                    }
                }
            } else {
                make_leaf(node, t.begin, t.end);
                return false;
            }
        }
        // partition, utils.rs:76-96 (Lomuto; left side keeps its order)
        size_t count_left = 0;
        for (size_t i = 0; i < work; ++i) {
            const float* c = centers + size_t(idx[i]) * stride;
            if (bin_index(c, bo, ctb, best_axis) < split_index) {
                std::swap(idx[i], idx[count_left]);
                ++count_left;
            }
        }
        const size_t begin_right = t.begin + count_left;
        if (begin_right > t.begin && begin_right < t.end) {
            const size_t l = counter;  // AtomicNodeStack::allocate, builders/mod.rs:59-76
            counter += 2;
            const size_t r = l + 1;
            node.extra2 = int32_t(l);
            node.extra1 = -1;
            RefAabb left_box = aabb_new(), right_box = aabb_new();
            // Stale left-box count: the left box spans bins [0, best[best_axis].count) where
            // best_axis may have been replaced by the fallback axis.
            const size_t left_end = contained ? split_index : size_t(best[best_axis].count);
            for (size_t i = 0; i < left_end; ++i) grow_bb(left_box, bins[best_axis][i].aabb);
            for (size_t i = split_index; i < bin_count; ++i)
                grow_bb(right_box, bins[best_axis][i].aabb);
            (*nodes)[l] = left_box;
            (*nodes)[r] = right_box;
            left = Task{l, t.begin, begin_right, t.depth + 1};
            right = Task{r, begin_right, t.end, t.depth + 1};
            return true;
        }
        make_leaf(node, t.begin, t.end);
        return false;
    }
};

size_t build_binned(const RefAabb* aabbs, const float* centers, size_t stride, size_t n,
                    size_t prims_per_leaf, bool contained, RefAabb* nodes_out,
                    uint32_t* indices_out) {
    if (n == 0) return 0;
    std::vector<RefAabb> nodes(2 * n - 1, aabb_new());  // BvhNode::new()
    for (size_t i = 0; i < n; ++i) indices_out[i] = uint32_t(i);
    BinnedBuilder b;
    b.aabbs = aabbs;
    b.centers = centers;
    b.stride = stride;
    b.max_leaf_size = prims_per_leaf == 0 ? 1 : prims_per_leaf;  // NonZeroUsize, binned_sah.rs:315
    b.contained = contained;
    b.nodes = &nodes;
    b.prim_indices = indices_out;
    nodes[0] = union_of_list(aabbs, n);  // binned_sah.rs:361
    // TaskSpawner::run_task (utils.rs:243-288) on one thread: the larger child
    // is pushed last, so it is run first.
    std::vector<BinnedBuilder::Task> stack;
    stack.push_back(BinnedBuilder::Task{0, 0, n, 0});
    while (!stack.empty()) {
        BinnedBuilder::Task t = stack.back();
        stack.pop_back();
        BinnedBuilder::Task a, c;
        if (b.run(t, a, c)) {
            if ((a.end - a.begin) < (c.end - c.begin)) std::swap(a, c);
            stack.push_back(c);
            stack.push_back(a);
        }
    }
    std::memcpy(nodes_out, nodes.data(), b.counter * sizeof(RefAabb));
    return b.counter;
}

// -------------------------------------------------------- MBVH collapse --
inline RefMbvhNode mbvh_new() {
    RefMbvhNode m;
    for (int i = 0; i < 4; ++i) {
        m.min_x[i] = m.min_y[i] = m.min_z[i] = 1e34f;
        m.max_x[i] = m.max_y[i] = m.max_z[i] = -1e34f;
        m.children[i] = -1;
        m.counts[i] = -1;
    }
    return m;
}

inline void set_bounds_bb(RefMbvhNode& m, int i, const RefAabb& b) {
    m.min_x[i] = b.min[0];
    m.min_y[i] = b.min[1];
    m.min_z[i] = b.min[2];
    m.max_x[i] = b.max[0];
    m.max_y[i] = b.max[1];
    m.max_z[i] = b.max[2];
}

// MbvhNode::merge_nodes, src/mbvh_node.rs:297-411.
void merge_nodes(size_t m_index, size_t cur, const RefAabb* pool, size_t pool_len,
                 RefMbvhNode* mpool, size_t& pool_ptr) {
    for (int i = 0; i < 4; ++i) set_bounds_bb(mpool[m_index], i, pool[cur]);
    int32_t nodes[4] = {-1, -1, -1, -1};
    int32_t leafs[4] = {-1, -1, -1, -1};
    const int32_t cur_lf = pool[cur].extra2;
    if (cur_lf >= 0) {
        for (int side = 0; side < 2; ++side) {
            const size_t child = size_t(cur_lf) + size_t(side);
            const int s0 = side * 2, s1 = side * 2 + 1;
            if (child >= pool_len) continue;  // bvh_pool.get(..) is None
            const RefAabb& c = pool[child];
            if (c.extra2 < 0) continue;  // get_left_first() is None
            const int32_t lf = c.extra2;
            if (c.extra1 >= 0) {
                // leaf child: one slot keeps the parent's bounds (mbvh_node.rs:305,319-321)
                nodes[s0] = lf;
                leafs[s0] = c.extra1;
            } else {
                const RefAabb& g0 = pool[size_t(lf)];
                const RefAabb& g1 = pool[size_t(lf) + 1];
                if (g0.extra1 >= 0) {
                    nodes[s0] = g0.extra2;
                    leafs[s0] = g0.extra1;
                } else {
                    nodes[s0] = lf;
                }
                set_bounds_bb(mpool[m_index], s0, g0);
                if (g1.extra1 >= 0) {
                    nodes[s1] = g1.extra2;
                    leafs[s1] = g1.extra1;
                } else {
                    nodes[s1] = lf + 1;
                }
                set_bounds_bb(mpool[m_index], s1, g1);
            }
        }
    }
    for (int i = 0; i < 4; ++i) {
        const int32_t node = nodes[i];
        const int32_t count = leafs[i];
        if (node >= 0 && count >= 0) {
            mpool[m_index].children[i] = node;
            mpool[m_index].counts[i] = count;
            continue;
        } else if (node == -1) {
            continue;
        }
        const RefAabb& bn = pool[size_t(node)];
        if (bn.extra1 >= 0) {
            mpool[m_index].children[i] = bn.extra2;
            mpool[m_index].counts[i] = bn.extra1;
            set_bounds_bb(mpool[m_index], i, bn);
        } else {
            const size_t new_m = pool_ptr;
            pool_ptr += 1;
            mpool[m_index].children[i] = int32_t(new_m);
            set_bounds_bb(mpool[m_index], i, bn);
            merge_nodes(new_m, size_t(node), pool, pool_len, mpool, pool_ptr);
        }
    }
}

// ------------------------------------------------------------- checkers --
bool contains_box(const RefAabb& outer, const RefAabb& inner) {
    for (int k = 0; k < 3; ++k) {
        if (!(outer.min[k] <= inner.min[k])) return false;
        if (!(outer.max[k] >= inner.max[k])) return false;
    }
    return true;
}

}  // namespace

// =================================================================== C API ==
extern "C" {

uint32_t ref_morton_split(uint32_t v) { return morton_split(v); }

void ref_world_bounds(const RefAabb* aabbs, size_t n, RefAabb* out) {
    *out = union_of_list(aabbs, n);
}

void ref_morton_codes(const RefAabb* world, const float* centers, size_t stride, size_t n,
                      uint32_t* codes) {
    const MortonEncoder e = make_encoder(*world);
    for (size_t i = 0; i < n; ++i) codes[i] = encode(e, centers + i * stride);
}

void ref_sorted_indices(const uint32_t* codes, size_t n, uint32_t* indices) {
    for (size_t i = 0; i < n; ++i) indices[i] = uint32_t(i);
    std::stable_sort(indices, indices + n, [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
}

// prefix_sum, src/utils.rs:42-57 (count 0 returns first[0]).
uint32_t ref_prefix_sum_u32(const uint32_t* first, size_t count, uint32_t* output) {
    if (count == 0) return first[0];
    uint32_t sum = 0;
    for (size_t i = 0; i < count; ++i) {
        sum += first[i];
        output[i] = sum;
    }
    return sum;
}

float ref_half_area(const RefAabb* a) { return half_area(*a); }

// shared/src/lib.rs:27-38: center = (v0 + v1 + v2) * (1/3); aabb grown from
// the three vertices without padding.
void ref_triangle_aabbs_centers(const float* verts, size_t n, RefAabb* aabbs, float* centers) {
    const float third = 1.0f / 3.0f;
    for (size_t i = 0; i < n; ++i) {
        const float* p = verts + i * 9;
        RefAabb a = aabb_new();
        grow_pt(a, p);
        grow_pt(a, p + 3);
        grow_pt(a, p + 6);
        aabbs[i] = a;
        for (int k = 0; k < 3; ++k) centers[i * 3 + k] = ((p[k] + p[3 + k]) + p[6 + k]) * third;
    }
}

size_t ref_build_locb(const RefAabb* aabbs, const float* centers, size_t stride, size_t n,
                      RefAabb* nodes, uint32_t* indices) {
    return build_locb(aabbs, centers, stride, n, 14, nodes, indices, nullptr);  // locb.rs:27
}

size_t ref_build_locb_radius(const RefAabb* aabbs, const float* centers, size_t stride,
                             size_t n, size_t radius, RefAabb* nodes, uint32_t* indices,
                             uint32_t* iterations) {
    return build_locb(aabbs, centers, stride, n, radius, nodes, indices, iterations);
}

size_t ref_build_binned_sah(const RefAabb* aabbs, const float* centers, size_t stride, size_t n,
                            size_t prims_per_leaf, RefAabb* nodes, uint32_t* indices) {
    return build_binned(aabbs, centers, stride, n, prims_per_leaf, false, nodes, indices);
}

size_t ref_build_binned_sah_contained(const RefAabb* aabbs, const float* centers, size_t stride,
                                      size_t n, size_t prims_per_leaf, RefAabb* nodes,
                                      uint32_t* indices) {
    return build_binned(aabbs, centers, stride, n, prims_per_leaf, true, nodes, indices);
}

// Mbvh::construct, src/bvh.rs:381-404.
size_t ref_build_mbvh(const RefAabb* nodes, size_t node_count, RefMbvhNode* m_nodes) {
    if (node_count == 0) return 0;
    for (size_t i = 0; i < node_count; ++i) m_nodes[i] = mbvh_new();
    size_t pool_ptr = 1;
    merge_nodes(0, 0, nodes, node_count, m_nodes, pool_ptr);
    return pool_ptr;
}

// Bvh::refit, src/bvh.rs:176-205, children before parents (reverse index
// order).  Deliberate difference: the reference assigns a fresh Aabb::new()
// (count = left_first = 0) to every node, which turns every node into an empty
// leaf; count and left_first are kept here.
void ref_refit(RefAabb* nodes, size_t node_count, const uint32_t* indices, size_t index_count,
               const RefAabb* aabbs) {
    for (size_t ii = node_count; ii-- > 0;) {
        RefAabb a = aabb_new();
        const int32_t lf = nodes[ii].extra2;
        const int32_t count = nodes[ii].extra1;
        if (lf >= 0) {
            if (count >= 0) {
                for (int32_t k = 0; k < count; ++k) {
                    const size_t pos = size_t(lf) + size_t(k);
                    if (pos >= index_count) break;
                    grow_bb(a, aabbs[indices[pos]]);
                }
            } else {
                grow_bb(a, nodes[lf]);
                grow_bb(a, nodes[size_t(lf) + 1]);
            }
            offset_by(a, 0.0001f);
        }
        a.extra1 = count;
        a.extra2 = lf;
        nodes[ii] = a;
    }
}

int ref_mt(const float* v9, const float o[3], const float d[3], float t_min, float t_cur, float* t,
           float* u, float* v) {
    MT m = mt_test(v9, o, d, t_min, t_cur);
    if (!m.ok) return 0;
    *t = m.t;
    *u = m.u;
    *v = m.v;
    return 1;
}

void ref_trace_mbvh(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                    size_t index_count, const float* verts, size_t prim_count,
                    const RefRay* rays, size_t n, RefHit* hits, int mode, int threads,
                    uint64_t* stats) {
    std::atomic<uint64_t> nv{0}, tt{0};
    parallel_rays(n, threads, [&](size_t b, size_t e, int) {
        std::vector<int32_t> stack;
        std::vector<std::pair<int32_t, float>> fstack;
        stack.reserve(256);
        fstack.reserve(256);
        uint64_t lnv = 0, ltt = 0;  // per-chunk counters: no shared cache lines
        for (size_t i = b; i < e; ++i) {
            if (mode == REF_MODE_VISIT_ORDER)
                trace_mbvh_one(m, m_count, indices, index_count, verts, prim_count, rays[i],
                               hits[i], mode, lnv, ltt, stack);
            else
                trace_mbvh_lowest_one(m, m_count, indices, index_count, verts, prim_count,
                                      rays[i], hits[i], lnv, ltt, fstack);
        }
        nv += lnv;
        tt += ltt;
    });
    if (stats) {
        stats[0] = nv.load();
        stats[1] = tt.load();
    }
}

void ref_trace_bvh(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                   size_t index_count, const float* verts, size_t prim_count,
                   const RefRay* rays, size_t n, RefHit* hits, int mode, int threads,
                   uint64_t* stats) {
    std::atomic<uint64_t> nv{0}, tt{0};
    parallel_rays(n, threads, [&](size_t b, size_t e, int) {
        std::vector<int32_t> stack;
        stack.reserve(256);
        uint64_t lnv = 0, ltt = 0;
        for (size_t i = b; i < e; ++i)
            trace_bvh_one(nodes, node_count, indices, index_count, verts, prim_count, rays[i],
                          hits[i], mode, lnv, ltt, stack);
        nv += lnv;
        tt += ltt;
    });
    if (stats) {
        stats[0] = nv.load();
        stats[1] = tt.load();
    }
}

// Brute force over every triangle in ascending id with strict '<': the
// lexicographic minimum of (t, prim).
void ref_trace_brute(const float* verts, size_t prim_count, const RefRay* rays, size_t n,
                     RefHit* hits, int threads) {
    parallel_rays(n, threads, [&](size_t b, size_t e, int) {
        for (size_t i = b; i < e; ++i) {
            Ray ray = make_ray(rays[i]);
            RefHit h{ray.t, 0.f, 0.f, 0xFFFFFFFFu};
            if (!(is_nan3(ray.o) || is_nan3(ray.d))) {
                for (size_t p = 0; p < prim_count; ++p) {
                    MT m = mt_test(verts + p * 9, ray.o, ray.d, ray.t_min, h.t);
                    if (m.ok) h = RefHit{m.t, m.u, m.v, uint32_t(p)};
                }
            }
            hits[i] = h;
        }
    });
}

void ref_occluded_mbvh(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                       size_t index_count, const float* verts, size_t prim_count,
                       const RefRay* rays, size_t n, uint8_t* occluded, int threads) {
    parallel_rays(n, threads, [&](size_t b, size_t e, int) {
        std::vector<int32_t> stack;
        stack.reserve(256);
        for (size_t i = b; i < e; ++i)
            occluded[i] = occluded_mbvh_one(m, m_count, indices, index_count, verts, prim_count,
                                            rays[i], stack)
                              ? 1
                              : 0;
    });
}

void ref_occluded_bvh(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                      size_t index_count, const float* verts, size_t prim_count,
                      const RefRay* rays, size_t n, uint8_t* occluded, int threads) {
    parallel_rays(n, threads, [&](size_t b, size_t e, int) {
        std::vector<int32_t> stack;
        stack.reserve(256);
        for (size_t i = b; i < e; ++i)
            occluded[i] = occluded_bvh_one(nodes, node_count, indices, index_count, verts,
                                           prim_count, rays[i], stack)
                              ? 1
                              : 0;
    });
}

void ref_occluded_brute(const float* verts, size_t prim_count, const RefRay* rays, size_t n,
                        uint8_t* occluded, int threads) {
    parallel_rays(n, threads, [&](size_t b, size_t e, int) {
        for (size_t i = b; i < e; ++i) {
            Ray ray = make_ray(rays[i]);
            uint8_t occ = 0;
            if (!(is_nan3(ray.o) || is_nan3(ray.d))) {
                for (size_t p = 0; p < prim_count && !occ; ++p) {
                    MT m = mt_test(verts + p * 9, ray.o, ray.d, ray.t_min, ray.t);
                    if (m.ok) occ = 1;
                }
            }
            occluded[i] = occ;
        }
    });
}

int ref_check_bvh(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                  size_t index_count, size_t prim_count, uint64_t* contain_violations,
                  uint32_t* max_depth) {
    if (contain_violations) *contain_violations = 0;
    if (max_depth) *max_depth = 0;
    if (node_count == 0) return prim_count == 0 ? 0 : 1;
    std::vector<uint32_t> seen(prim_count, 0);
    std::vector<uint8_t> visited(node_count, 0);
    std::vector<std::pair<int32_t, uint32_t>> stack;
    stack.push_back({0, 0});
    uint64_t viol = 0;
    uint32_t maxd = 0;
    while (!stack.empty()) {
        auto [ni, depth] = stack.back();
        stack.pop_back();
        if (ni < 0 || size_t(ni) >= node_count) return 2;
        if (visited[ni]) return 3;  // a node reached twice: not a tree
        visited[ni] = 1;
        maxd = std::max(maxd, depth);
        const RefAabb& nd = nodes[ni];
        if (nd.extra1 >= 0) {
            if (nd.extra2 < 0) return 4;
            for (int32_t k = 0; k < nd.extra1; ++k) {
                const size_t pos = size_t(nd.extra2) + size_t(k);
                if (pos >= index_count) return 5;
                const uint32_t p = indices[pos];
                if (p >= prim_count) return 6;
                seen[p] += 1;
            }
        } else {
            if (nd.extra2 < 0 || size_t(nd.extra2) + 1 >= node_count) return 7;
            for (int c = 0; c < 2; ++c) {
                const int32_t ci = nd.extra2 + c;
                if (!contains_box(nd, nodes[ci])) ++viol;
                stack.push_back({ci, depth + 1});
            }
        }
    }
    if (contain_violations) *contain_violations = viol;
    if (max_depth) *max_depth = maxd;
    for (size_t p = 0; p < prim_count; ++p)
        if (seen[p] != 1) return 8;
    return 0;
}

int ref_check_mbvh(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                   size_t index_count, size_t prim_count, uint32_t* max_depth) {
    if (max_depth) *max_depth = 0;
    if (m_count == 0) return prim_count == 0 ? 0 : 1;
    std::vector<uint32_t> seen(prim_count, 0);
    std::vector<uint8_t> visited(m_count, 0);
    std::vector<std::pair<int32_t, uint32_t>> stack;
    stack.push_back({0, 1});
    uint32_t maxd = 0;
    while (!stack.empty()) {
        auto [ni, depth] = stack.back();
        stack.pop_back();
        if (ni < 0 || size_t(ni) >= m_count) return 2;
        if (visited[ni]) return 3;
        visited[ni] = 1;
        maxd = std::max(maxd, depth);
        const RefMbvhNode& nd = m[ni];
        for (int i = 0; i < 4; ++i) {
            if (nd.counts[i] >= 0) {
                for (int32_t k = 0; k < nd.counts[i]; ++k) {
                    const size_t pos = size_t(nd.children[i]) + size_t(k);
                    if (nd.children[i] < 0 || pos >= index_count) return 5;
                    const uint32_t p = indices[pos];
                    if (p >= prim_count) return 6;
                    seen[p] += 1;
                }
            } else if (nd.children[i] >= 0) {
                stack.push_back({nd.children[i], depth + 1});
            }
        }
    }
    if (max_depth) *max_depth = maxd;
    for (size_t p = 0; p < prim_count; ++p)
        if (seen[p] != 1) return 8;
    return 0;
}

double ref_sah_cost(const RefAabb* nodes, size_t node_count) {
    if (node_count == 0) return 0.0;
    auto ha = [](const RefAabb& a) {
        const double d0 = double(a.max[0]) - a.min[0];
        const double d1 = double(a.max[1]) - a.min[1];
        const double d2 = double(a.max[2]) - a.min[2];
        return (d0 + d1) * d2 + d0 * d1;
    };
    const double root = ha(nodes[0]);
    double cost = 0.0;
    std::vector<int32_t> stack{0};
    while (!stack.empty()) {
        const int32_t ni = stack.back();
        stack.pop_back();
        if (ni < 0 || size_t(ni) >= node_count) continue;
        const RefAabb& nd = nodes[ni];
        if (nd.extra1 >= 0) {
            cost += ha(nd) * double(nd.extra1);
        } else {
            cost += ha(nd);
            if (nd.extra2 >= 0) {
                stack.push_back(nd.extra2);
                stack.push_back(nd.extra2 + 1);
            }
        }
    }
    return root > 0 ? cost / root : cost;
}

static bool bits_equal(const float* a, const float* b, int n) {
    return std::memcmp(a, b, sizeof(float) * size_t(n)) == 0;
}

int ref_compare_bvh(const RefAabb* a, size_t na, const uint32_t* ia, const RefAabb* b, size_t nb,
                    const uint32_t* ib, int as_sets) {
    if (na == 0 || nb == 0) return (na == nb) ? 0 : 1;
    std::vector<std::pair<int32_t, int32_t>> stack{{0, 0}};
    while (!stack.empty()) {
        auto [x, y] = stack.back();
        stack.pop_back();
        if (x < 0 || y < 0 || size_t(x) >= na || size_t(y) >= nb) return 2;
        const RefAabb& p = a[x];
        const RefAabb& q = b[y];
        if (!bits_equal(p.min, q.min, 3) || !bits_equal(p.max, q.max, 3)) return 3;
        if ((p.extra1 >= 0) != (q.extra1 >= 0)) return 4;
        if (p.extra1 >= 0) {
            if (p.extra1 != q.extra1) return 5;
            std::vector<uint32_t> la(ia + p.extra2, ia + p.extra2 + p.extra1);
            std::vector<uint32_t> lb(ib + q.extra2, ib + q.extra2 + q.extra1);
            if (as_sets) {
                std::sort(la.begin(), la.end());
                std::sort(lb.begin(), lb.end());
            }
            if (la != lb) return 6;
        } else {
            stack.push_back({p.extra2, q.extra2});
            stack.push_back({p.extra2 + 1, q.extra2 + 1});
        }
    }
    return 0;
}

int ref_compare_mbvh(const RefMbvhNode* a, size_t na, const uint32_t* ia, const RefMbvhNode* b,
                     size_t nb, const uint32_t* ib) {
    if (na == 0 || nb == 0) return (na == nb) ? 0 : 1;
    std::vector<std::pair<int32_t, int32_t>> stack{{0, 0}};
    while (!stack.empty()) {
        auto [x, y] = stack.back();
        stack.pop_back();
        if (x < 0 || y < 0 || size_t(x) >= na || size_t(y) >= nb) return 2;
        const RefMbvhNode& p = a[x];
        const RefMbvhNode& q = b[y];
        if (!bits_equal(p.min_x, q.min_x, 24)) return 3;  // all six bound arrays
        for (int i = 0; i < 4; ++i) {
            if (p.counts[i] != q.counts[i]) return 4;
            if (p.counts[i] >= 0) {
                for (int32_t k = 0; k < p.counts[i]; ++k)
                    if (ia[p.children[i] + k] != ib[q.children[i] + k]) return 6;
            } else if ((p.children[i] >= 0) != (q.children[i] >= 0)) {
                return 5;
            } else if (p.children[i] >= 0) {
                stack.push_back({p.children[i], q.children[i]});
            }
        }
    }
    return 0;
}

}  // extern "C"
