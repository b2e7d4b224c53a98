// C ABI: rtbvh_ffi's callback traversal (intersect, intersect_packet,
// intersect_mbvh, intersect_mbvh_packet; rtbvh_ffi/src/lib.rs:540-835).
//
// These entry points hand every candidate primitive to a host callback,
// which may lower t and may stop the walk; a host callback cannot run inside
// a GPU kernel, so this walk runs on the calling thread over the handle's
// node and index arrays (host memory for handles returned by create_bvh /
// create_mbvh).  The device path for bulk rays is rtbvh_gpu_intersect*.
//
// Visit order is the reference's index iterators (src/iter_indices.rs):
//   BvhIndexIterator        :66-107   root without a box test, far child first
//   BvhPacketIndexIterator  :160-209  any-lane box tests (Aabb::intersect4)
//   MbvhIndexIterator       :264-312  slot masks taken at pop, ids[3]..ids[0]
//   MbvhPacketIndexIterator :366-414  unsorted slots, any-lane mask
// with the reference's 32-entry stacks replaced by growable ones.
#include <math.h>

#include <vector>

#include "internal.hpp"

namespace rtbvh {
namespace {

typedef bool (*Callback)(uint32_t, float*, void*);

inline float pmin(float a, float b) { return a < b ? a : b; }  // _mm_min_ps lane
inline float pmax(float a, float b) { return a > b ? a : b; }  // _mm_max_ps lane

struct Ray1 {
    float o[3], d[3], inv[3];
    int sign[3];
    float t_min, t;
};

Ray1 make_ray1(const float* o, const float* d, float t) {
    Ray1 r;
    for (int k = 0; k < 3; ++k) {
        r.o[k] = o[k];
        r.d[k] = d[k];
        r.inv[k] = 1.0f / d[k];  // Ray::new, src/ray.rs:166-182
        r.sign[k] = d[k] < 0.0f ? 1 : 0;
    }
    r.t_min = 1e-4f;
    r.t = t;  // rtbvh_ffi/src/lib.rs:565-566
    return r;
}

// Aabb::intersect (src/aabb.rs:146-181): Some(t_far) when hit.
bool slab1(const RTAabb& b, const Ray1& r, float* t_far) {
    const float* lo[2] = {b.min, b.max};
    float ray_min = (lo[r.sign[0]][0] - r.o[0]) * r.inv[0];
    float ray_max = (lo[1 - r.sign[0]][0] - r.o[0]) * r.inv[0];
    const float y_min = (lo[r.sign[1]][1] - r.o[1]) * r.inv[1];
    const float y_max = (lo[1 - r.sign[1]][1] - r.o[1]) * r.inv[1];
    if ((ray_min > y_max) || (y_min > ray_max)) return false;
    if (y_min > ray_min) ray_min = y_min;
    if (y_max < ray_max) ray_max = y_max;
    const float z_min = (lo[r.sign[2]][2] - r.o[2]) * r.inv[2];
    const float z_max = (lo[1 - r.sign[2]][2] - r.o[2]) * r.inv[2];
    if ((ray_min > z_max) || (z_min > ray_max)) return false;
    if (z_max < ray_max) ray_max = z_max;
    if (ray_max > r.t_min) {
        *t_far = ray_max;
        return true;
    }
    return false;
}

struct Packet {
    float ox[4], oy[4], oz[4], dx[4], dy[4], dz[4], ix[4], iy[4], iz[4];
    float t[4];
};

// Aabb::intersect4 (src/aabb.rs:218-244): per-lane t_min when any lane hits.
bool slab4(const RTAabb& b, const Packet& p, float tmin_out[4]) {
    bool any = false;
    for (int l = 0; l < 4; ++l) {
        const float t1x = (b.min[0] - p.ox[l]) * p.ix[l];
        const float t1y = (b.min[1] - p.oy[l]) * p.iy[l];
        const float t1z = (b.min[2] - p.oz[l]) * p.iz[l];
        const float t2x = (b.max[0] - p.ox[l]) * p.ix[l];
        const float t2y = (b.max[1] - p.oy[l]) * p.iy[l];
        const float t2z = (b.max[2] - p.oz[l]) * p.iz[l];
        const float tn = pmax(pmin(t1x, t2x), pmax(pmin(t1y, t2y), pmin(t1z, t2z)));
        const float tf = pmin(pmax(t1x, t2x), pmin(pmax(t1y, t2y), pmax(t1z, t2z)));
        tmin_out[l] = tn;
        if ((tf > 0.0f) && (tf > tn) && (tn < p.t[l])) any = true;
    }
    return any;
}

// MbvhNode::intersect (src/mbvh_node.rs:177-240).
void mbvh_hit1(const RTMbvhNode& n, const Ray1& r, int ids[4], bool res[4]) {
    float tmin[4];
    for (int i = 0; i < 4; ++i) {
        const float tx0 = (n.min_x[i] - r.o[0]) * r.inv[0];
        const float tx1 = (n.max_x[i] - r.o[0]) * r.inv[0];
        const float ty0 = (n.min_y[i] - r.o[1]) * r.inv[1];
        const float ty1 = (n.max_y[i] - r.o[1]) * r.inv[1];
        const float tz0 = (n.min_z[i] - r.o[2]) * r.inv[2];
        const float tz1 = (n.max_z[i] - r.o[2]) * r.inv[2];
        const float tn = pmax(pmin(tx0, tx1), pmax(pmin(ty0, ty1), pmin(tz0, tz1)));
        const float tf = pmin(pmax(tx0, tx1), pmin(pmax(ty0, ty1), pmax(tz0, tz1)));
        res[i] = (tf >= tn) && (tn < r.t);
        tmin[i] = tn;
        ids[i] = i;
    }
    auto cs = [&](int a, int b, bool swap_t) {
        if (tmin[a] > tmin[b]) {
            if (swap_t) {
                const float t = tmin[a];
                tmin[a] = tmin[b];
                tmin[b] = t;
            }
            const int x = ids[a];
            ids[a] = ids[b];
            ids[b] = x;
        }
    };
    cs(0, 1, true);
    cs(2, 3, true);
    cs(0, 2, true);
    cs(1, 3, true);
    cs(2, 3, false);
}

// MbvhNode::intersect4 (src/mbvh_node.rs:243-295).
void mbvh_hit4(const RTMbvhNode& n, const Packet& p, bool res[4]) {
    for (int s = 0; s < 4; ++s) res[s] = false;
    for (int l = 0; l < 4; ++l) {
        for (int s = 0; s < 4; ++s) {
            float t1 = (n.min_x[s] - p.ox[l]) * p.ix[l];
            float t2 = (n.max_x[s] - p.ox[l]) * p.ix[l];
            float tn = pmin(t1, t2);
            float tf = pmax(t1, t2);
            t1 = (n.min_y[s] - p.oy[l]) * p.iy[l];
            t2 = (n.max_y[s] - p.oy[l]) * p.iy[l];
            tn = pmax(tn, pmin(t1, t2));
            tf = pmin(tf, pmax(t1, t2));
            t1 = (n.min_z[s] - p.oz[l]) * p.iz[l];
            t2 = (n.max_z[s] - p.oz[l]) * p.iz[l];
            tn = pmax(tn, pmin(t1, t2));
            tf = pmin(tf, pmax(t1, t2));
            if ((tf > tn) && (tn < p.t[l])) res[s] = true;
        }
    }
}

bool any_nan(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (v[i] != v[i]) return true;
    return false;
}

// Walks a BVH, calling `visit(prim) -> stop` per candidate.
template <class Test, class Visit>
void walk_bvh(const RTBvhNode* nodes, uint32_t node_count, const uint32_t* indices,
              uint32_t index_count, Test&& test, Visit&& visit) {
    if (node_count == 0) return;
    std::vector<int32_t> stack;
    stack.reserve(64);
    stack.push_back(0);
    while (!stack.empty()) {
        const int32_t ni = stack.back();
        stack.pop_back();
        if (ni < 0 || uint32_t(ni) >= node_count) continue;
        const RTAabb& nd = nodes[ni].aabb;
        if (nd.count > -1) {
            for (int32_t j = 0; j < nd.count; ++j) {
                const uint32_t pos = uint32_t(nd.left_first) + uint32_t(j);
                if (pos >= index_count) break;
                if (visit(indices[pos])) return;
            }
        } else if (nd.left_first > -1) {
            const int32_t lf = nd.left_first;
            if (uint32_t(lf) + 1 >= node_count) continue;
            int order = 0;  // 0: neither, 1: left, 2: right, 3: both L-then-R, 4: both R-then-L
            test(nodes[lf].aabb, nodes[lf + 1].aabb, &order);
            switch (order) {
                case 1: stack.push_back(lf); break;
                case 2: stack.push_back(lf + 1); break;
                case 3:
                    stack.push_back(lf);
                    stack.push_back(lf + 1);
                    break;
                case 4:
                    stack.push_back(lf + 1);
                    stack.push_back(lf);
                    break;
                default: break;
            }
        }
    }
}

template <class Hit, class Visit>
void walk_mbvh(const RTMbvhNode* nodes, uint32_t node_count, const uint32_t* indices,
               uint32_t index_count, Hit&& hit, Visit&& visit) {
    if (node_count == 0) return;
    std::vector<int32_t> stack;
    stack.reserve(64);
    int32_t cur = 0;
    for (;;) {
        const RTMbvhNode& n = nodes[cur];
        int ids[4];
        bool res[4];
        hit(n, ids, res);
        for (int i = 0; i < 4; ++i) {
            const int id = ids[3 - i];
            if (!res[id]) continue;
            const int32_t count = n.counts[id];
            const int32_t lf = n.children[id];
            if (count > -1) {
                for (int32_t j = 0; j < count; ++j) {
                    const uint32_t pos = uint32_t(lf) + uint32_t(j);
                    if (pos >= index_count) break;
                    if (visit(indices[pos])) return;
                }
            } else if (lf > -1) {
                stack.push_back(lf);
            }
        }
        if (stack.empty()) return;
        cur = stack.back();
        stack.pop_back();
        if (uint32_t(cur) >= node_count) return;
    }
}

Packet make_packet(const float* ox, const float* oy, const float* oz, const float* dx,
                   const float* dy, const float* dz, const float* t) {
    Packet p;
    for (int l = 0; l < 4; ++l) {
        p.ox[l] = ox[l];
        p.oy[l] = oy[l];
        p.oz[l] = oz[l];
        p.dx[l] = dx[l];
        p.dy[l] = dy[l];
        p.dz[l] = dz[l];
        p.ix[l] = 1.0f / dx[l];
        p.iy[l] = 1.0f / dy[l];
        p.iz[l] = 1.0f / dz[l];
        p.t[l] = t[l];
    }
    return p;
}

}  // namespace
}  // namespace rtbvh

using namespace rtbvh;

extern "C" {

ResultCode intersect(RTBvh bvh, const float* origin, const float* direction, float* t,
                     void* user_data, bool (*cb)(uint32_t, float*, void*)) {
    if (!origin || !direction || !t || !cb) {
        set_error("null origin, direction, t or callback");
        return Error;
    }
    if (any_nan(origin, 3) || any_nan(direction, 3)) return Nan;  // lib.rs:561-563
    Ray1 r = make_ray1(origin, direction, *t);
    walk_bvh(
        bvh.nodes, bvh.nodes ? bvh.node_count : 0, bvh.indices, bvh.indices ? bvh.index_count : 0,
        [&](const RTAabb& l, const RTAabb& rr, int* order) {
            float tl = 0.f, tr = 0.f;
            const bool hl = slab1(l, r, &tl), hr = slab1(rr, r, &tr);
            *order = (hl && hr) ? (tl < tr ? 3 : 4) : hl ? 1 : hr ? 2 : 0;
        },
        [&](uint32_t prim) { return cb(prim, &r.t, user_data); });
    *t = r.t;
    return Ok;
}

ResultCode intersect_packet(RTBvh bvh, const float* origin_x, const float* origin_y,
                            const float* origin_z, const float* direction_x,
                            const float* direction_y, const float* direction_z, float* t,
                            void* user_data, bool (*cb)(uint32_t, float*, void*)) {
    if (!origin_x || !origin_y || !origin_z || !direction_x || !direction_y || !direction_z ||
        !t || !cb) {
        set_error("null packet array or callback");
        return Error;
    }
    if (any_nan(origin_x, 4) || any_nan(origin_y, 4) || any_nan(origin_z, 4) ||
        any_nan(direction_x, 4) || any_nan(direction_y, 4) || any_nan(direction_z, 4))
        return Nan;  // lib.rs:650-658
    Packet p = make_packet(origin_x, origin_y, origin_z, direction_x, direction_y, direction_z, t);
    walk_bvh(
        bvh.nodes, bvh.nodes ? bvh.node_count : 0, bvh.indices, bvh.indices ? bvh.index_count : 0,
        [&](const RTAabb& l, const RTAabb& rr, int* order) {
            float tl[4], tr[4];
            const bool hl = slab4(l, p, tl), hr = slab4(rr, p, tr);
            if (hl && hr) {
                // BvhNode::sort_nodes4 (src/bvh_node.rs:180-211)
                bool lt = false;
                for (int k = 0; k < 4; ++k) lt = lt || (tl[k] < tr[k]);
                *order = lt ? 3 : 4;
            } else {
                *order = hl ? 1 : hr ? 2 : 0;
            }
        },
        [&](uint32_t prim) { return cb(prim, p.t, user_data); });
    for (int l = 0; l < 4; ++l) t[l] = p.t[l];
    return Ok;
}

ResultCode intersect_mbvh(RTMbvh bvh, const float* origin, const float* direction, float* t,
                          void* user_data, bool (*cb)(uint32_t, float*, void*)) {
    if (!origin || !direction || !t || !cb) {
        set_error("null origin, direction, t or callback");
        return Error;
    }
    if (any_nan(origin, 3) || any_nan(direction, 3)) return Nan;  // lib.rs:711-713
    Ray1 r = make_ray1(origin, direction, *t);
    walk_mbvh(
        bvh.nodes, bvh.nodes ? bvh.node_count : 0, bvh.indices, bvh.indices ? bvh.index_count : 0,
        [&](const RTMbvhNode& n, int ids[4], bool res[4]) { mbvh_hit1(n, r, ids, res); },
        [&](uint32_t prim) { return cb(prim, &r.t, user_data); });
    *t = r.t;
    return Ok;
}

ResultCode intersect_mbvh_packet(RTMbvh bvh, const float* origin_x, const float* origin_y,
                                 const float* origin_z, const float* direction_x,
                                 const float* direction_y, const float* direction_z, float* t,
                                 void* user_data, bool (*cb)(uint32_t, float*, void*)) {
    if (!origin_x || !origin_y || !origin_z || !direction_x || !direction_y || !direction_z ||
        !t || !cb) {
        set_error("null packet array or callback");
        return Error;
    }
    if (any_nan(origin_x, 4) || any_nan(origin_y, 4) || any_nan(origin_z, 4) ||
        any_nan(direction_x, 4) || any_nan(direction_y, 4) || any_nan(direction_z, 4))
        return Nan;  // lib.rs:799-807
    Packet p = make_packet(origin_x, origin_y, origin_z, direction_x, direction_y, direction_z, t);
    walk_mbvh(
        bvh.nodes, bvh.nodes ? bvh.node_count : 0, bvh.indices, bvh.indices ? bvh.index_count : 0,
        [&](const RTMbvhNode& n, int ids[4], bool res[4]) {
            for (int i = 0; i < 4; ++i) ids[i] = i;  // unsorted (mbvh_node.rs:293)
            mbvh_hit4(n, p, res);
        },
        [&](uint32_t prim) { return cb(prim, p.t, user_data); });
    for (int l = 0; l < 4; ++l) t[l] = p.t[l];
    return Ok;
}

}  // extern "C"
