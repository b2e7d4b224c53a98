// Device-side geometry for gfx950: the reference's float arithmetic restated
// operation by operation so that results are bit-identical to rtbvh on x86-64.
//
// Requirements met by the build (rtbvh_amd/Makefile): -ffp-contract=off (no
// FMA contraction; Rust never contracts), no fast-math, IEEE division
// (hipcc's default correctly rounded fp32 divide), fp32 denormals preserved
// (.amdhsa_float_denorm_mode_32 3).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtbvh {

// glam Vec4::min/max on SSE2 (_mm_min_ps/_mm_max_ps): `a<b?a:b`, `a>b?a:b`;
// the second operand wins on NaN.  Used by the MBVH slab test
// (src/mbvh_node.rs:195-203).
__device__ __forceinline__ float ps_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float ps_max(float a, float b) { return a > b ? a : b; }

// f32::min/max (glam Vec3 component-wise; src/aabb.rs:74-86, 252-273) as
// LLVM lowers llvm.minnum/maxnum on x86-64: NaN-ignoring, ties return `a`.
__device__ __forceinline__ float rs_min(float a, float b) { return (a != a) ? b : (b < a ? b : a); }
__device__ __forceinline__ float rs_max(float a, float b) { return (a != a) ? b : (b > a ? b : a); }

struct alignas(16) DevAabb {
    float min[3];
    int32_t extra1;  // count
    float max[3];
    int32_t extra2;  // left_first
};

struct alignas(16) DevMbvhNode {
    float4 min_x, max_x, min_y, max_y, min_z, max_z;
    int4 children, counts;
};

// Compressed 8-wide traversal-tree node (80 B).  Slot s of the node covers
// [p + qlo[s] * 2^(e-127), p + qhi[s] * 2^(e-127)] per axis, a box that holds
// the slot's subtree with at least one grid step to spare on every side.
// Interior children are stored from child_base on, triangles (one per leaf
// slot, 64 B records) from tri_base on, both in slot order.
struct alignas(16) DevCw8Node {
    float px, py, pz;
    uint32_t emask;       // bytes 0..2: biased exponents of x, y, z; byte 3: interior slots
    uint32_t child_base;  // first interior child
    uint32_t tri_base;    // first triangle record
    uint32_t lmask;       // byte 0: triangle slots
    uint32_t pad;
    uint32_t qlo_x[2], qlo_y[2], qlo_z[2];  // byte s (of the pair) = slot s
    uint32_t qhi_x[2], qhi_y[2], qhi_z[2];
};
static_assert(sizeof(DevCw8Node) == 80, "compressed node is 80 bytes");

// Lowest-(t, id) rule: a candidate (t, prim) replaces the best (bt, bp) when
// it is nearer, or as near with a lower id -- the latter only against an
// accepted hit, never against the ray's t_max (bp still RTBVH_MISS), which
// the reference's test (t < ray.t, src/builders/spatial_sah.rs:157)
// excludes.  This is what brute force in id order gives.
__device__ __forceinline__ bool lowest_first(float t, uint32_t prim, float bt, uint32_t bp) {
    return (t < bt) || (t == bt && bp != 0xFFFFFFFFu && prim < bp);
}

struct DevRay {
    float o[3];
    float t_min;
    float d[3];
    float t_max;
};

struct DevHit {
    float t, u, v;
    uint32_t prim;
};

// Moller-Trumbore, SpatialTriangle::intersect (src/builders/spatial_sah.rs:
// 132-163).  glam's scalar Vec3: cross(a,b) = (a.y*b.z - b.y*a.z,
// a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y); dot = (a.x*b.x + a.y*b.y) + a.z*b.z.
// Returns the candidate t (and u, v) when the reference would accept it for
// some t_cur > t; `ok` is false when it rejects regardless of t_cur.
struct MtResult {
    bool ok;
    float t, u, v;
};

// The same test with the edges e1 = v1 - v0, e2 = v2 - v0 already formed
// (identical values: the subtraction is the same IEEE operation wherever it
// is done).
__device__ __forceinline__ MtResult mt_candidate_e(float v0x, float v0y, float v0z, float e1x,
                                                   float e1y, float e1z, float e2x, float e2y,
                                                   float e2z, float ox, float oy, float oz,
                                                   float dx, float dy, float dz, float t_min) {
    MtResult r;
    r.ok = false;
    r.t = 0.f;
    r.u = 0.f;
    r.v = 0.f;
    const float hx = dy * e2z - e2y * dz;
    const float hy = dz * e2x - e2z * dx;
    const float hz = dx * e2y - e2x * dy;
    const float a = (e1x * hx + e1y * hy) + e1z * hz;
    if (a > -1e-5f && a < 1e-5f) return r;
    const float f = 1.0f / a;
    const float sx = ox - v0x, sy = oy - v0y, sz = oz - v0z;
    const float u = f * ((sx * hx + sy * hy) + sz * hz);
    if (!(0.0f <= u && u <= 1.0f)) return r;
    const float qx = sy * e1z - e1y * sz;
    const float qy = sz * e1x - e1z * sx;
    const float qz = sx * e1y - e1x * sy;
    const float v = f * ((dx * qx + dy * qy) + dz * qz);
    if (v < 0.0f || u + v > 1.0f) return r;
    const float t = f * ((e2x * qx + e2y * qy) + e2z * qz);
    r.ok = t > t_min;
    r.t = t;
    r.u = u;
    r.v = v;
    return r;
}

__device__ __forceinline__ MtResult mt_candidate(float v0x, float v0y, float v0z, float v1x,
                                                 float v1y, float v1z, float v2x, float v2y,
                                                 float v2z, float ox, float oy, float oz,
                                                 float dx, float dy, float dz, float t_min) {
    return mt_candidate_e(v0x, v0y, v0z, v1x - v0x, v1y - v0y, v1z - v0z, v2x - v0x, v2y - v0y,
                          v2z - v0z, ox, oy, oz, dx, dy, dz, t_min);
}

// Leaf-ordered triangle record (kTriRecord float4 per index entry): v0 with the
// primitive id in .w (RTBVH_MISS: no triangle), e1, e2, and 16 bytes of
// padding, so that a record never straddles a 64-byte line (measured: 2 %
// faster incoherent traces than packed 48-byte records).
constexpr uint32_t kTriRecord = 4;

__device__ __forceinline__ MtResult mt_prepared(const float4* tris, uint32_t pos, float ox,
                                                float oy, float oz, float dx, float dy, float dz,
                                                float t_min, uint32_t* prim) {
    const float4 a = tris[size_t(pos) * kTriRecord];
    const float4 e1 = tris[size_t(pos) * kTriRecord + 1];
    const float4 e2 = tris[size_t(pos) * kTriRecord + 2];
    // all 48 bytes in one round trip: without this use the compiler moves the
    // v0 load behind the determinant test, a second dependent memory access
    asm volatile("" : : "v"(a.x), "v"(a.y), "v"(a.z));
    *prim = __float_as_uint(a.w);
    return mt_candidate_e(a.x, a.y, a.z, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z, ox, oy, oz, dx, dy,
                          dz, t_min);
}

}  // namespace rtbvh
