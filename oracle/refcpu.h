/*
 * refcpu — CPU restatement of meirbon/rtbvh's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the GPU
 * implementation in rtbvh_amd/ is checked against.  Only tests/, the smoke
 * check in __graft_entry__.py and bench.py's CPU-baseline leg may load it.
 * The product library never links or calls it.
 *
 * Every function follows the reference file:line cited beside it.  The
 * reference is Rust (glam/rayon); it cannot be compiled in this environment,
 * so this file is the restatement that the reference's own known-answer
 * tests, closed forms, invariants and brute force pin (tests/test_oracle_*).
 *
 * Build: oracle/Makefile (g++ -O3 -ffp-contract=off, SSE2 baseline, no FMA).
 */
#ifndef RTBVH_REFCPU_H
#define RTBVH_REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layouts identical to rtbvh_ffi's RTAabb / RTBvhNode / RTMbvhNode
 * (rtbvh_ffi/src/lib.rs:144-208). */
typedef struct RefAabb {
    float min[3];
    int32_t extra1; /* count (BvhNode) */
    float max[3];
    int32_t extra2; /* left_first (BvhNode) */
} RefAabb;

typedef struct RefMbvhNode {
    float min_x[4], max_x[4], min_y[4], max_y[4], min_z[4], max_z[4];
    int32_t children[4];
    int32_t counts[4];
} RefMbvhNode;

/* 32-byte ray record shared with the GPU ABI: origin, t_min, direction, t_max. */
typedef struct RefRay {
    float o[3];
    float t_min;
    float d[3];
    float t_max;
} RefRay;

/* 16-byte hit record: t, u, v, prim (UINT32_MAX on miss, t = t_max). */
typedef struct RefHit {
    float t, u, v;
    uint32_t prim;
} RefHit;

/* Tie rules for closest hit. */
enum {
    REF_MODE_LOWEST_ID = 0,  /* lexicographic min of (t, prim id) */
    REF_MODE_VISIT_ORDER = 1 /* reference: first accepted candidate in visit order wins */
};

/* ---- geometry (src/morton.rs, src/utils.rs, src/aabb.rs) ---- */
uint32_t ref_morton_split(uint32_t v);
void ref_world_bounds(const RefAabb* aabbs, size_t n, RefAabb* out);
void ref_morton_codes(const RefAabb* world, const float* centers, size_t center_stride_floats,
                      size_t n, uint32_t* codes);
void ref_sorted_indices(const uint32_t* codes, size_t n, uint32_t* indices);
uint32_t ref_prefix_sum_u32(const uint32_t* first, size_t count, uint32_t* output);
float ref_half_area(const RefAabb* a);

/* ---- triangle helpers (shared/src/lib.rs:27-38) ---- */
void ref_triangle_aabbs_centers(const float* verts, size_t n, RefAabb* aabbs, float* centers3);

/* ---- builders ---- */
/* LOCB (src/builders/locb.rs:249-329). nodes must hold 2n-1 (or 1) entries,
 * indices n.  Returns the node count. */
size_t ref_build_locb(const RefAabb* aabbs, const float* centers, size_t center_stride_floats,
                      size_t n, RefAabb* nodes, uint32_t* indices);
/* Same, with a caller-chosen search radius (locb.rs:31-34). */
size_t ref_build_locb_radius(const RefAabb* aabbs, const float* centers,
                             size_t center_stride_floats, size_t n, size_t radius,
                             RefAabb* nodes, uint32_t* indices, uint32_t* iterations);
/* Binned SAH (src/builders/binned_sah.rs:346-400), single-threaded DFS
 * allocation order.  Returns the node count. */
size_t ref_build_binned_sah(const RefAabb* aabbs, const float* centers,
                            size_t center_stride_floats, size_t n, size_t prims_per_leaf,
                            RefAabb* nodes, uint32_t* indices);
/* The same with the reference's two box defects fixed (stale left-box count,
 * binned_sah.rs:188-235; double leaf pad, :133,135), so every child box lies
 * inside its parent; the topology rule is unchanged. */
size_t ref_build_binned_sah_contained(const RefAabb* aabbs, const float* centers,
                                      size_t center_stride_floats, size_t n,
                                      size_t prims_per_leaf, RefAabb* nodes, uint32_t* indices);
/* MBVH collapse (src/bvh.rs:381-404, src/mbvh_node.rs:297-411).  m_nodes must
 * hold node_count entries.  Returns the MBVH node count. */
size_t ref_build_mbvh(const RefAabb* nodes, size_t node_count, RefMbvhNode* m_nodes);
/* Bvh::refit (src/bvh.rs:176-205), in place. */
void ref_refit(RefAabb* nodes, size_t node_count, const uint32_t* indices, size_t index_count,
               const RefAabb* aabbs);

/* ---- traversal + Moller-Trumbore (closest hit, any hit) ----
 * verts: 9 floats per primitive (v0, v1, v2).  stats (nullable) receives
 * {node visits, triangle tests} summed over the rays. */
void ref_trace_mbvh(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                    size_t index_count, const float* verts, size_t prim_count,
                    const RefRay* rays, size_t n, RefHit* hits, int mode, int threads,
                    uint64_t* stats);
void ref_trace_bvh(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                   size_t index_count, const float* verts, size_t prim_count,
                   const RefRay* rays, size_t n, RefHit* hits, int mode, int threads,
                   uint64_t* stats);
void ref_trace_brute(const float* verts, size_t prim_count, const RefRay* rays, size_t n,
                     RefHit* hits, int threads);
void ref_occluded_mbvh(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                       size_t index_count, const float* verts, size_t prim_count,
                       const RefRay* rays, size_t n, uint8_t* occluded, int threads);
void ref_occluded_bvh(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                      size_t index_count, const float* verts, size_t prim_count,
                      const RefRay* rays, size_t n, uint8_t* occluded, int threads);
void ref_occluded_brute(const float* verts, size_t prim_count, const RefRay* rays, size_t n,
                        uint8_t* occluded, int threads);
/* Single Moller-Trumbore test (spatial_sah.rs:132-163).  Returns 1 and writes
 * t,u,v when t_min < t < t_cur. */
int ref_mt(const float* v9, const float o[3], const float d[3], float t_min, float t_cur,
           float* t, float* u, float* v);

/* ---- structure checks (strengthened Bvh::validate, src/bvh.rs:207-244) ---- */
/* Returns 0 when every primitive is referenced exactly once from the root,
 * every interior child index is in range and the walk terminates; else a
 * nonzero code.  contain_violations (nullable) receives the number of
 * parent/child pairs whose child box is not inside the parent box. */
int ref_check_bvh(const RefAabb* nodes, size_t node_count, const uint32_t* indices,
                  size_t index_count, size_t prim_count, uint64_t* contain_violations,
                  uint32_t* max_depth);
int ref_check_mbvh(const RefMbvhNode* m, size_t m_count, const uint32_t* indices,
                   size_t index_count, size_t prim_count, uint32_t* max_depth);
/* SAH cost with C_trav = C_isect = 1, relative to the root's half area. */
double ref_sah_cost(const RefAabb* nodes, size_t node_count);
/* Structural equality from the roots (numbering-independent).  Returns 0 when
 * equal.  Leaf primitive lists are compared as sets when as_sets != 0. */
int ref_compare_bvh(const RefAabb* a, size_t na, const uint32_t* ia, const RefAabb* b,
                    size_t nb, const uint32_t* ib, int as_sets);
int ref_compare_mbvh(const RefMbvhNode* a, size_t na, const uint32_t* ia,
                     const RefMbvhNode* b, size_t nb, const uint32_t* ib);

#ifdef __cplusplus
}
#endif
#endif
