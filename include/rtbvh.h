/*
 * rtbvh.h — C ABI of librtbvh_amd.so, the MI355X (gfx950) implementation of
 * meirbon/rtbvh's hot path: BVH construction (locally ordered clustering /
 * PLOC, binned SAH, 4-wide MBVH collapse) and closest-hit / any-hit ray
 * traversal.
 *
 * Part 1 is the reference's own C ABI (the cbindgen output of
 * rtbvh_ffi/src/lib.rs, include guard RTBVH_H, rtbvh_ffi/build.rs:15-45):
 * same structs, enum values, symbols and argument meaning, so existing C/C++
 * callers link unchanged.  Trees are built on the GPU and mirrored to host
 * memory owned by the library.
 *
 * Part 2 is additive: device-resident, stream-ordered entry points that take
 * device pointers and a HIP stream (passed as void*).  They have no
 * counterpart in the reference; each cites the reference behaviour it
 * reproduces.
 *
 * Error convention: every entry point returns a ResultCode and never aborts.
 * Codes 0-4 keep the reference's values (rtbvh_ffi/src/lib.rs:17-25); codes
 * 5-9 are appended.  rtbvh_last_error() describes the last failure on the
 * calling thread.
 *
 * Threading (as rtbvh_ffi's StructureManager, rtbvh_ffi/src/lib.rs:8-15):
 * handles live in one registry guarded by a reader-writer lock.  Creation
 * and free_* take it exclusively; the traversals that look a handle up (the
 * prepared, traversal-tree and batch calls) hold it shared until their
 * kernels are enqueued, so a concurrent free waits for them and then for the
 * device before releasing memory.  Calls that take raw node/index pointers
 * (intersect*, rtbvh_gpu_intersect*, rtbvh_gpu_occluded*) read them without a
 * lock, as the reference's do: freeing or refitting a tree while another
 * thread traverses it is the caller's race.  Calls on one batch plan must not
 * overlap, its free included.
 *
 * Memory: the builders behind create_bvh, rtbvh_gpu_create_bvh,
 * rtbvh_gpu_create_spatial_bvh, rtbvh_gpu_create_batch and
 * rtbvh_gpu_create_trace_tree*, and traversals with RTBVH_TRACE_ORDER_RAYS,
 * keep their scratch between calls, one block per kind (PLOC, binned SAH,
 * spatial SAH, traversal tree, ray order) and device, grown to the largest
 * request up to 1 GiB each and held until the process ends; concurrent
 * calls on one device and larger requests allocate and free their own.
 */
#ifndef RTBVH_H
#define RTBVH_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================== */
/* Part 1 — rtbvh_ffi ABI                                                 */
/* ===================================================================== */

/* rtbvh_ffi/src/lib.rs:17-25 (values 0-4); 5-9 appended. */
typedef enum ResultCode {
    Ok = 0,
    Error = 1,
    NoPrimitives = 2,
    InequalAabbsAndPrimitives = 3,
    Nan = 4,
    DeviceError = 5,     /* a HIP call failed */
    OutOfMemory = 6,     /* device or host allocation failed */
    InvalidArgument = 7, /* bad stride, size, flag or pointer kind */
    StackOverflow = 8,   /* a ray needed more than RTBVH_MAX_STACK entries */
    NoDevice = 9,        /* no usable gfx950 device */
} ResultCode;

/* rtbvh_ffi/src/lib.rs:129-133 (#[repr(u32)]): a 32-bit value in both
 * languages. */
#ifdef __cplusplus
enum BvhType : uint32_t {
    LocallyOrderedClustered = 0,
    BinnedSAH = 1,
};
#else
enum BvhType {
    LocallyOrderedClustered = 0,
    BinnedSAH = 1,
};
typedef uint32_t BvhType;
#endif

/* rtbvh_ffi/src/lib.rs:144-151 — 32 bytes, identical to rtbvh::Aabb<i32>. */
typedef struct RTAabb {
    float min[3];
    int32_t count;
    float max[3];
    int32_t left_first;
} RTAabb;

/* rtbvh_ffi/src/lib.rs:175-179 — count >= 0: leaf with `count` primitives at
 * indices[left_first..]; count == -1: interior, children at left_first and
 * left_first + 1. */
typedef struct RTBvhNode {
    RTAabb aabb;
} RTBvhNode;

/* rtbvh_ffi/src/lib.rs:197-208 — 128 bytes, four child slots in SoA form.
 * counts[i] >= 0: leaf slot (children[i] = first index); counts[i] == -1 and
 * children[i] >= 0: interior slot (children[i] = MBVH node); both -1: empty. */
typedef struct RTMbvhNode {
    float min_x[4];
    float max_x[4];
    float min_y[4];
    float max_y[4];
    float min_z[4];
    float max_z[4];
    int32_t children[4];
    int32_t counts[4];
} RTMbvhNode;

/* rtbvh_ffi/src/lib.rs:210-220 — passed by value; the pointers refer to
 * library-owned memory (host memory for Part 1 handles, device memory for
 * Part 2 handles). */
typedef struct RTBvh {
    uint32_t id;
    uint32_t node_count;
    const RTBvhNode *nodes;
    uint32_t index_count;
    const uint32_t *indices;
} RTBvh;

/* rtbvh_ffi/src/lib.rs:234-244 */
typedef struct RTMbvh {
    uint32_t id;
    uint32_t node_count;
    const RTMbvhNode *nodes;
    uint32_t index_count;
    const uint32_t *indices;
} RTMbvh;

/* rtbvh_ffi/src/lib.rs:428-493.  Builds a BVH over prim_count primitives
 * given by their centres (center_stride 12 or 16 bytes) and, optionally, their
 * aabbs (NULL: point boxes at the centres).  Null centres/result -> Error;
 * prim_count 0 -> NoPrimitives; a stride other than 12 or 16 ->
 * InvalidArgument (the reference panics).  The tree is built on the GPU and
 * mirrored to host memory; *result points at that mirror. */
ResultCode create_bvh(const RTAabb *aabbs, size_t prim_count, const float *centers,
                      size_t center_stride, size_t prims_per_leaf, BvhType bvh_type,
                      RTBvh *result);

/* rtbvh_ffi/src/lib.rs:499-513.  Collapses a tree made by create_bvh into the
 * 4-wide MBVH (on the GPU).  Unknown id / null pointers -> Error. */
ResultCode create_mbvh(RTBvh bvh, RTMbvh *mbvh);

/* rtbvh_ffi/src/lib.rs:519-538.  Recomputes every node's bounds from new
 * per-primitive aabbs (prim_count of them), in place, keeping the topology.
 * The reference's Bvh::refit also resets count/left_first to 0 (src/bvh.rs:
 * 179,203); this implementation keeps them. */
ResultCode refit(const RTAabb *aabbs, RTBvh bvh);

/* rtbvh_ffi/src/lib.rs:551-581, 599-686, 700-731, 749-835.  Candidate
 * iteration with a host callback per candidate primitive, in the reference's
 * visit order; the callback may lower *t and returns true to stop early.
 * These run on the host because the callback is host code. */
ResultCode intersect(RTBvh bvh, const float *origin, const float *direction, float *t,
                     void *user_data, bool (*intersect)(uint32_t, float *, void *));
ResultCode intersect_packet(RTBvh bvh, const float *origin_x, const float *origin_y,
                            const float *origin_z, const float *direction_x,
                            const float *direction_y, const float *direction_z, float *t,
                            void *user_data, bool (*intersect)(uint32_t, float *, void *));
ResultCode intersect_mbvh(RTMbvh bvh, const float *origin, const float *direction, float *t,
                          void *user_data, bool (*intersect)(uint32_t, float *, void *));
ResultCode intersect_mbvh_packet(RTMbvh bvh, const float *origin_x, const float *origin_y,
                                 const float *origin_z, const float *direction_x,
                                 const float *direction_y, const float *direction_z, float *t,
                                 void *user_data,
                                 bool (*intersect)(uint32_t, float *, void *));

/* rtbvh_ffi/src/lib.rs:838-849.  Unknown ids print to stderr, as in the
 * reference. */
void free_bvh(RTBvh bvh);
void free_mbvh(RTMbvh bvh);

/* ===================================================================== */
/* Part 2 — device-resident GPU entry points (additive)                   */
/* ===================================================================== */

/* 32-byte ray: Ray::new (src/ray.rs:166-182) takes origin and direction,
 * t_min defaults to 1e-4 and t (the current maximum) to 1e34; the FFI sets t
 * from the caller (rtbvh_ffi/src/lib.rs:565-566). */
typedef struct RTRay {
    float origin[3];
    float t_min;
    float direction[3];
    float t_max;
} RTRay;

/* 16-byte closest-hit record.  prim_id == RTBVH_MISS and t == t_max on a
 * miss; u, v are the barycentrics of the Moller-Trumbore test
 * (src/builders/spatial_sah.rs:132-163). */
typedef struct RTHit {
    float t;
    float u;
    float v;
    uint32_t prim_id;
} RTHit;

#define RTBVH_MISS 0xFFFFFFFFu
/* Maximum traversal stack entries per ray (the reference has 32, unchecked:
 * src/iter.rs:25). */
#define RTBVH_MAX_STACK 256

/* Trace flags.  Default (0): the closest hit is the lexicographic minimum of
 * (t, prim_id) over all triangles — identical to brute force.  With
 * RTBVH_TRACE_REFERENCE_ORDER the kernel follows the reference iterator's
 * visit order exactly (src/iter.rs:308-354 for the MBVH, :88-127 for the
 * BVH) and, on exactly equal t, keeps the first accepted candidate as the
 * reference's closest-hit driver does (examples/benchmark.rs:33-37). */
#define RTBVH_TRACE_REFERENCE_ORDER 0x1u
/* Traversal-tree calls only (rtbvh_gpu_trace_tree_intersect/_occluded; no
 * counterpart in the reference, which traces in the caller's order): trace
 * the rays grouped by the cell of their origin in a grid over the origins'
 * bounds (up to 256^3 cells, 4 to 16 rays per cell on average), cells in
 * Morton order, and write each result at the
 * ray's own index — the same results as without the flag, bit for bit.  For
 * large sets of unrelated rays (2^28 shadow rays from random surface points
 * of a 20M-triangle scene: ordering and any-hit traversal 85 ms against
 * 157 ms); it costs a stable radix sort of the rays' cells.  The call takes device memory from
 * a per-device cache — about 16 bytes per ray for a compressed tree, whose
 * kernel reads the rays through the order; 64 for a 4-wide tree, traced
 * over a sorted copy whose results are gathered back.  The call is
 * stream-ordered like the others: a later call that reuses the cached block
 * waits for this one on the device; a request beyond the cache's 1 GiB is
 * allocated and freed per call, and then the call waits for its work.  At
 * most 2^32 - 4096 rays per call. */
#define RTBVH_TRACE_ORDER_RAYS 0x2u

/* GPU build, same arguments as create_bvh but every pointer is device memory
 * and the work is enqueued on `stream` (hipStream_t; NULL = default stream).
 * *result receives device pointers owned by the library; free with
 * rtbvh_gpu_free_bvh.  The tree is complete when the stream reaches this
 * point.  bvh_type LocallyOrderedClustered = PLOC with search radius 14
 * (src/builders/locb.rs), BinnedSAH = 16-bin SAH (src/builders/binned_sah.rs). */
ResultCode rtbvh_gpu_create_bvh(const RTAabb *aabbs, size_t prim_count, const float *centers,
                                size_t center_stride, size_t prims_per_leaf, BvhType bvh_type,
                                void *stream, RTBvh *result);

/* GPU spatial-split SAH build, the capability of create_spatial_Bvh
 * (rtbvh_ffi/src/lib.rs:337-389, SpatialSahBuilder,
 * src/builders/spatial_sah.rs:247-1034): nodes are split by reference
 * centroids or by planes that cut straddling triangles into one reference
 * per side, bounded by the clipped triangle; up to 0.75 * prim_count extra
 * references.  Arguments as rtbvh_gpu_create_bvh plus `vertices` (device, 9
 * floats per primitive, the triangles the references clip).  result->index_count
 * counts the references (at least prim_count; a primitive may sit in several
 * leaves).  Synchronises the stream. */
ResultCode rtbvh_gpu_create_spatial_bvh(const RTAabb *aabbs, size_t prim_count,
                                        const float *centers, size_t center_stride,
                                        const float *vertices, size_t prims_per_leaf,
                                        void *stream, RTBvh *result);

/* GPU MBVH collapse (Mbvh::construct, src/bvh.rs:381-404) of a tree made by
 * rtbvh_gpu_create_bvh. */
ResultCode rtbvh_gpu_create_mbvh(RTBvh bvh, void *stream, RTMbvh *result);

/* GPU refit of a tree made by rtbvh_gpu_create_bvh (aabbs: device, one per
 * primitive). */
ResultCode rtbvh_gpu_refit(const RTAabb *aabbs, RTBvh bvh, void *stream);

/* Closest-hit traversal of ray_count rays.  `bvh`/`mbvh` may be any tree whose
 * nodes/indices point at device memory (from rtbvh_gpu_create_*, or uploaded
 * by the caller).  vertices: device, 9 floats (v0, v1, v2) per primitive,
 * prim_count primitives.  rays/hits: device arrays of ray_count records.
 * Asynchronous on `stream`; call rtbvh_gpu_status() after synchronising to
 * learn whether any ray overflowed its stack. */
ResultCode rtbvh_gpu_intersect(RTBvh bvh, const float *vertices, size_t prim_count,
                               const RTRay *rays, size_t ray_count, RTHit *hits, uint32_t flags,
                               void *stream);
ResultCode rtbvh_gpu_intersect_mbvh(RTMbvh mbvh, const float *vertices, size_t prim_count,
                                    const RTRay *rays, size_t ray_count, RTHit *hits,
                                    uint32_t flags, void *stream);

/* rtbvh_gpu_intersect_mbvh that also adds {MBVH nodes visited, triangles
 * tested} over all rays into counters[0..1] (device, uint64; not cleared) —
 * the per-ray work behind the traversal's memory traffic. */
ResultCode rtbvh_gpu_intersect_mbvh_stats(RTMbvh mbvh, const float *vertices, size_t prim_count,
                                          const RTRay *rays, size_t ray_count, RTHit *hits,
                                          uint32_t flags, uint64_t *counters, void *stream);

/* Attaches to an MBVH made by rtbvh_gpu_create_mbvh a device copy of its
 * triangles in leaf order (v0 and the edges v1 - v0, v2 - v0 per index
 * entry, 64 bytes each: 48 used, 16 padding), taken from `vertices`
 * (device, 9 floats per primitive).  The copy is a snapshot: attach again after the vertices
 * change.  Freed with the MBVH. */
ResultCode rtbvh_gpu_mbvh_attach_triangles(RTMbvh mbvh, const float *vertices, size_t prim_count,
                                           void *stream);

/* rtbvh_gpu_intersect_mbvh / rtbvh_gpu_occluded_mbvh over the attached
 * triangles: identical results, one dependent memory access less per
 * triangle test.  Error if none are attached. */
ResultCode rtbvh_gpu_intersect_mbvh_prepared(RTMbvh mbvh, const RTRay *rays, size_t ray_count,
                                             RTHit *hits, uint32_t flags, void *stream);
ResultCode rtbvh_gpu_occluded_mbvh_prepared(RTMbvh mbvh, const RTRay *rays, size_t ray_count,
                                            uint8_t *occluded, uint32_t flags, void *stream);

/* ---- Ray packets ------------------------------------------------------------
 * RayPacket4 (src/ray.rs:44-59): four rays as ten 4-wide vectors.  On the GPU
 * a wavefront is the packet, so packets are unpacked into ray records and
 * traced like any other rays; hit i belongs to packet i / 4, lane i % 4. */
typedef struct RTRayPacket4 {
    float origin_x[4], origin_y[4], origin_z[4];
    float direction_x[4], direction_y[4], direction_z[4];
    float inv_direction_x[4], inv_direction_y[4], inv_direction_z[4];
    float t[4];
} RTRayPacket4;

/* rays[4 p + i] = RayPacket4::ray(i) of packets[p] (src/ray.rs:118-146): origin,
 * direction, t_min 1e-4, t_max = t[i].  packets, rays: device, 16-byte aligned. */
ResultCode rtbvh_gpu_rays_from_packets(const RTRayPacket4 *packets, size_t packet_count,
                                       RTRay *rays, void *stream);

/* ---- Traversal trees --------------------------------------------------------
 * A 4-wide tree built from a binary tree for the lowest-(t, id) traversal
 * only (no counterpart in the reference): each node opens the child with the
 * largest surface area until it has four slots, slot bounds are the
 * children's own, and subtrees of at most max_leaf primitives become one
 * leaf whose triangles are stored contiguously (64 bytes each).  Hits are
 * identical to rtbvh_gpu_intersect_mbvh's default mode on the same
 * triangles; the tree is a snapshot of `vertices`. */
typedef struct RTTraceTree {
    uint32_t id;
    uint32_t node_count;
    uint32_t prim_count;
    uint32_t max_leaf;
    uint32_t width;     /* 4, or 8: the compressed tree of rtbvh_gpu_create_trace_tree8 */
    uint32_t pad_;
    const void *nodes;       /* device: width 4, 128 B RTMbvhNode-layout records; width 8,
                                80 B compressed records (rtbvh_amd/csrc/device_math.hpp) */
    const float *triangles;  /* device: 16 floats per primitive, leaf order: v0 with the
                                primitive id's bits in [3], e1 = v1 - v0, e2 = v2 - v0,
                                4 floats of padding */
} RTTraceTree;

/* bvh: made by rtbvh_gpu_create_bvh; vertices: device, 9 floats per
 * primitive; max_leaf 1..8 (0: 1, the fastest on the benchmark scenes). */
ResultCode rtbvh_gpu_create_trace_tree(RTBvh bvh, const float *vertices, size_t prim_count,
                                       uint32_t max_leaf, void *stream, RTTraceTree *result);
/* The compressed 8-wide traversal tree: 80-byte nodes whose eight slot boxes
 * are 8-bit steps on a per-node power-of-two grid, rounded outward with one
 * step to spare; the binary tree spread over the slots by an SAH-optimal
 * distribution; slots placed by direction octant so that a ray visits the
 * near side first; one triangle per leaf slot.  Same hits, bit for bit, as
 * the 4-wide tree; traced and freed by the same calls.  Fails with Error on
 * non-finite bounds or binary leaves of more than 8 primitives (the 4-wide
 * tree accepts both). */
ResultCode rtbvh_gpu_create_trace_tree8(RTBvh bvh, const float *vertices, size_t prim_count,
                                        void *stream, RTTraceTree *result);
ResultCode rtbvh_gpu_trace_tree_intersect(RTTraceTree tree, const RTRay *rays, size_t ray_count,
                                          RTHit *hits, uint32_t flags, void *stream);
/* rtbvh_gpu_trace_tree_intersect that also adds {nodes visited, triangles
 * tested} into counters[0..1] (device, uint64; not cleared). */
ResultCode rtbvh_gpu_trace_tree_intersect_stats(RTTraceTree tree, const RTRay *rays,
                                                size_t ray_count, RTHit *hits, uint64_t *counters,
                                                void *stream);
ResultCode rtbvh_gpu_trace_tree_occluded(RTTraceTree tree, const RTRay *rays, size_t ray_count,
                                         uint8_t *occluded, uint32_t flags, void *stream);
void rtbvh_gpu_free_trace_tree(RTTraceTree tree);

/* ---- Launch shape of the lowest-id MBVH traversal ---------------------------
 * (no counterpart in the reference).  Per device; the default is
 * {64, 8, 800, 0, 16, 1, 0}.  Results never depend on
 * it, only speed. */
typedef struct RTTraceConfig {
    uint32_t block;          /* threads per block: 64, 128 or 256 */
    uint32_t lds_stack;      /* traversal stack entries per ray kept in LDS, 0..32 */
    uint32_t occupancy_pct;  /* persistent grid as a share of resident blocks, 10..3200 */
    uint32_t kernel;         /* 4-wide trees with leaf-ordered triangles: 0 one node or one
                                triangle per ray per iteration, 1 a node and all its leaf
                                slots (compressed trees have one kernel) */
    uint32_t refill;         /* kernel 0 and compressed trees: idle lanes of a wavefront
                                that take new rays together, 1..64 */
    uint32_t schedule;       /* 4-wide kernel 0 and compressed trees: 0 each lane walks a
                                fixed stride of rays, 1 each wavefront owns a contiguous
                                range of rays and its lanes draw the next one as they
                                free up; the launcher lowers refill and the grid when
                                the launch has few rays per resident lane, and runs
                                schedule 0 with refill 1 below 2 rays per lane */
    uint32_t postpone;       /* compressed trees: a wavefront tests triangles only once this
                                many of its lanes have one waiting, or a lane has nothing
                                else to do; 0..64 (0: whenever any lane has one) */
} RTTraceConfig;

ResultCode rtbvh_gpu_set_trace_config(const RTTraceConfig *config); /* NULL: default */
ResultCode rtbvh_gpu_get_trace_config(RTTraceConfig *config);
/* Autotuner: times 252 launch shapes (block 64/128/256 x LDS stack 4..16 x
 * grid 200/400/800 % x {kernel 0 with refill 1/4/16 and schedule 0/1,
 * kernel 1}) of the closest-hit traversal of `rays` against
 * `mbvh`'s attached triangles, writing `hits` (device, ray_count records),
 * and makes the fastest the current device's shape (also returned in *best).
 * Synchronises the stream. */
ResultCode rtbvh_gpu_autotune_mbvh(RTMbvh mbvh, const RTRay *rays, size_t ray_count, RTHit *hits,
                                   void *stream, RTTraceConfig *best);
/* The same over a traversal tree; for a compressed (width 8) tree, 55 shapes:
 * schedule 1 with block 64/128 x LDS stack 6/8/10 x grid 400/800/1600 % x
 * refill 8/16/32, and the static schedule at 64/8/800 %. */
ResultCode rtbvh_gpu_autotune_trace_tree(RTTraceTree tree, const RTRay *rays, size_t ray_count,
                                         RTHit *hits, void *stream, RTTraceConfig *best);

/* ---- Adjoint of the hit ----------------------------------------------------
 * Gradients of a loss L through closest hits made by the rtbvh_gpu_intersect*
 * calls (no counterpart in the reference, whose triangle test is forward
 * only).  The hit triangle is fixed by the forward pass; (t, u, v) solves
 * o + t d = (1-u-v) v0 + u v1 + v v2.  Inputs (device): vertices (9 floats per
 * primitive), the rays and hits of the forward pass, grad_tuv (3 floats per
 * ray: dL/dt, dL/du, dL/dv).  Outputs (device, each may be NULL):
 * grad_vertices (9 floats per primitive) is ACCUMULATED into (zero it first),
 * grad_origin / grad_direction (3 floats per ray) are written; rays that
 * missed get zeros. */
ResultCode rtbvh_gpu_adjoint(const float *vertices, size_t prim_count, const RTRay *rays,
                             const RTHit *hits, size_t ray_count, const float *grad_tuv,
                             float *grad_vertices, float *grad_origin, float *grad_direction,
                             void *stream);

/* ---- Batches of scenes ---------------------------------------------------
 * Many scenes of uneven size built and traced together (no counterpart in
 * the reference, which builds and traces one scene per call).  The scenes'
 * primitives are concatenated; scene s owns primitives
 * [scene_offsets[s], scene_offsets[s+1]).  Each scene's tree is the PLOC tree
 * a build of that scene alone makes; the trees share one node array. */
#define RTBVH_MAX_BATCH_SCENES 4096u

typedef struct RTBatch {
    uint32_t id;
    uint32_t scene_count;
    uint32_t bvh_node_count;      /* 2 * index_count - scene_count */
    uint32_t node_count;          /* traversal-tree nodes (0 until triangles are attached) */
    const RTBvhNode *bvh_nodes;   /* device; node s is the root of scene s */
    const void *nodes;            /* device traversal tree (compressed 8-wide records, as
                                     RTTraceTree width 8), node s the root of scene s; NULL
                                     in the handle returned by rtbvh_gpu_create_batch */
    uint32_t index_count;
    uint32_t pad_;
    const uint32_t *indices;      /* device; batch-wide primitive ids in leaf order */
} RTBatch;

/* Builds (PLOC, radius 14) every scene of the batch.
 * aabbs/centers: device, one per primitive of the batch (aabbs may be
 * NULL); scene_offsets: HOST array of scene_count + 1 entries, starting at 0,
 * every scene at least 3 primitives, scene_count <= RTBVH_MAX_BATCH_SCENES. */
ResultCode rtbvh_gpu_create_batch(const RTAabb *aabbs, const float *centers, size_t center_stride,
                                  const uint32_t *scene_offsets, size_t scene_count,
                                  void *stream, RTBatch *result);

/* Builds the batch's traversal tree (as rtbvh_gpu_create_trace_tree8, one
 * tree per scene, scene-local ids) from vertices: device, 9 floats per
 * primitive of the batch, and updates *batch.  Needed before tracing the
 * batch. */
ResultCode rtbvh_gpu_batch_attach_triangles(RTBatch *batch, const float *vertices,
                                            size_t prim_count, void *stream);

/* One persistent launch over the rays of every scene: ray i belongs to the
 * scene s with ray_offsets[s] <= i < ray_offsets[s+1] (ray_offsets: DEVICE
 * array of scene_count + 1 entries, ray_offsets[0] = 0, ray_offsets[scene_count]
 * = ray_count).  Hits carry scene-local primitive ids and equal, bit for bit,
 * those of tracing each scene on its own.  Lowest-(t, id) rule only (flags 0). */
ResultCode rtbvh_gpu_intersect_batch(RTBatch batch, const RTRay *rays, const uint32_t *ray_offsets,
                                     size_t ray_count, RTHit *hits, uint32_t flags, void *stream);
ResultCode rtbvh_gpu_occluded_batch(RTBatch batch, const RTRay *rays, const uint32_t *ray_offsets,
                                    size_t ray_count, uint8_t *occluded, uint32_t flags,
                                    void *stream);
/* rtbvh_gpu_autotune_trace_tree over a batch with attached triangles,
 * timing rtbvh_gpu_intersect_batch on `rays` with per-scene `ray_offsets`;
 * the shape list follows the batch's tree (compressed or 4-wide). */
ResultCode rtbvh_gpu_autotune_batch(RTBatch batch, const RTRay *rays, const uint32_t *ray_offsets,
                                    size_t ray_count, RTHit *hits, void *stream,
                                    RTTraceConfig *best);
void rtbvh_gpu_free_batch(RTBatch batch);

/* ---- Capturable batch build + trace ----------------------------------------
 * A plan holds every buffer for one batch shape (scene offsets fixed at
 * creation) and enqueues PLOC build, traversal tree (as
 * rtbvh_gpu_create_trace_tree with max_leaf 0) and traversal
 * without allocating and without waiting for the device, so the whole frame
 * can live in a hipGraph (no counterpart in the reference).  Clustering
 * iterations and tree levels are fixed at creation by an eager build of the
 * given inputs with headroom; when later inputs need more, the traversal of
 * that frame does nothing and rtbvh_gpu_batch_plan_status returns Error.
 * Calls on one plan must not overlap. */
typedef struct RTBatchPlan {
    uint32_t id;
    uint32_t scene_count;
    uint32_t prim_count;
    uint32_t iterations;  /* multi-block clustering iterations per build */
} RTBatchPlan;

/* aabbs (nullable), centers, vertices: device, as rtbvh_gpu_create_batch
 * and rtbvh_gpu_batch_attach_triangles (an eager build of these sizes the
 * budgets); scene_offsets: host, scene_count + 1 entries (scene_count may be
 * 1). */
ResultCode rtbvh_gpu_batch_plan_create(const RTAabb *aabbs, const float *centers,
                                       size_t center_stride, const float *vertices,
                                       const uint32_t *scene_offsets, size_t scene_count,
                                       void *stream, RTBatchPlan *result);
/* Enqueues build + collapse + triangles + traversal on `stream`.  Exactly one
 * of hits (closest hit, scene-local ids) and occluded (any hit) is non-NULL;
 * ray_offsets (device, scene_count + 1 entries) may be NULL for one scene.
 * aabbs may be NULL only if it was NULL at creation. */
ResultCode rtbvh_gpu_batch_plan_enqueue(RTBatchPlan plan, const RTAabb *aabbs,
                                        const float *centers, const float *vertices,
                                        const RTRay *rays, const uint32_t *ray_offsets,
                                        size_t ray_count, RTHit *hits, uint8_t *occluded,
                                        void *stream);
/* The same work as one hipGraph: recorded on the first call for a set of
 * argument pointers (after one eager pass), replayed by later calls with the
 * same pointers; ordered after prior work on `stream` and before later work. */
ResultCode rtbvh_gpu_batch_plan_launch(RTBatchPlan plan, const RTAabb *aabbs, const float *centers,
                                       const float *vertices, const RTRay *rays,
                                       const uint32_t *ray_offsets, size_t ray_count, RTHit *hits,
                                       uint8_t *occluded, void *stream);
/* Synchronises; Error if the last build ran out of clustering iterations,
 * else rtbvh_gpu_status. */
ResultCode rtbvh_gpu_batch_plan_status(RTBatchPlan plan, void *stream);
/* Traces later frames of a one-scene plan as RTBVH_TRACE_ORDER_RAYS does,
 * from buffers for up to max_ray_count rays (about 16 bytes per ray, at
 * most 2^32 - 4096 rays) that the plan
 * allocates here and keeps; launches of more rays then fail with
 * InvalidArgument.  max_ray_count 0 returns to the caller's order and frees
 * the buffers.  A recorded graph is recorded again on the next launch.  No
 * counterpart in the reference; for frames of many unrelated rays. */
ResultCode rtbvh_gpu_batch_plan_order_rays(RTBatchPlan plan, size_t max_ray_count, void *stream);
void rtbvh_gpu_batch_plan_free(RTBatchPlan plan);

/* Any-hit (shadow) traversal: occluded[i] = 1 when some triangle is hit with
 * t_min < t < t_max, else 0 (the FFI's early-out callback, rtbvh_ffi/src/
 * lib.rs:571-576). */
ResultCode rtbvh_gpu_occluded(RTBvh bvh, const float *vertices, size_t prim_count,
                              const RTRay *rays, size_t ray_count, uint8_t *occluded,
                              uint32_t flags, void *stream);
ResultCode rtbvh_gpu_occluded_mbvh(RTMbvh mbvh, const float *vertices, size_t prim_count,
                                   const RTRay *rays, size_t ray_count, uint8_t *occluded,
                                   uint32_t flags, void *stream);

/* ---- Allocation-free build ---------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, code that must not call hipMalloc): the PLOC build of
 * rtbvh_gpu_create_bvh (LocallyOrderedClustered; the same nodes and indices,
 * bit for bit) into caller arrays, its workspace taken from a static buffer of
 * RTBVH_NOALLOC_SCRATCH_BYTES per device that the library holds until the
 * process ends.  aabbs (nullable), centers (stride 12 or 16): device, as
 * rtbvh_gpu_create_bvh; nodes: device, at least max(1, 2 * prim_count - 1)
 * RTBvhNode records (for prim_count <= 2 the tree is the single root leaf in
 * nodes[0]); indices: device, prim_count entries.  prim_count 0 ->
 * NoPrimitives; above RTBVH_NOALLOC_MAX_PRIMS -> InvalidArgument.  Calls no
 * hipMalloc/hipFree.  Synchronises `stream` (host reads between clustering
 * batches).  Calls that overlap on one device race on the shared buffer. */
#define RTBVH_NOALLOC_MAX_PRIMS 262144u
#define RTBVH_NOALLOC_SCRATCH_BYTES (40u << 20)
ResultCode rtbvh_gpu_create_bvh_noalloc(const RTAabb *aabbs, size_t prim_count,
                                        const float *centers, size_t center_stride,
                                        RTBvhNode *nodes, uint32_t *indices, void *stream);

/* Synchronises `stream`, then returns and clears the device status of the
 * current device: Ok, or StackOverflow when a traversal ran out of stack
 * since the last call. */
ResultCode rtbvh_gpu_status(void *stream);

void rtbvh_gpu_free_bvh(RTBvh bvh);
void rtbvh_gpu_free_mbvh(RTMbvh mbvh);

/* Human-readable description of the last error on this thread ("" if none). */
const char *rtbvh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* RTBVH_H */
