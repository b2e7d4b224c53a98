// Internal interfaces shared by the translation units of librtbvh_amd.so.
#pragma once

#include <algorithm>
#include <shared_mutex>
#include <vector>

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/rtbvh.h"
#include "device_math.hpp"

namespace rtbvh {

constexpr uint32_t kStatusStackOverflow = 1u;

struct TraceParams {
    const DevMbvhNode* mnodes;
    const DevCw8Node* cnodes;  // non-null: compressed 8-wide tree (with tris)
    const DevAabb* bnodes;
    uint32_t node_count;
    uint32_t index_count;
    const uint32_t* indices;
    const float* verts;
    uint32_t prim_count;
    uint32_t pad_;
    const float4* tris;  // non-null: leaf-ordered triangles (3 float4 per index entry)
    const DevRay* rays;
    uint64_t ray_count;
    DevHit* hits;
    uint8_t* occluded;
    uint32_t* status;
    uint64_t* stats;  // non-null: {node visits, triangle tests} accumulated (MBVH closest hit)
    const uint32_t* ray_offsets;  // non-null: batch of scene_count scenes (MBVH forest)
    uint32_t scene_count;
    const uint32_t* gate;  // non-null: the launch does nothing while *gate != 0
    uint32_t lds_depth;    // set by the launcher: stack entries per lane held in LDS
    uint32_t refill;       // set by the launcher: idle lanes that trigger taking new rays
    uint32_t postpone;     // set by the launcher (compressed tree): triangle tests wait until
                           // this many lanes of the wave have one (0: never wait)
    bool small_leaves;     // every leaf holds at most 16 triangles (one-item kernel)
    uint32_t wave_ranges;  // set by the launcher: 1 each wave draws from its own ray range
    const uint32_t* ray_ids;  // non-null (compressed tree, one scene): position j traces ray
                              // ray_ids[j] and writes its result at ray_ids[j]
};


// Launch shape of the lowest-id MBVH traversal (per device).
struct TraceConfig {
    uint32_t block;          // threads per block: 64, 128 or 256
    uint32_t lds_depth;      // stack entries per lane in LDS (0..32)
    uint32_t occupancy_pct;  // persistent grid as a share of resident blocks (10..3200)
    uint32_t kernel;         // with leaf-ordered triangles: 0 one item per iteration,
                             // 1 a node and all its leaf slots per iteration
    uint32_t refill;         // one-item kernel: idle lanes of a wave that take new rays
                             // together (1..64)
    uint32_t schedule;       // one-item and compressed kernels: 0 fixed ray stride per lane,
                             // 1 each wave's lanes draw from the wave's own ray range
    uint32_t postpone;       // compressed tree: triangle tests wait until this many lanes
                             // of a wave have one, unless a lane has nothing else (0..64)
};
TraceConfig default_trace_config();
bool valid_trace_config(const TraceConfig& c);
void set_trace_config(const TraceConfig* c);  // null: back to the default
TraceConfig get_trace_config();

hipError_t launch_trace(const TraceParams& p, bool mbvh, bool ref_order, bool any_hit,
                        hipStream_t stream);
// Lowest-(t, id) MBVH traversal (trace_fast.hip).
hipError_t launch_trace_fast(const TraceParams& p, bool any_hit, hipStream_t stream);

// Device status word of the current device (kStatus* bits).
uint32_t* device_status_ptr();

// ---- GPU builders ----------------------------------------------------------
// Device arrays owned by one built tree.
// A batch of scenes is a forest: root_count roots at nodes [0, root_count),
// root s being scene s's tree.
struct DeviceBvh {
    DevAabb* nodes = nullptr;
    uint32_t* indices = nullptr;
    int32_t* parents = nullptr;  // lazily built by ensure_parents()
    uint32_t* order = nullptr;   // breadth-first node order, lazily built for the refit
    uint32_t node_count = 0;
    uint32_t prim_count = 0;  // entries of indices (above the primitives after spatial splits)
    uint32_t source_prims = 0;  // primitives built over, when that differs from prim_count
    uint32_t root_count = 1;
    std::vector<uint32_t> level_off;  // order[level_off[L], level_off[L+1]) is depth L
};

struct DeviceMbvh {
    DevMbvhNode* nodes = nullptr;
    DevCw8Node* cnodes = nullptr;  // compressed 8-wide traversal tree (nodes is then null)
    uint32_t* indices = nullptr;
    float4* tris = nullptr;  // attached leaf-ordered triangles (optional)
    uint32_t max_leaf = 0;   // largest leaf, measured when triangles are attached
    uint32_t node_count = 0;
    uint32_t prim_count = 0;
    uint32_t root_count = 1;  // MBVH node s is the root of scene s
};

constexpr uint32_t kMaxBatchScenes = RTBVH_MAX_BATCH_SCENES;

// Build a binary BVH on the device.  aabbs may be null (point boxes at the
// centres).  On success the DeviceBvh holds newly allocated device arrays.
ResultCode gpu_build_ploc(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                          size_t n, int radius, hipStream_t stream, DeviceBvh* out);
// Batched build: scene s owns primitives [scene_off[s], scene_off[s+1])
// (host array, scenes + 1 entries, each scene at least 3 primitives).
ResultCode gpu_build_ploc_batch(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                                const uint32_t* scene_off, uint32_t scenes, hipStream_t stream,
                                DeviceBvh* out);
ResultCode check_batch_offsets(const uint32_t* scene_off, uint32_t scenes);
// Workspace bytes of one single-scene PLOC build of n primitives.
size_t ploc_workspace_bytes(size_t n, bool point_aabbs);
// The same build as gpu_build_ploc (radius 14, 1 <= n) into caller arrays
// (nodes: max(1, 2n - 1) records, indices: n) with a caller workspace of
// ws_bytes; allocates nothing.  Synchronises the stream.
ResultCode gpu_build_ploc_into(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                               uint32_t n, void* ws, size_t ws_bytes, RTBvhNode* nodes,
                               uint32_t* indices, hipStream_t stream);

// Capturable PLOC build: every buffer allocated up front, a fixed number of
// clustering iterations (sized by one eager build at creation), no host
// synchronisation in ploc_plan_enqueue.  The tree's arrays stay owned by the
// plan.  The fail word is nonzero after an enqueued build whose iteration
// budget ran out (the tree is then incomplete).
struct PlocPlanImpl;
ResultCode ploc_plan_create(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                            const uint32_t* scene_off, uint32_t scenes, hipStream_t stream,
                            PlocPlanImpl** out, DeviceBvh* tree);
hipError_t ploc_plan_enqueue(const PlocPlanImpl* p, const RTAabb* aabbs, const float* centers,
                             hipStream_t stream);
const uint32_t* ploc_plan_fail_word(const PlocPlanImpl* p);
const uint32_t* ploc_plan_scene_offsets(const PlocPlanImpl* p);
uint32_t ploc_plan_iterations(const PlocPlanImpl* p);
void ploc_plan_destroy(PlocPlanImpl* p);

// Capturable collapse of a plan's tree into at most mcap MBVH nodes.
struct CollapsePlanImpl;
ResultCode collapse_plan_create(const DeviceBvh& tree, CollapsePlanImpl** out, DeviceMbvh* view);
hipError_t collapse_plan_enqueue(const CollapsePlanImpl* c, const DeviceBvh& tree,
                                 hipStream_t stream);
void collapse_plan_destroy(CollapsePlanImpl* c);

// parent[child] for every interior node (parents pre-filled with -1).
hipError_t launch_parent_links(const DevAabb* nodes, uint32_t count, int32_t* parents,
                               hipStream_t stream);

// Traversal tree (trace_tree.hip): width 4, a tree with tight slot bounds and
// leaves of up to max_leaf primitives; width 8, the compressed 8-wide tree
// with one triangle per leaf slot (max_leaf must be 1).  Triangles are stored
// in leaf order; traced by the lowest-id kernels with `tris`.  *out receives
// newly allocated nodes or cnodes (capacity prim_count + roots) and tris;
// indices is null.  cached: the plan's workspace comes from the workspace
// cache (a plan destroyed right after its build).
struct TraceTreePlan;
ResultCode trace_tree_plan_create(const DeviceBvh& tree, uint32_t max_leaf, uint32_t width,
                                  TraceTreePlan** out_plan, DeviceMbvh* out, bool cached = false);
ResultCode trace_tree_build(TraceTreePlan* t, const DeviceBvh& tree, const float* verts,
                            uint32_t prim_count, const uint32_t* scene_off, uint32_t scenes,
                            DeviceMbvh* out, hipStream_t stream);
hipError_t trace_tree_enqueue(const TraceTreePlan* t, const DeviceBvh& tree, const float* verts,
                              uint32_t prim_count, const uint32_t* scene_off, uint32_t scenes,
                              DeviceMbvh* out, uint32_t levels, uint32_t* fail,
                              hipStream_t stream);
uint32_t trace_tree_levels(const TraceTreePlan* t);
void trace_tree_plan_destroy(TraceTreePlan* t);

// Exclusive scans whose lengths are read on the device (scan.hip): up to
// kScanArraysMax arrays at once, each of its own length, totals into device
// memory.
constexpr int kScanArraysMax = 6;
struct ScanArrays {
    uint32_t* p[kScanArraysMax];
};
struct ScanCounts {
    const uint32_t* p[kScanArraysMax];
};
struct ScanTotals {
    uint32_t* p[kScanArraysMax];
};
size_t scan_dev_scratch_bytes(uint32_t max_count, int k);
hipError_t scan_exclusive_dev(const ScanArrays& a, int k, const ScanCounts& counts,
                              uint32_t max_count, const ScanTotals& totals, void* scratch,
                              size_t scratch_bytes, hipStream_t stream);

// Build workspaces.  Eager builds take their scratch from a per-device
// cache that keeps one block per builder kind between builds (grown to the
// largest request up to kWorkspaceCacheMax bytes, held until the process
// ends): a hipMalloc and a hipFree of a few hundred MB cost more than a 1M
// build's clustering iterations.  A block already in use, or a larger
// request, is allocated and freed per build as before.  The work that used a
// block must be complete when it is released (workspace_release waits for
// the given stream, or for the device when stream is null).
enum WorkspaceKind : int {
    kWsPloc = 0,
    kWsBinned = 1,
    kWsTree = 2,
    kWsSpatial = 3,
    kWsRayOrder = 4,
    kWsKinds = 5
};
constexpr size_t kWorkspaceCacheMax = size_t(1) << 30;
struct Workspace {
    void* ptr = nullptr;
    bool cached = false;
    int dev = -1;
    int kind = 0;
};
hipError_t workspace_acquire(WorkspaceKind kind, size_t bytes, Workspace* out);
void workspace_release(Workspace* w, hipStream_t stream, bool wait);
// Stream-ordered use (workspace.hip): no host wait for a cached block.
hipError_t workspace_acquire_on(WorkspaceKind kind, size_t bytes, hipStream_t stream,
                                Workspace* out);
void workspace_release_on(Workspace* w, hipStream_t stream);

// Stable radix sort of (key, value) pairs (build_ploc.hip): the low `bits`
// bits of the keys, 8 a pass; *second: the result is in (k1, v1).
constexpr uint32_t kRadixPairsMax = 0xFFFFF000u;
size_t radix_pairs_scratch_bytes(uint32_t n);
hipError_t radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t n,
                            uint32_t bits, void* scratch, size_t scratch_bytes,
                            hipStream_t stream, bool* second);

// Rays traced in the order of their origins' cells (ray_order.hip).  One
// block of ray_order_bytes(capacity, copies) holds every buffer; the passes
// allocate nothing and never wait for the host.  ray_order_sort leaves the
// caller's ray index of each sorted slot (*ids): compressed trees trace
// through it (TraceParams::ray_ids); with copies (4-wide trees) it also
// copies the rays into sorted[], the traversal writes results[] and
// ray_order_gather puts them in the caller's order.
struct RayOrder {
    float* frame;          // grid origin, cells per unit length, cells per axis
    float* parts;          // per-block bounds of the origin sample
    uint32_t *k0, *v0;     // (cell, ray) pairs
    uint32_t *k1, *v1;     // the sort's other buffers
    void* radix_scratch;
    size_t radix_bytes;
    DevRay* sorted;        // copies: capacity rays
    DevHit* results;       // copies: capacity hit records (or bytes, for any hit)
    uint64_t capacity;
};
size_t ray_order_bytes(uint64_t ray_count, bool copies);
RayOrder ray_order_carve(void* base, uint64_t ray_count, bool copies);
hipError_t ray_order_sort(const RayOrder& r, const DevRay* rays, uint64_t n, hipStream_t stream,
                          const uint32_t** ids);
hipError_t ray_order_gather(const RayOrder& r, const uint32_t* ids, uint64_t n, DevHit* hits,
                            uint8_t* occluded, hipStream_t stream);

ResultCode gpu_build_binned(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                            size_t n, size_t max_leaf, hipStream_t stream, DeviceBvh* out);
// Spatial-split SAH; verts: device, 9 floats per primitive.  The result's
// prim_count is its number of index entries (primitives and duplicates).
ResultCode gpu_build_spatial(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                             const float* verts, size_t n, size_t max_leaf, hipStream_t stream,
                             DeviceBvh* out);
ResultCode gpu_build_mbvh(DeviceBvh& bvh, hipStream_t stream, DeviceMbvh* out);
ResultCode gpu_refit(DeviceBvh& bvh, const RTAabb* aabbs, hipStream_t stream);
ResultCode ensure_parents(DeviceBvh& bvh, hipStream_t stream);
// Writes the record of every index entry pos (leaf order) at
// tris[kTriRecord * pos].
// scene_off (device, scenes + 1 entries) makes the stored ids scene-local.
hipError_t launch_prepare_triangles(const uint32_t* indices, uint32_t index_count,
                                    const float* verts, uint32_t prim_count, float4* tris,
                                    hipStream_t stream, const uint32_t* scene_off = nullptr,
                                    uint32_t scenes = 0);

// Adjoint of the hit (t, u, v) with respect to vertices and rays (adjoint.hip).
hipError_t launch_adjoint(const float* verts, uint32_t prim_count, const DevRay* rays,
                          const DevHit* hits, uint64_t ray_count, const float* grad_tuv,
                          float* grad_verts, float* grad_origin, float* grad_dir,
                          hipStream_t stream);

hipError_t mbvh_max_leaf(const DevMbvhNode* nodes, uint32_t count, uint32_t* result,
                         hipStream_t stream);

// A shared lock on the library's handle registry.  The lookups below take it
// into the caller's hold, and the caller keeps it until its kernels are
// enqueued: every free_* takes the registry's exclusive lock, so the arrays a
// lookup returned stay allocated until then (a free then waits for the
// device before releasing them).
using RegistryHold = std::shared_lock<std::shared_mutex>;

// small: every leaf of the MBVH holds at most 16 triangles.
const float4* attached_triangles(uint32_t id, uint32_t* count, int* device, bool* small,
                                 RegistryHold& hold);

struct BatchView {
    const DevMbvhNode* nodes;
    const DevCw8Node* cnodes;
    uint32_t node_count;
    const uint32_t* indices;
    uint32_t index_count;
    const float4* tris;
    uint32_t scenes;
    int device;
};
bool batch_view(uint32_t id, BatchView* out, RegistryHold& hold);
bool trace_tree_view(uint32_t id, BatchView* out, RegistryHold& hold);

constexpr uint32_t kDefaultMaxLeaf = 1;
// Traversal trees of batches and batch plans: 4 or 8 (the compressed tree).
constexpr uint32_t kBatchTreeWidth = 8;

void set_error(const char* fmt, ...);
ResultCode hip_fail(hipError_t e, const char* what);

}  // namespace rtbvh
