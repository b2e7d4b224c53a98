// Internal interfaces shared by the translation units of librtbvh_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/rtbvh.h"
#include "device_math.hpp"

namespace rtbvh {

constexpr uint32_t kStatusStackOverflow = 1u;

struct TraceParams {
    const DevMbvhNode* mnodes;
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
};

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
    uint32_t node_count = 0;
    uint32_t prim_count = 0;
    uint32_t root_count = 1;
};

struct DeviceMbvh {
    DevMbvhNode* nodes = nullptr;
    uint32_t* indices = nullptr;
    float4* tris = nullptr;  // attached leaf-ordered triangles (optional)
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
ResultCode gpu_build_binned(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                            size_t n, size_t max_leaf, hipStream_t stream, DeviceBvh* out);
ResultCode gpu_build_mbvh(DeviceBvh& bvh, hipStream_t stream, DeviceMbvh* out);
ResultCode gpu_refit(DeviceBvh& bvh, const RTAabb* aabbs, hipStream_t stream);
ResultCode ensure_parents(DeviceBvh& bvh, hipStream_t stream);
// Writes tris[3*pos..] for every index entry pos (leaf order).
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

const float4* attached_triangles(uint32_t id, uint32_t* count, int* device);

struct BatchView {
    const DevMbvhNode* nodes;
    uint32_t node_count;
    const uint32_t* indices;
    uint32_t index_count;
    const float4* tris;
    uint32_t scenes;
    int device;
};
bool batch_view(uint32_t id, BatchView* out);

void set_error(const char* fmt, ...);
ResultCode hip_fail(hipError_t e, const char* what);

}  // namespace rtbvh
