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
    const DevRay* rays;
    uint64_t ray_count;
    DevHit* hits;
    uint8_t* occluded;
    uint32_t* status;
    uint64_t* stats;  // non-null: {node visits, triangle tests} accumulated (MBVH closest hit)
};

hipError_t launch_trace(const TraceParams& p, bool mbvh, bool ref_order, bool any_hit,
                        hipStream_t stream);
// Lowest-(t, id) MBVH traversal (trace_fast.hip).
hipError_t launch_trace_fast(const TraceParams& p, bool any_hit, hipStream_t stream);

// Device status word of the current device (kStatus* bits).
uint32_t* device_status_ptr();

// ---- GPU builders ----------------------------------------------------------
// Device arrays owned by one built tree.
struct DeviceBvh {
    DevAabb* nodes = nullptr;
    uint32_t* indices = nullptr;
    int32_t* parents = nullptr;  // lazily built by ensure_parents()
    uint32_t node_count = 0;
    uint32_t prim_count = 0;
};

struct DeviceMbvh {
    DevMbvhNode* nodes = nullptr;
    uint32_t* indices = nullptr;
    uint32_t node_count = 0;
    uint32_t prim_count = 0;
};

// Build a binary BVH on the device.  aabbs may be null (point boxes at the
// centres).  On success the DeviceBvh holds newly allocated device arrays.
ResultCode gpu_build_ploc(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                          size_t n, int radius, hipStream_t stream, DeviceBvh* out);
ResultCode gpu_build_binned(const RTAabb* aabbs, const float* centers, size_t stride_floats,
                            size_t n, size_t max_leaf, hipStream_t stream, DeviceBvh* out);
ResultCode gpu_build_mbvh(DeviceBvh& bvh, hipStream_t stream, DeviceMbvh* out);
ResultCode gpu_refit(DeviceBvh& bvh, const RTAabb* aabbs, hipStream_t stream);
ResultCode ensure_parents(DeviceBvh& bvh, hipStream_t stream);

void set_error(const char* fmt, ...);
ResultCode hip_fail(hipError_t e, const char* what);

}  // namespace rtbvh
