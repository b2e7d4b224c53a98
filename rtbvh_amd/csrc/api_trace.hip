// C ABI: device-resident traversal entry points and library-wide helpers.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "internal.hpp"

namespace rtbvh {

namespace {
thread_local char g_last_error[512] = {0};

// One status word per device (kStatus* bits), cleared by rtbvh_gpu_status().
__device__ uint32_t g_device_status;

bool device_accessible(const void* p) {
    if (p == nullptr) return false;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    switch (attr.type) {
        case hipMemoryTypeDevice:
        case hipMemoryTypeManaged:
        case hipMemoryTypeUnified:
            return true;
        case hipMemoryTypeHost:
            return attr.devicePointer != nullptr;
        default:
            return false;
    }
}

ResultCode check_device(const void* p, const char* name) {
    if (!device_accessible(p)) {
        set_error("%s (%p) is not device-accessible memory", name, p);
        return InvalidArgument;
    }
    return Ok;
}

}  // namespace

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

ResultCode hip_fail(hipError_t e, const char* what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return OutOfMemory;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return NoDevice;
    return DeviceError;
}

uint32_t* device_status_ptr() {
    static uint32_t* ptrs[64] = {nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (ptrs[dev] == nullptr) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_device_status)) != hipSuccess) return nullptr;
        ptrs[dev] = static_cast<uint32_t*>(p);
    }
    return ptrs[dev];
}

namespace {

ResultCode trace_entry(const void* nodes, uint32_t node_count, const uint32_t* indices,
                       uint32_t index_count, const float* vertices, size_t prim_count,
                       const RTRay* rays, size_t ray_count, RTHit* hits, uint8_t* occluded,
                       uint32_t flags, void* stream, bool mbvh, uint64_t* stats = nullptr,
                       const float4* tris = nullptr, bool small_leaves = false) {
    if (flags & ~RTBVH_TRACE_REFERENCE_ORDER) {
        set_error("unknown trace flags 0x%x", flags);
        return InvalidArgument;
    }
    if (prim_count > 0xFFFFFFFFull) {
        set_error("prim_count %zu exceeds 2^32-1", prim_count);
        return InvalidArgument;
    }
    if (ray_count == 0) return Ok;
    ResultCode rc;
    if ((rc = check_device(rays, "rays")) != Ok) return rc;
    if (hits && (rc = check_device(hits, "hits")) != Ok) return rc;
    if (occluded && (rc = check_device(occluded, "occluded")) != Ok) return rc;
    if (stats && (rc = check_device(stats, "stats")) != Ok) return rc;
    if (node_count > 0) {
        if ((rc = check_device(nodes, "nodes")) != Ok) return rc;
        if (index_count > 0 && (rc = check_device(indices, "indices")) != Ok) return rc;
        if (!tris && prim_count > 0 && (rc = check_device(vertices, "vertices")) != Ok) return rc;
    }
    if ((reinterpret_cast<uintptr_t>(rays) & 15) || (hits && (reinterpret_cast<uintptr_t>(hits) & 15)) ||
        (reinterpret_cast<uintptr_t>(nodes) & 15)) {
        set_error("rays, hits and nodes must be 16-byte aligned");
        return InvalidArgument;
    }
    TraceParams p{};
    p.mnodes = mbvh ? static_cast<const DevMbvhNode*>(nodes) : nullptr;
    p.bnodes = mbvh ? nullptr : static_cast<const DevAabb*>(nodes);
    p.node_count = node_count;
    p.index_count = indices ? index_count : 0;
    p.indices = indices;
    p.verts = vertices;
    p.prim_count = vertices ? uint32_t(prim_count) : 0;
    p.rays = reinterpret_cast<const DevRay*>(rays);
    p.ray_count = ray_count;
    p.hits = reinterpret_cast<DevHit*>(hits);
    p.occluded = occluded;
    p.stats = stats;
    p.tris = tris;
    p.small_leaves = small_leaves;
    p.status = device_status_ptr();
    if (p.status == nullptr) {
        set_error("device status word unavailable");
        return DeviceError;
    }
    const hipError_t e = launch_trace(p, mbvh, (flags & RTBVH_TRACE_REFERENCE_ORDER) != 0,
                                      occluded != nullptr, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "trace launch");
    return Ok;
}

ResultCode prepared_entry(RTMbvh mbvh, const RTRay* rays, size_t ray_count, RTHit* hits,
                          uint8_t* occluded, uint32_t flags, void* stream) {
    uint32_t count = 0;
    int device = 0;
    bool small = false;
    RegistryHold hold;  // held until the traversal is enqueued
    const float4* tris = attached_triangles(mbvh.id, &count, &device, &small, hold);
    if (tris == nullptr) {
        set_error("mbvh %u has no attached triangles (rtbvh_gpu_mbvh_attach_triangles)", mbvh.id);
        return Error;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
        set_error("mbvh %u lives on device %d, current device is %d", mbvh.id, device, cur);
        return InvalidArgument;
    }
    if (count != mbvh.index_count) {
        set_error("handle does not match the attached triangles");
        return InvalidArgument;
    }
    return trace_entry(mbvh.nodes, mbvh.node_count, mbvh.indices, mbvh.index_count, nullptr,
                       count, rays, ray_count, hits, occluded, flags, stream, true, nullptr, tris,
                       small);
}

// Lowest-id traversal of a traversal tree or a batch's tree (leaf-ordered
// triangles, no index array).
ResultCode view_entry(const BatchView& v, const char* what, uint32_t id, const RTRay* rays,
                      const uint32_t* ray_offsets, size_t ray_count, RTHit* hits,
                      uint8_t* occluded, uint32_t flags, void* stream,
                      uint64_t* stats = nullptr) {
    const bool order_rays = (flags & RTBVH_TRACE_ORDER_RAYS) != 0;
    if ((flags & ~RTBVH_TRACE_ORDER_RAYS) != 0) {
        set_error("this traversal supports the lowest-(t, id) rule only (flags 0 or "
                  "RTBVH_TRACE_ORDER_RAYS)");
        return InvalidArgument;
    }
    if (order_rays && (ray_offsets || stats)) {
        set_error("RTBVH_TRACE_ORDER_RAYS applies to one traversal tree without counters");
        return InvalidArgument;
    }
    if (order_rays && ray_count > kRadixPairsMax) {
        set_error("RTBVH_TRACE_ORDER_RAYS orders at most %u rays per call (got %zu)",
                  kRadixPairsMax, ray_count);
        return InvalidArgument;
    }
    if (ray_count > 0xFFFFFFFFull) {
        set_error("at most 2^32-1 rays per call (got %zu)", ray_count);
        return InvalidArgument;
    }
    if (v.tris == nullptr) {
        set_error("%s %u has no attached triangles", what, id);
        return Error;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != v.device) {
        set_error("%s %u lives on device %d, current device is %d", what, id, v.device, cur);
        return InvalidArgument;
    }
    if (ray_count == 0) return Ok;
    ResultCode rc;
    if (ray_offsets && (rc = check_device(ray_offsets, "ray_offsets")) != Ok) return rc;
    if ((rc = check_device(rays, "rays")) != Ok) return rc;
    if (hits && (rc = check_device(hits, "hits")) != Ok) return rc;
    if (occluded && (rc = check_device(occluded, "occluded")) != Ok) return rc;
    if (stats && (rc = check_device(stats, "counters")) != Ok) return rc;
    if ((reinterpret_cast<uintptr_t>(rays) & 15) || (hits && (reinterpret_cast<uintptr_t>(hits) & 15))) {
        set_error("rays and hits must be 16-byte aligned");
        return InvalidArgument;
    }
    if (stats && (occluded || ray_offsets)) {
        set_error("counters are available for single-tree closest hit only");
        return InvalidArgument;
    }
    TraceParams p{};
    p.stats = stats;
    p.small_leaves = true;  // traversal-tree leaves hold at most 8 triangles
    p.mnodes = v.nodes;
    p.cnodes = v.cnodes;
    p.node_count = v.node_count;
    p.index_count = v.index_count;
    p.indices = nullptr;
    p.tris = v.tris;
    p.rays = reinterpret_cast<const DevRay*>(rays);
    p.ray_count = ray_count;
    p.hits = reinterpret_cast<DevHit*>(hits);
    p.occluded = occluded;
    p.ray_offsets = ray_offsets;
    p.scene_count = v.scenes;
    p.status = device_status_ptr();
    if (p.status == nullptr) {
        set_error("device status word unavailable");
        return DeviceError;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (!order_rays) {
        const hipError_t e = launch_trace_fast(p, occluded != nullptr, st);
        if (e != hipSuccess) return hip_fail(e, "trace launch");
        return Ok;
    }
    // compressed trees read each ray through the order; 4-wide trees trace
    // a sorted copy whose results are gathered back
    const bool copies = v.cnodes == nullptr;
    Workspace ws;
    hipError_t e = workspace_acquire_on(kWsRayOrder, ray_order_bytes(ray_count, copies), st, &ws);
    if (e != hipSuccess) return hip_fail(e, "ray order workspace");
    const RayOrder ro = ray_order_carve(ws.ptr, ray_count, copies);
    const uint32_t* ids = nullptr;
    e = ray_order_sort(ro, reinterpret_cast<const DevRay*>(rays), ray_count, st, &ids);
    if (copies) {
        p.rays = ro.sorted;
        p.hits = hits ? ro.results : nullptr;
        p.occluded = occluded ? reinterpret_cast<uint8_t*>(ro.results) : nullptr;
    } else {
        p.ray_ids = ids;
    }
    if (e == hipSuccess) e = launch_trace_fast(p, occluded != nullptr, st);
    if (e == hipSuccess && copies)
        e = ray_order_gather(ro, ids, ray_count, reinterpret_cast<DevHit*>(hits), occluded, st);
    workspace_release_on(&ws, st);  // the next user waits for this call's work
    if (e != hipSuccess) return hip_fail(e, "ordered trace");
    return Ok;
}

ResultCode batch_entry(RTBatch batch, const RTRay* rays, const uint32_t* ray_offsets,
                       size_t ray_count, RTHit* hits, uint8_t* occluded, uint32_t flags,
                       void* stream) {
    BatchView v;
    RegistryHold hold;  // held until the traversal is enqueued
    if (!batch_view(batch.id, &v, hold)) {
        set_error("unknown batch id %u", batch.id);
        return Error;
    }
    if (ray_offsets == nullptr && ray_count > 0) {
        set_error("ray_offsets must not be null");
        return Error;
    }
    return view_entry(v, "batch", batch.id, rays, ray_offsets, ray_count, hits, occluded, flags,
                      stream);
}

ResultCode tree_entry(RTTraceTree tree, const RTRay* rays, size_t ray_count, RTHit* hits,
                      uint8_t* occluded, uint32_t flags, void* stream, uint64_t* stats = nullptr) {
    BatchView v;
    RegistryHold hold;  // held until the traversal is enqueued
    if (!trace_tree_view(tree.id, &v, hold)) {
        set_error("unknown trace tree id %u", tree.id);
        return Error;
    }
    v.scenes = 1;
    return view_entry(v, "trace tree", tree.id, rays, nullptr, ray_count, hits, occluded, flags,
                      stream, stats);
}

// RayPacket4 (src/ray.rs:44-59): ten Vec4 — origin x/y/z, direction x/y/z,
// inverse direction x/y/z, t — for four rays; RayPacket4::ray (src/ray.rs:
// 118-146) gives lane i the ray (origin_i, direction_i) with t_min 1e-4 and
// t_i.
__global__ void packets_to_rays_kernel(const float4* packets, uint64_t packet_count,
                                       float4* rays) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // one ray per thread
    if (i >= packet_count * 4) return;
    const uint64_t pk = i >> 2;
    const uint32_t lane = uint32_t(i & 3u);
    const float* q = reinterpret_cast<const float*>(packets + pk * 10);
    rays[i * 2] = make_float4(q[0 + lane], q[4 + lane], q[8 + lane], 1e-4f);
    rays[i * 2 + 1] = make_float4(q[12 + lane], q[16 + lane], q[20 + lane], q[36 + lane]);
}

}  // namespace
}  // namespace rtbvh

using namespace rtbvh;

extern "C" {

const char* rtbvh_last_error(void) { return g_last_error; }

ResultCode rtbvh_gpu_intersect(RTBvh bvh, const float* vertices, size_t prim_count,
                               const RTRay* rays, size_t ray_count, RTHit* hits, uint32_t flags,
                               void* stream) {
    if (hits == nullptr && ray_count > 0) {
        set_error("hits is null");
        return Error;
    }
    return trace_entry(bvh.nodes, bvh.node_count, bvh.indices, bvh.index_count, vertices,
                       prim_count, rays, ray_count, hits, nullptr, flags, stream, false);
}

ResultCode rtbvh_gpu_intersect_mbvh(RTMbvh mbvh, const float* vertices, size_t prim_count,
                                    const RTRay* rays, size_t ray_count, RTHit* hits,
                                    uint32_t flags, void* stream) {
    if (hits == nullptr && ray_count > 0) {
        set_error("hits is null");
        return Error;
    }
    return trace_entry(mbvh.nodes, mbvh.node_count, mbvh.indices, mbvh.index_count, vertices,
                       prim_count, rays, ray_count, hits, nullptr, flags, stream, true);
}

ResultCode rtbvh_gpu_intersect_mbvh_stats(RTMbvh mbvh, const float* vertices, size_t prim_count,
                                          const RTRay* rays, size_t ray_count, RTHit* hits,
                                          uint32_t flags, uint64_t* counters, void* stream) {
    if ((hits == nullptr || counters == nullptr) && ray_count > 0) {
        set_error("hits and counters must not be null");
        return Error;
    }
    return trace_entry(mbvh.nodes, mbvh.node_count, mbvh.indices, mbvh.index_count, vertices,
                       prim_count, rays, ray_count, hits, nullptr, flags, stream, true, counters);
}

ResultCode rtbvh_gpu_intersect_mbvh_prepared(RTMbvh mbvh, const RTRay* rays, size_t ray_count,
                                             RTHit* hits, uint32_t flags, void* stream) {
    if (hits == nullptr && ray_count > 0) {
        set_error("hits is null");
        return Error;
    }
    return prepared_entry(mbvh, rays, ray_count, hits, nullptr, flags, stream);
}

ResultCode rtbvh_gpu_adjoint(const float* vertices, size_t prim_count, const RTRay* rays,
                             const RTHit* hits, size_t ray_count, const float* grad_tuv,
                             float* grad_vertices, float* grad_origin, float* grad_direction,
                             void* stream) {
    if (prim_count > 0xFFFFFFFFull) {
        set_error("prim_count %zu exceeds 2^32-1", prim_count);
        return InvalidArgument;
    }
    if (ray_count == 0) return Ok;
    if (rays == nullptr || hits == nullptr || grad_tuv == nullptr ||
        (vertices == nullptr && prim_count > 0)) {
        set_error("vertices, rays, hits and grad_tuv must not be null");
        return Error;
    }
    ResultCode rc;
    if ((rc = check_device(rays, "rays")) != Ok || (rc = check_device(hits, "hits")) != Ok ||
        (rc = check_device(grad_tuv, "grad_tuv")) != Ok)
        return rc;
    if (prim_count > 0 && (rc = check_device(vertices, "vertices")) != Ok) return rc;
    if (grad_vertices && (rc = check_device(grad_vertices, "grad_vertices")) != Ok) return rc;
    if (grad_origin && (rc = check_device(grad_origin, "grad_origin")) != Ok) return rc;
    if (grad_direction && (rc = check_device(grad_direction, "grad_direction")) != Ok) return rc;
    if ((reinterpret_cast<uintptr_t>(rays) & 15) || (reinterpret_cast<uintptr_t>(hits) & 15)) {
        set_error("rays and hits must be 16-byte aligned");
        return InvalidArgument;
    }
    if (ray_count > (uint64_t(0x7FFFFFFF) * 256)) {
        set_error("too many rays (%zu)", ray_count);
        return InvalidArgument;
    }
    const hipError_t e = launch_adjoint(
        vertices, uint32_t(prim_count), reinterpret_cast<const DevRay*>(rays),
        reinterpret_cast<const DevHit*>(hits), ray_count, grad_tuv, grad_vertices, grad_origin,
        grad_direction, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "adjoint launch");
    return Ok;
}

ResultCode rtbvh_gpu_trace_tree_intersect(RTTraceTree tree, const RTRay* rays, size_t ray_count,
                                          RTHit* hits, uint32_t flags, void* stream) {
    if (hits == nullptr && ray_count > 0) {
        set_error("hits is null");
        return Error;
    }
    return tree_entry(tree, rays, ray_count, hits, nullptr, flags, stream);
}

ResultCode rtbvh_gpu_trace_tree_intersect_stats(RTTraceTree tree, const RTRay* rays,
                                                size_t ray_count, RTHit* hits, uint64_t* counters,
                                                void* stream) {
    if ((hits == nullptr && ray_count > 0) || counters == nullptr) {
        set_error("hits and counters must not be null");
        return Error;
    }
    return tree_entry(tree, rays, ray_count, hits, nullptr, 0, stream, counters);
}

ResultCode rtbvh_gpu_trace_tree_occluded(RTTraceTree tree, const RTRay* rays, size_t ray_count,
                                         uint8_t* occluded, uint32_t flags, void* stream) {
    if (occluded == nullptr && ray_count > 0) {
        set_error("occluded is null");
        return Error;
    }
    return tree_entry(tree, rays, ray_count, nullptr, occluded, flags, stream);
}

ResultCode rtbvh_gpu_intersect_batch(RTBatch batch, const RTRay* rays, const uint32_t* ray_offsets,
                                     size_t ray_count, RTHit* hits, uint32_t flags, void* stream) {
    if (hits == nullptr && ray_count > 0) {
        set_error("hits is null");
        return Error;
    }
    return batch_entry(batch, rays, ray_offsets, ray_count, hits, nullptr, flags, stream);
}

ResultCode rtbvh_gpu_occluded_batch(RTBatch batch, const RTRay* rays, const uint32_t* ray_offsets,
                                    size_t ray_count, uint8_t* occluded, uint32_t flags,
                                    void* stream) {
    if (occluded == nullptr && ray_count > 0) {
        set_error("occluded is null");
        return Error;
    }
    return batch_entry(batch, rays, ray_offsets, ray_count, nullptr, occluded, flags, stream);
}

ResultCode rtbvh_gpu_occluded_mbvh_prepared(RTMbvh mbvh, const RTRay* rays, size_t ray_count,
                                            uint8_t* occluded, uint32_t flags, void* stream) {
    if (occluded == nullptr && ray_count > 0) {
        set_error("occluded is null");
        return Error;
    }
    return prepared_entry(mbvh, rays, ray_count, nullptr, occluded, flags, stream);
}

ResultCode rtbvh_gpu_occluded(RTBvh bvh, const float* vertices, size_t prim_count,
                              const RTRay* rays, size_t ray_count, uint8_t* occluded,
                              uint32_t flags, void* stream) {
    if (occluded == nullptr && ray_count > 0) {
        set_error("occluded is null");
        return Error;
    }
    return trace_entry(bvh.nodes, bvh.node_count, bvh.indices, bvh.index_count, vertices,
                       prim_count, rays, ray_count, nullptr, occluded, flags, stream, false);
}

ResultCode rtbvh_gpu_occluded_mbvh(RTMbvh mbvh, const float* vertices, size_t prim_count,
                                   const RTRay* rays, size_t ray_count, uint8_t* occluded,
                                   uint32_t flags, void* stream) {
    if (occluded == nullptr && ray_count > 0) {
        set_error("occluded is null");
        return Error;
    }
    return trace_entry(mbvh.nodes, mbvh.node_count, mbvh.indices, mbvh.index_count, vertices,
                       prim_count, rays, ray_count, nullptr, occluded, flags, stream, true);
}

ResultCode rtbvh_gpu_rays_from_packets(const RTRayPacket4* packets, size_t packet_count,
                                       RTRay* rays, void* stream) {
    if (packet_count == 0) return Ok;
    if (packets == nullptr || rays == nullptr) {
        set_error("packets and rays must not be null");
        return Error;
    }
    ResultCode rc;
    if ((rc = check_device(packets, "packets")) != Ok || (rc = check_device(rays, "rays")) != Ok)
        return rc;
    if ((reinterpret_cast<uintptr_t>(packets) & 15) || (reinterpret_cast<uintptr_t>(rays) & 15)) {
        set_error("packets and rays must be 16-byte aligned");
        return InvalidArgument;
    }
    if (packet_count > (uint64_t(0x7FFFFFFF) * 64)) {
        set_error("too many packets (%zu)", packet_count);
        return InvalidArgument;
    }
    const uint64_t n = uint64_t(packet_count) * 4;
    hipLaunchKernelGGL(packets_to_rays_kernel, dim3(uint32_t((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), reinterpret_cast<const float4*>(packets),
                       uint64_t(packet_count), reinterpret_cast<float4*>(rays));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "packets to rays");
    return Ok;
}

ResultCode rtbvh_gpu_set_trace_config(const RTTraceConfig* config) {
    if (config == nullptr) {
        set_trace_config(nullptr);
        return Ok;
    }
    const TraceConfig c{config->block, config->lds_stack, config->occupancy_pct, config->kernel,
                        config->refill, config->schedule, config->postpone};
    if (!valid_trace_config(c)) {
        set_error("trace config: block 64/128/256, lds_stack <= 32, occupancy_pct 10..3200, "
                  "kernel 0/1, refill 1..64, schedule 0/1, postpone 0..64");
        return InvalidArgument;
    }
    set_trace_config(&c);
    return Ok;
}

ResultCode rtbvh_gpu_get_trace_config(RTTraceConfig* config) {
    if (config == nullptr) {
        set_error("config must not be null");
        return Error;
    }
    const TraceConfig c = get_trace_config();
    config->block = c.block;
    config->lds_stack = c.lds_depth;
    config->occupancy_pct = c.occupancy_pct;
    config->kernel = c.kernel;
    config->refill = c.refill;
    config->schedule = c.schedule;
    config->postpone = c.postpone;
    return Ok;
}

}  // extern "C"

namespace rtbvh {
namespace {

// Times every candidate launch shape of `trace` (closest hit of the caller's
// rays, best of two runs after a warm-up) and keeps the fastest as the
// device's configuration.
template <typename Trace>
ResultCode autotune(const RTRay* rays, size_t ray_count, RTHit* hits, void* stream,
                    RTTraceConfig* best, bool compressed, Trace trace) {
    if (rays == nullptr || hits == nullptr || ray_count == 0) {
        set_error("autotuning needs rays and hits");
        return Error;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return hip_fail(e, "autotune events");
    }
    const TraceConfig before = get_trace_config();
    // compressed tree (one kernel): wave-range schedule with block x LDS
    // depth x grid x refill, and the static schedule at the default shape;
    // 4-wide trees, one-item kernel: block x LDS depth x grid x refill x
    // schedule; per-node-leaves kernel: block x LDS depth x grid
    std::vector<TraceConfig> shapes;
    if (compressed) {
        for (uint32_t b : {64u, 128u})
            for (uint32_t d : {6u, 8u, 10u})
                for (uint32_t o : {400u, 800u, 1600u})
                    for (uint32_t rf : {8u, 16u, 32u})
                        shapes.push_back(TraceConfig{b, d, o, 0, rf, 1, 0});
        shapes.push_back(TraceConfig{64, 8, 800, 0, 1, 0, 0});
    } else {
        for (uint32_t b : {64u, 128u, 256u})
            for (uint32_t d : {4u, 8u, 12u, 16u})
                for (uint32_t o : {200u, 400u, 800u}) {
                    for (uint32_t rf : {1u, 4u, 16u})
                        for (uint32_t sc : {0u, 1u})
                            shapes.push_back(TraceConfig{b, d, o, 0, rf, sc});
                    shapes.push_back(TraceConfig{b, d, o, 1, 1, 0});
                }
    }
    TraceConfig top = before;
    float top_ms = 3.4e38f;
    ResultCode rc = Ok;
    for (const TraceConfig& c : shapes) {
        set_trace_config(&c);
        float ms_best = 3.4e38f;
        for (int rep = 0; rep < 3 && rc == Ok; ++rep) {
            if ((e = hipEventRecord(e0, st)) != hipSuccess) break;
            rc = trace();
            if (rc != Ok) break;
            if ((e = hipEventRecord(e1, st)) != hipSuccess ||
                (e = hipEventSynchronize(e1)) != hipSuccess)
                break;
            float ms = 0.f;
            if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            if (rep > 0 && ms < ms_best) ms_best = ms;  // first run warms up
        }
        if (rc != Ok || e != hipSuccess) break;
        if (ms_best < top_ms) {
            top_ms = ms_best;
            top = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != Ok || e != hipSuccess) {
        set_trace_config(&before);
        return rc != Ok ? rc : hip_fail(e, "autotune");
    }
    set_trace_config(&top);
    if (best) {
        best->block = top.block;
        best->lds_stack = top.lds_depth;
        best->occupancy_pct = top.occupancy_pct;
        best->kernel = top.kernel;
        best->refill = top.refill;
        best->schedule = top.schedule;
        best->postpone = top.postpone;
    }
    return Ok;
}

}  // namespace
}  // namespace rtbvh

extern "C" {

ResultCode rtbvh_gpu_autotune_mbvh(RTMbvh mbvh, const RTRay* rays, size_t ray_count, RTHit* hits,
                                   void* stream, RTTraceConfig* best) {
    return autotune(rays, ray_count, hits, stream, best, false, [&]() {
        return prepared_entry(mbvh, rays, ray_count, hits, nullptr, 0, stream);
    });
}

ResultCode rtbvh_gpu_autotune_trace_tree(RTTraceTree tree, const RTRay* rays, size_t ray_count,
                                         RTHit* hits, void* stream, RTTraceConfig* best) {
    BatchView v;
    {
        RegistryHold hold;  // released before the timed calls take their own
        if (!trace_tree_view(tree.id, &v, hold)) {
            set_error("unknown trace tree id %u", tree.id);
            return Error;
        }
    }
    return autotune(rays, ray_count, hits, stream, best, v.cnodes != nullptr, [&]() {
        return tree_entry(tree, rays, ray_count, hits, nullptr, 0, stream);
    });
}

ResultCode rtbvh_gpu_autotune_batch(RTBatch batch, const RTRay* rays, const uint32_t* ray_offsets,
                                    size_t ray_count, RTHit* hits, void* stream,
                                    RTTraceConfig* best) {
    BatchView v;
    {
        RegistryHold hold;  // released before the timed calls take their own
        if (!batch_view(batch.id, &v, hold)) {
            set_error("unknown batch id %u", batch.id);
            return Error;
        }
    }
    return autotune(rays, ray_count, hits, stream, best, v.cnodes != nullptr, [&]() {
        return batch_entry(batch, rays, ray_offsets, ray_count, hits, nullptr, 0, stream);
    });
}

ResultCode rtbvh_gpu_status(void* stream) {
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    uint32_t* p = device_status_ptr();
    if (p == nullptr) {
        set_error("device status word unavailable");
        return DeviceError;
    }
    uint32_t v = 0;
    e = hipMemcpy(&v, p, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "status read");
    if (v != 0) {
        const uint32_t zero = 0;
        e = hipMemcpy(p, &zero, sizeof(zero), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "status reset");
    }
    if (v & kStatusStackOverflow) {
        set_error("a ray exceeded RTBVH_MAX_STACK traversal stack entries");
        return StackOverflow;
    }
    return Ok;
}

}  // extern "C"
