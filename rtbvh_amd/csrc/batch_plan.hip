// C ABI: capturable build + trace of a batch of scenes (rtbvh_gpu_batch_plan_*).
//
// No counterpart in the reference (one scene per call, built on the CPU).  A
// plan owns every buffer of the PLOC build, the traversal tree (4-wide, leaf-
// ordered triangles) and the traversal for fixed sizes, so that
// rtbvh_gpu_batch_plan_enqueue allocates nothing and never waits for the
// device; rtbvh_gpu_batch_plan_launch records that sequence once into a
// hipGraph and replays it.  The numbers of clustering iterations and of
// tree levels are fixed at creation from an eager build of the given inputs
// plus headroom; a later build that needs more sets the plan's fail word,
// the traversal behind it then does nothing and rtbvh_gpu_batch_plan_status
// reports the failure.
#include <stdio.h>

#include <memory>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace rtbvh {

namespace {

struct GraphKey {
    const void* aabbs;
    const void* centers;
    const void* vertices;
    const void* rays;
    const void* ray_offsets;
    size_t ray_count;
    void* hits;
    void* occluded;
    bool operator==(const GraphKey& o) const {
        return aabbs == o.aabbs && centers == o.centers && vertices == o.vertices &&
               rays == o.rays && ray_offsets == o.ray_offsets && ray_count == o.ray_count &&
               hits == o.hits && occluded == o.occluded;
    }
};

struct PlanEntry {
    int device = 0;
    uint32_t scenes = 1, prims = 0;
    uint32_t levels = 0;  // traversal-tree levels per build
    PlocPlanImpl* ploc = nullptr;
    TraceTreePlan* tt = nullptr;
    DeviceBvh tree;    // view into the PLOC plan
    DeviceMbvh ttree;  // traversal tree: nodes (capacity) and triangles, owned
    hipStream_t stream = nullptr;  // graph launches
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t exec = nullptr;
    GraphKey key{};
    void* order_buf = nullptr;  // rtbvh_gpu_batch_plan_order_rays buffers (null: caller's order)
    RayOrder order{};
    ~PlanEntry() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (order_buf) (void)hipFree(order_buf);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) (void)hipStreamDestroy(stream);
        if (ttree.nodes) (void)hipFree(ttree.nodes);
        if (ttree.cnodes) (void)hipFree(ttree.cnodes);
        if (ttree.tris) (void)hipFree(ttree.tris);
        trace_tree_plan_destroy(tt);
        ploc_plan_destroy(ploc);
    }
};

struct Plans {
    std::mutex mu;
    std::vector<std::unique_ptr<PlanEntry>> v;
};

Plans& plans() {
    static Plans p;
    return p;
}

PlanEntry* find_plan(uint32_t id) {
    Plans& p = plans();
    std::lock_guard<std::mutex> lk(p.mu);
    if (id >= p.v.size()) return nullptr;
    return p.v[id].get();
}

bool on_device(const void* ptr) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified;
}

ResultCode check_args(const PlanEntry& pe, const RTAabb* aabbs, const float* centers,
                      const float* vertices, const RTRay* rays, const uint32_t* ray_offsets,
                      size_t ray_count, RTHit* hits, uint8_t* occluded) {
    if (centers == nullptr || vertices == nullptr || (rays == nullptr && ray_count > 0)) {
        set_error("centers, vertices and rays must not be null");
        return Error;
    }
    if ((hits == nullptr) == (occluded == nullptr)) {
        set_error("exactly one of hits and occluded must be given");
        return InvalidArgument;
    }
    if (pe.scenes > 1 && ray_offsets == nullptr) {
        set_error("ray_offsets must not be null for a plan of %u scenes", pe.scenes);
        return Error;
    }
    if (ray_count > 0xFFFFFFFFull) {
        set_error("a batch traces at most 2^32-1 rays (got %zu)", ray_count);
        return InvalidArgument;
    }
    if (pe.order_buf && ray_count > pe.order.capacity) {
        set_error("the plan orders at most %llu rays (got %zu)",
                  (unsigned long long)pe.order.capacity, ray_count);
        return InvalidArgument;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != pe.device) {
        set_error("plan lives on device %d, current device is %d", pe.device, cur);
        return InvalidArgument;
    }
    const void* ptrs[] = {aabbs, centers, vertices, rays, ray_offsets, hits, occluded};
    for (const void* q : ptrs) {
        if (q && !on_device(q)) {
            set_error("argument %p is not device memory", q);
            return InvalidArgument;
        }
    }
    if ((reinterpret_cast<uintptr_t>(rays) & 15) || (reinterpret_cast<uintptr_t>(hits) & 15)) {
        set_error("rays and hits must be 16-byte aligned");
        return InvalidArgument;
    }
    return Ok;
}

// Build, traversal tree and traversal, stream-ordered, without allocation or
// host synchronisation.
hipError_t enqueue_all(PlanEntry& pe, const RTAabb* aabbs, const float* centers,
                       const float* vertices, const RTRay* rays, const uint32_t* ray_offsets,
                       size_t ray_count, RTHit* hits, uint8_t* occluded, hipStream_t stream) {
    hipError_t e = ploc_plan_enqueue(pe.ploc, aabbs, centers, stream);
    if (e != hipSuccess) return e;
    uint32_t* fail = const_cast<uint32_t*>(ploc_plan_fail_word(pe.ploc));
    if ((e = trace_tree_enqueue(pe.tt, pe.tree, vertices, pe.prims,
                                pe.scenes > 1 ? ploc_plan_scene_offsets(pe.ploc) : nullptr,
                                pe.scenes, &pe.ttree, pe.levels, fail, stream)) != hipSuccess)
        return e;
    if (ray_count == 0) return hipSuccess;
    TraceParams p{};
    p.mnodes = pe.ttree.nodes;
    p.cnodes = pe.ttree.cnodes;
    p.node_count = pe.ttree.node_count;
    p.index_count = pe.prims;
    p.indices = nullptr;
    p.tris = pe.ttree.tris;
    p.rays = reinterpret_cast<const DevRay*>(rays);
    p.ray_count = ray_count;
    p.hits = reinterpret_cast<DevHit*>(hits);
    p.occluded = occluded;
    p.ray_offsets = pe.scenes > 1 ? ray_offsets : nullptr;
    p.scene_count = pe.scenes;
    p.gate = fail;
    p.small_leaves = true;  // traversal-tree leaves
    p.status = device_status_ptr();
    if (p.status == nullptr) return hipErrorInvalidValue;
    if (!pe.order_buf) return launch_trace_fast(p, occluded != nullptr, stream);
    // one scene, compressed tree: the traversal reads the rays through the
    // order the sort leaves and writes each result at the ray's own index
    static_assert(kBatchTreeWidth == 8, "ordered plans need the compressed tree's kernel");
    const uint32_t* ids = nullptr;
    if ((e = ray_order_sort(pe.order, reinterpret_cast<const DevRay*>(rays), ray_count, stream,
                            &ids)) != hipSuccess)
        return e;
    p.ray_ids = ids;
    return launch_trace_fast(p, occluded != nullptr, stream);
}

}  // namespace
}  // namespace rtbvh

using namespace rtbvh;

extern "C" {

ResultCode rtbvh_gpu_batch_plan_create(const RTAabb* aabbs, const float* centers,
                                       size_t center_stride, const float* vertices,
                                       const uint32_t* scene_offsets, size_t scene_count,
                                       void* stream, RTBatchPlan* result) {
    if (centers == nullptr || scene_offsets == nullptr || result == nullptr) {
        set_error("centers, scene_offsets and result must not be null");
        return Error;
    }
    if (center_stride != 12 && center_stride != 16) {
        set_error("center_stride must be 12 or 16 bytes, got %zu", center_stride);
        return InvalidArgument;
    }
    if (scene_count > RTBVH_MAX_BATCH_SCENES) {
        set_error("a batch holds at most %u scenes (got %zu)", RTBVH_MAX_BATCH_SCENES,
                  scene_count);
        return InvalidArgument;
    }
    const uint32_t scenes = uint32_t(scene_count);
    ResultCode rc = check_batch_offsets(scene_offsets, scenes);
    if (rc != Ok) return rc;
    std::unique_ptr<PlanEntry> pe(new PlanEntry());
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return NoDevice;
    }
    if (hipGetDevice(&pe->device) != hipSuccess) return hip_fail(hipErrorInvalidDevice, "device");
    if (!on_device(centers) || (aabbs && !on_device(aabbs)) ||
        (vertices && !on_device(vertices))) {
        set_error("centers, aabbs and vertices must be device memory");
        return InvalidArgument;
    }
    pe->scenes = scenes;
    pe->prims = scene_offsets[scenes];
    hipStream_t st = static_cast<hipStream_t>(stream);
    rc = ploc_plan_create(aabbs, centers, center_stride / 4, scene_offsets, scenes, st, &pe->ploc,
                          &pe->tree);
    if (rc != Ok) return rc;
    // the traversal tree, built once eagerly to size its level budget
    rc = trace_tree_plan_create(pe->tree, kDefaultMaxLeaf, kBatchTreeWidth, &pe->tt, &pe->ttree);
    if (rc != Ok) return rc;
    const uint32_t capacity = pe->ttree.node_count;
    if (vertices == nullptr) {
        set_error("vertices must not be null");
        return Error;
    }
    rc = trace_tree_build(pe->tt, pe->tree, vertices, pe->prims,
                          scenes > 1 ? ploc_plan_scene_offsets(pe->ploc) : nullptr, scenes,
                          &pe->ttree, st);
    if (rc != Ok) return rc;
    pe->ttree.node_count = capacity;  // replays may need more nodes than this build
    const uint32_t lv = trace_tree_levels(pe->tt);
    pe->levels = lv + (lv / 4 > 4 ? lv / 4 : 4);
    hipError_t e = hipStreamCreateWithFlags(&pe->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pe->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pe->ev_out, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "batch plan");
    Plans& p = plans();
    std::lock_guard<std::mutex> lk(p.mu);
    p.v.push_back(std::move(pe));
    result->id = uint32_t(p.v.size() - 1);
    result->scene_count = scenes;
    result->prim_count = p.v.back()->prims;
    result->iterations = ploc_plan_iterations(p.v.back()->ploc);
    return Ok;
}

ResultCode rtbvh_gpu_batch_plan_enqueue(RTBatchPlan plan, const RTAabb* aabbs,
                                        const float* centers, const float* vertices,
                                        const RTRay* rays, const uint32_t* ray_offsets,
                                        size_t ray_count, RTHit* hits, uint8_t* occluded,
                                        void* stream) {
    PlanEntry* pe = find_plan(plan.id);
    if (pe == nullptr) {
        set_error("unknown plan id %u", plan.id);
        return Error;
    }
    ResultCode rc =
        check_args(*pe, aabbs, centers, vertices, rays, ray_offsets, ray_count, hits, occluded);
    if (rc != Ok) return rc;
    const hipError_t e = enqueue_all(*pe, aabbs, centers, vertices, rays, ray_offsets, ray_count,
                                     hits, occluded, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "batch plan enqueue");
    return Ok;
}

ResultCode rtbvh_gpu_batch_plan_launch(RTBatchPlan plan, const RTAabb* aabbs, const float* centers,
                                       const float* vertices, const RTRay* rays,
                                       const uint32_t* ray_offsets, size_t ray_count, RTHit* hits,
                                       uint8_t* occluded, void* stream) {
    PlanEntry* pe = find_plan(plan.id);
    if (pe == nullptr) {
        set_error("unknown plan id %u", plan.id);
        return Error;
    }
    ResultCode rc =
        check_args(*pe, aabbs, centers, vertices, rays, ray_offsets, ray_count, hits, occluded);
    if (rc != Ok) return rc;
    const GraphKey key{aabbs, centers, vertices, rays, ray_offsets, ray_count, hits, occluded};
    hipStream_t user = static_cast<hipStream_t>(stream);
    // everything below runs on the plan's stream after the caller's prior
    // work: the inputs may still be written there, and the outputs may be
    // memory the caller's stream is still using
    hipError_t e;
    if ((e = hipEventRecord(pe->ev_in, user)) != hipSuccess ||
        (e = hipStreamWaitEvent(pe->stream, pe->ev_in, 0)) != hipSuccess)
        return hip_fail(e, "order after the caller's stream");
    if (pe->exec == nullptr || !(pe->key == key)) {
        if (pe->exec) {
            (void)hipGraphExecDestroy(pe->exec);
            pe->exec = nullptr;
        }
        // one eager pass first: every launch configuration (kernel attributes,
        // occupancy) is resolved outside the capture
        if ((e = enqueue_all(*pe, aabbs, centers, vertices, rays, ray_offsets, ray_count, hits,
                             occluded, pe->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(pe->stream)) != hipSuccess)
            return hip_fail(e, "batch plan warm-up");
        if ((e = hipStreamBeginCapture(pe->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return hip_fail(e, "begin capture");
        const hipError_t ce = enqueue_all(*pe, aabbs, centers, vertices, rays, ray_offsets,
                                          ray_count, hits, occluded, pe->stream);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(pe->stream, &graph);
        if (ce != hipSuccess || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_fail(ce != hipSuccess ? ce : e, "capture");
        }
        e = hipGraphInstantiate(&pe->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            pe->exec = nullptr;
            return hip_fail(e, "graph instantiate");
        }
        pe->key = key;
    }
    if ((e = hipGraphLaunch(pe->exec, pe->stream)) != hipSuccess ||
        (e = hipEventRecord(pe->ev_out, pe->stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(user, pe->ev_out, 0)) != hipSuccess)
        return hip_fail(e, "graph launch");
    return Ok;
}

ResultCode rtbvh_gpu_batch_plan_status(RTBatchPlan plan, void* stream) {
    PlanEntry* pe = find_plan(plan.id);
    if (pe == nullptr) {
        set_error("unknown plan id %u", plan.id);
        return Error;
    }
    uint32_t fail = 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = hipStreamSynchronize(pe->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&fail, ploc_plan_fail_word(pe->ploc), 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "plan status");
    if (fail) {
        set_error("the build needed more than the plan's %u clustering iterations; create a new "
                  "plan", ploc_plan_iterations(pe->ploc));
        return Error;
    }
    return rtbvh_gpu_status(stream);
}

ResultCode rtbvh_gpu_batch_plan_order_rays(RTBatchPlan plan, size_t max_ray_count, void* stream) {
    PlanEntry* pe = find_plan(plan.id);
    if (pe == nullptr) {
        set_error("unknown plan id %u", plan.id);
        return Error;
    }
    if (max_ray_count > 0 && pe->scenes != 1) {
        set_error("rays are ordered for one-scene plans only (this plan has %u scenes)",
                  pe->scenes);
        return InvalidArgument;
    }
    if (max_ray_count > kRadixPairsMax) {
        set_error("a plan orders at most %u rays (got %zu)", kRadixPairsMax, max_ray_count);
        return InvalidArgument;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != pe->device) {
        set_error("plan lives on device %d, current device is %d", pe->device, cur);
        return InvalidArgument;
    }
    // earlier frames (graph replays on the plan's stream, eager enqueues on
    // the caller's) may still use the old buffers
    hipError_t e = hipStreamSynchronize(pe->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "order rays");
    if (pe->exec) {
        (void)hipGraphExecDestroy(pe->exec);
        pe->exec = nullptr;
    }
    if (pe->order_buf) {
        (void)hipFree(pe->order_buf);
        pe->order_buf = nullptr;
        pe->order = RayOrder{};
    }
    if (max_ray_count == 0) return Ok;
    void* buf = nullptr;
    if ((e = hipMalloc(&buf, ray_order_bytes(max_ray_count, false))) != hipSuccess)
        return hip_fail(e, "ray order buffers");
    pe->order_buf = buf;
    pe->order = ray_order_carve(buf, max_ray_count, false);
    return Ok;
}

void rtbvh_gpu_batch_plan_free(RTBatchPlan plan) {
    Plans& p = plans();
    std::lock_guard<std::mutex> lk(p.mu);
    if (plan.id >= p.v.size() || !p.v[plan.id]) {
        fprintf(stderr, "Could not free plan with id: %u\n", plan.id);
        return;
    }
    (void)hipStreamSynchronize(p.v[plan.id]->stream);
    (void)hipDeviceSynchronize();  // eager enqueues may still read the buffers
    p.v[plan.id].reset();
}

}  // extern "C"
