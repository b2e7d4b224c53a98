// C ABI: tree construction, collapse, refit and release — the rtbvh_ffi
// entry points (create_bvh / create_mbvh / refit / free_*) and their
// device-resident rtbvh_gpu_* counterparts.
//
// Trees live in a process-wide registry mirroring rtbvh_ffi's
// StructureManager (rtbvh_ffi/src/lib.rs:8-127): ids are handed out in order
// and never reused, a freed slot stays empty, creation and release take the
// exclusive lock and lookups the shared lock.  Every tree's arrays live on
// the device it was built on; handles returned by the rtbvh_ffi entry points
// additionally carry a host mirror that their pointers refer to.
#include <stdio.h>

#include <memory>
#include <mutex>
#include <shared_mutex>
#include <vector>

#include "internal.hpp"

namespace rtbvh {

namespace {

struct BvhEntry {
    DeviceBvh dev;
    int device = 0;
    std::vector<RTBvhNode> host_nodes;
    std::vector<uint32_t> host_indices;
};

struct MbvhEntry {
    DeviceMbvh dev;
    int device = 0;
    std::vector<RTMbvhNode> host_nodes;
    std::vector<uint32_t> host_indices;
};

struct TraceTreeEntry {
    DeviceMbvh t;  // nodes + leaf-ordered triangles
    int device = 0;
};

struct BatchEntry {
    DeviceBvh bvh;    // forest, roots [0, scenes)
    DeviceMbvh ttree; // traversal tree over the batch (after attach)
    uint32_t* scene_off = nullptr;  // device, scenes + 1 primitive offsets
    uint32_t scenes = 0;
    int device = 0;
};

struct Registry {
    std::shared_mutex mu;
    std::vector<std::unique_ptr<BvhEntry>> bvhs;    // null = freed
    std::vector<std::unique_ptr<MbvhEntry>> mbvhs;
    std::vector<std::unique_ptr<BatchEntry>> batches;
    std::vector<std::unique_ptr<TraceTreeEntry>> ttrees;
};

Registry& reg() {
    static Registry r;
    return r;
}

void release(DeviceBvh& d) {
    if (d.nodes) (void)hipFree(d.nodes);
    if (d.indices) (void)hipFree(d.indices);
    if (d.parents) (void)hipFree(d.parents);
    if (d.order) (void)hipFree(d.order);
    d = DeviceBvh{};
}

void release(DeviceMbvh& d) {
    if (d.nodes) (void)hipFree(d.nodes);
    if (d.cnodes) (void)hipFree(d.cnodes);
    if (d.indices) (void)hipFree(d.indices);
    if (d.tris) (void)hipFree(d.tris);
    d = DeviceMbvh{};
}

void release(BatchEntry& b) {
    release(b.bvh);
    release(b.ttree);
    if (b.scene_off) (void)hipFree(b.scene_off);
    b.scene_off = nullptr;
}

// Frees are ordered after all work already enqueued on the device, which may
// still read the arrays (hipFree alone is not relied on for that).
void drain_device() { (void)hipDeviceSynchronize(); }

// Switches to `device` for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int device) {
        if (hipGetDevice(&prev) == hipSuccess && prev != device) (void)hipSetDevice(device);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

ResultCode current_device(int* dev) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return NoDevice;
    }
    e = hipGetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    return Ok;
}

RTBvh handle_of(uint32_t id, const BvhEntry& e, bool host) {
    RTBvh h;
    h.id = id;
    h.node_count = e.dev.node_count;
    h.index_count = e.dev.prim_count;
    if (host) {
        h.nodes = e.host_nodes.data();
        h.indices = e.host_indices.data();
    } else {
        h.nodes = reinterpret_cast<const RTBvhNode*>(e.dev.nodes);
        h.indices = e.dev.indices;
    }
    return h;
}

RTMbvh handle_of(uint32_t id, const MbvhEntry& e, bool host) {
    RTMbvh h;
    h.id = id;
    h.node_count = e.dev.node_count;
    h.index_count = e.dev.prim_count;
    if (host) {
        h.nodes = e.host_nodes.data();
        h.indices = e.host_indices.data();
    } else {
        h.nodes = reinterpret_cast<const RTMbvhNode*>(e.dev.nodes);
        h.indices = e.dev.indices;
    }
    return h;
}

// Argument checks shared by create_bvh and rtbvh_gpu_create_bvh, done before
// any device call.
ResultCode validate_build(size_t prim_count, size_t center_stride, BvhType bvh_type) {
    if (center_stride != 12 && center_stride != 16) {
        // rtbvh_ffi asserts 12 <= stride <= 16, stride % 4 == 0 (lib.rs:441-449)
        set_error("center_stride must be 12 or 16 bytes, got %zu", center_stride);
        return InvalidArgument;
    }
    if (bvh_type != LocallyOrderedClustered && bvh_type != BinnedSAH) {
        set_error("unknown BvhType %u", unsigned(bvh_type));
        return InvalidArgument;
    }
    if (prim_count == 0) {
        set_error("no primitives");
        return NoPrimitives;  // Builder::construct_* (src/bvh.rs:88-90,114-116)
    }
    return Ok;
}

ResultCode build_device(const RTAabb* aabbs, size_t prim_count, const float* centers,
                        size_t center_stride, size_t prims_per_leaf, BvhType bvh_type,
                        hipStream_t stream, DeviceBvh* out) {
    ResultCode rc = validate_build(prim_count, center_stride, bvh_type);
    if (rc != Ok) return rc;
    if (bvh_type == LocallyOrderedClustered)
        return gpu_build_ploc(aabbs, centers, center_stride / 4, prim_count, 14, stream, out);
    if (bvh_type == BinnedSAH)
        return gpu_build_binned(aabbs, centers, center_stride / 4, prim_count, prims_per_leaf,
                                stream, out);
    set_error("unknown BvhType %u", unsigned(bvh_type));
    return InvalidArgument;
}

uint32_t register_bvh(std::unique_ptr<BvhEntry> e) {
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    r.bvhs.push_back(std::move(e));
    return uint32_t(r.bvhs.size() - 1);
}

uint32_t register_mbvh(std::unique_ptr<MbvhEntry> e) {
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    r.mbvhs.push_back(std::move(e));
    return uint32_t(r.mbvhs.size() - 1);
}

ResultCode download_bvh(BvhEntry& e) {
    e.host_nodes.resize(e.dev.node_count);
    e.host_indices.resize(e.dev.prim_count);
    hipError_t err = hipSuccess;
    if (e.dev.node_count)
        err = hipMemcpy(e.host_nodes.data(), e.dev.nodes, size_t(e.dev.node_count) * 32,
                        hipMemcpyDeviceToHost);
    if (err == hipSuccess && e.dev.prim_count)
        err = hipMemcpy(e.host_indices.data(), e.dev.indices, size_t(e.dev.prim_count) * 4,
                        hipMemcpyDeviceToHost);
    if (err != hipSuccess) return hip_fail(err, "download bvh");
    return Ok;
}

ResultCode download_mbvh(MbvhEntry& e) {
    e.host_nodes.resize(e.dev.node_count);
    e.host_indices.resize(e.dev.prim_count);
    hipError_t err = hipSuccess;
    if (e.dev.node_count)
        err = hipMemcpy(e.host_nodes.data(), e.dev.nodes, size_t(e.dev.node_count) * 128,
                        hipMemcpyDeviceToHost);
    if (err == hipSuccess && e.dev.prim_count)
        err = hipMemcpy(e.host_indices.data(), e.dev.indices, size_t(e.dev.prim_count) * 4,
                        hipMemcpyDeviceToHost);
    if (err != hipSuccess) return hip_fail(err, "download mbvh");
    return Ok;
}

// Finds the entry for `id`; the shared lock is held by the caller.
BvhEntry* find_bvh(uint32_t id) {
    Registry& r = reg();
    if (id >= r.bvhs.size()) return nullptr;
    return r.bvhs[id].get();
}

MbvhEntry* find_mbvh(uint32_t id) {
    Registry& r = reg();
    if (id >= r.mbvhs.size()) return nullptr;
    return r.mbvhs[id].get();
}

BatchEntry* find_batch(uint32_t id) {
    Registry& r = reg();
    if (id >= r.batches.size()) return nullptr;
    return r.batches[id].get();
}

RTBatch handle_of(uint32_t id, const BatchEntry& e) {
    RTBatch h{};
    h.id = id;
    h.scene_count = e.scenes;
    h.bvh_node_count = e.bvh.node_count;
    h.node_count = e.ttree.node_count;
    h.bvh_nodes = reinterpret_cast<const RTBvhNode*>(e.bvh.nodes);
    h.nodes = e.ttree.cnodes ? static_cast<const void*>(e.ttree.cnodes)
                             : static_cast<const void*>(e.ttree.nodes);
    h.index_count = e.bvh.prim_count;
    h.indices = e.bvh.indices;
    return h;
}

bool is_device_pointer(const void* p) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified;
}

ResultCode collapse_entry(uint32_t id, void* stream, bool host, RTMbvh* out) {
    std::unique_ptr<MbvhEntry> me(new MbvhEntry());
    {
        Registry& r = reg();
        // exclusive: the collapse may attach parent pointers to the tree
        std::unique_lock<std::shared_mutex> lk(r.mu);
        BvhEntry* be = find_bvh(id);
        if (be == nullptr) {
            set_error("unknown bvh id %u", id);
            return Error;
        }
        DeviceGuard g(be->device);
        me->device = be->device;
        ResultCode rc = gpu_build_mbvh(be->dev, static_cast<hipStream_t>(stream), &me->dev);
        if (rc != Ok) return rc;
    }
    if (host) {
        DeviceGuard g(me->device);
        ResultCode rc = download_mbvh(*me);
        if (rc != Ok) {
            release(me->dev);
            return rc;
        }
    }
    MbvhEntry* raw = me.get();
    const uint32_t mid = register_mbvh(std::move(me));
    *out = handle_of(mid, *raw, host);
    return Ok;
}

}  // namespace

// Attached leaf-ordered triangles of an MBVH made by this library (null if
// none); *count receives the number of index entries they cover.
const float4* attached_triangles(uint32_t id, uint32_t* count, int* device, bool* small,
                                 RegistryHold& hold) {
    Registry& r = reg();
    if (!hold.owns_lock()) hold = RegistryHold(r.mu);
    MbvhEntry* me = find_mbvh(id);
    if (me == nullptr || me->dev.tris == nullptr) return nullptr;
    *count = me->dev.prim_count;
    *device = me->device;
    *small = me->dev.max_leaf <= 16;
    return me->dev.tris;
}

// The arrays a batched traversal needs (false if the id is unknown); tris
// is null until triangles are attached.
bool batch_view(uint32_t id, BatchView* out, RegistryHold& hold) {
    Registry& r = reg();
    if (!hold.owns_lock()) hold = RegistryHold(r.mu);
    BatchEntry* be = find_batch(id);
    if (be == nullptr) return false;
    out->nodes = be->ttree.nodes;
    out->cnodes = be->ttree.cnodes;
    out->node_count = be->ttree.node_count;
    out->indices = nullptr;
    out->index_count = be->ttree.prim_count;
    out->tris = be->ttree.tris;
    out->scenes = be->scenes;
    out->device = be->device;
    return true;
}

bool trace_tree_view(uint32_t id, BatchView* out, RegistryHold& hold) {
    Registry& r = reg();
    if (!hold.owns_lock()) hold = RegistryHold(r.mu);
    if (id >= r.ttrees.size() || !r.ttrees[id]) return false;
    const TraceTreeEntry& te = *r.ttrees[id];
    out->nodes = te.t.nodes;
    out->cnodes = te.t.cnodes;
    out->node_count = te.t.node_count;
    out->indices = nullptr;
    out->index_count = te.t.prim_count;
    out->tris = te.t.tris;
    out->scenes = te.t.root_count;
    out->device = te.device;
    return true;
}

}  // namespace rtbvh

using namespace rtbvh;

extern "C" {

ResultCode rtbvh_gpu_create_bvh(const RTAabb* aabbs, size_t prim_count, const float* centers,
                                size_t center_stride, size_t prims_per_leaf, BvhType bvh_type,
                                void* stream, RTBvh* result) {
    if (centers == nullptr || result == nullptr) {
        set_error("centers and result must not be null");
        return Error;  // rtbvh_ffi/src/lib.rs:437-439
    }
    ResultCode rc = validate_build(prim_count, center_stride, bvh_type);
    if (rc != Ok) return rc;
    std::unique_ptr<BvhEntry> e(new BvhEntry());
    rc = current_device(&e->device);
    if (rc != Ok) return rc;
    if (!is_device_pointer(centers) || (aabbs && !is_device_pointer(aabbs))) {
        set_error("centers and aabbs must be device memory");
        return InvalidArgument;
    }
    rc = build_device(aabbs, prim_count, centers, center_stride, prims_per_leaf, bvh_type,
                      static_cast<hipStream_t>(stream), &e->dev);
    if (rc != Ok) return rc;
    BvhEntry* raw = e.get();
    const uint32_t id = register_bvh(std::move(e));
    *result = handle_of(id, *raw, false);
    return Ok;
}

ResultCode rtbvh_gpu_create_spatial_bvh(const RTAabb* aabbs, size_t prim_count,
                                        const float* centers, size_t center_stride,
                                        const float* vertices, size_t prims_per_leaf,
                                        void* stream, RTBvh* result) {
    if (centers == nullptr || vertices == nullptr || result == nullptr) {
        set_error("centers, vertices and result must not be null");
        return Error;
    }
    ResultCode rc = validate_build(prim_count, center_stride, LocallyOrderedClustered);
    if (rc != Ok) return rc;
    std::unique_ptr<BvhEntry> e(new BvhEntry());
    rc = current_device(&e->device);
    if (rc != Ok) return rc;
    if (!is_device_pointer(centers) || !is_device_pointer(vertices) ||
        (aabbs && !is_device_pointer(aabbs))) {
        set_error("centers, vertices and aabbs must be device memory");
        return InvalidArgument;
    }
    rc = gpu_build_spatial(aabbs, centers, center_stride / 4, vertices, prim_count, prims_per_leaf,
                           static_cast<hipStream_t>(stream), &e->dev);
    if (rc != Ok) return rc;
    BvhEntry* raw = e.get();
    const uint32_t id = register_bvh(std::move(e));
    *result = handle_of(id, *raw, false);
    return Ok;
}

ResultCode rtbvh_gpu_create_mbvh(RTBvh bvh, void* stream, RTMbvh* result) {
    if (result == nullptr) {
        set_error("result must not be null");
        return Error;
    }
    return collapse_entry(bvh.id, stream, false, result);
}

ResultCode rtbvh_gpu_refit(const RTAabb* aabbs, RTBvh bvh, void* stream) {
    if (aabbs == nullptr) {
        set_error("aabbs must not be null");
        return Error;
    }
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    BvhEntry* be = find_bvh(bvh.id);
    if (be == nullptr) {
        set_error("unknown bvh id %u", bvh.id);
        return Error;
    }
    DeviceGuard g(be->device);
    return gpu_refit(be->dev, aabbs, static_cast<hipStream_t>(stream));
}

ResultCode rtbvh_gpu_mbvh_attach_triangles(RTMbvh mbvh, const float* vertices, size_t prim_count,
                                           void* stream) {
    if (vertices == nullptr && prim_count > 0) {
        set_error("vertices must not be null");
        return Error;
    }
    if (prim_count > 0xFFFFFFFFull) {
        set_error("prim_count %zu exceeds 2^32-1", prim_count);
        return InvalidArgument;
    }
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    MbvhEntry* me = find_mbvh(mbvh.id);
    if (me == nullptr) {
        set_error("unknown mbvh id %u", mbvh.id);
        return Error;
    }
    DeviceGuard g(me->device);
    hipPointerAttribute_t attr;
    if (prim_count > 0 && (hipPointerGetAttributes(&attr, vertices) != hipSuccess ||
                           !(attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
                             attr.type == hipMemoryTypeUnified))) {
        (void)hipGetLastError();
        set_error("vertices (%p) is not device memory", static_cast<const void*>(vertices));
        return InvalidArgument;
    }
    const uint32_t entries = me->dev.prim_count;
    if (me->dev.tris == nullptr && entries > 0) {
        hipError_t e = hipMalloc(&me->dev.tris, size_t(entries) * kTriRecord * sizeof(float4));
        if (e != hipSuccess) {
            me->dev.tris = nullptr;
            return hip_fail(e, "hipMalloc(triangles)");
        }
    }
    hipError_t e = launch_prepare_triangles(me->dev.indices, entries, vertices,
                                            uint32_t(prim_count), me->dev.tris,
                                            static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "prepare triangles");
    e = mbvh_max_leaf(me->dev.nodes, me->dev.node_count, &me->dev.max_leaf,
                      static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "largest leaf");
    return Ok;
}

ResultCode rtbvh_gpu_create_batch(const RTAabb* aabbs, const float* centers, size_t center_stride,
                                  const uint32_t* scene_offsets, size_t scene_count, void* stream,
                                  RTBatch* result) {
    if (centers == nullptr || result == nullptr || scene_offsets == nullptr) {
        set_error("centers, scene_offsets and result must not be null");
        return Error;
    }
    if (center_stride != 12 && center_stride != 16) {
        set_error("center_stride must be 12 or 16 bytes, got %zu", center_stride);
        return InvalidArgument;
    }
    if (scene_count == 0) {
        set_error("no scenes");
        return NoPrimitives;
    }
    if (scene_count > RTBVH_MAX_BATCH_SCENES) {
        set_error("a batch holds at most %u scenes (got %zu)", RTBVH_MAX_BATCH_SCENES,
                  scene_count);
        return InvalidArgument;
    }
    std::unique_ptr<BatchEntry> be(new BatchEntry());
    ResultCode rc = current_device(&be->device);
    if (rc != Ok) return rc;
    if (!is_device_pointer(centers) || (aabbs && !is_device_pointer(aabbs))) {
        set_error("centers and aabbs must be device memory");
        return InvalidArgument;
    }
    const uint32_t scenes = uint32_t(scene_count);
    hipStream_t st = static_cast<hipStream_t>(stream);
    rc = gpu_build_ploc_batch(aabbs, centers, center_stride / 4, scene_offsets, scenes, st,
                              &be->bvh);
    if (rc != Ok) return rc;
    be->scenes = scenes;
    hipError_t e = hipMalloc(&be->scene_off, size_t(scenes + 1) * 4);
    if (e == hipSuccess)
        e = hipMemcpyAsync(be->scene_off, scene_offsets, size_t(scenes + 1) * 4,
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        release(*be);
        return hip_fail(e, "scene offsets");
    }
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    r.batches.push_back(std::move(be));
    const uint32_t id = uint32_t(r.batches.size() - 1);
    *result = handle_of(id, *r.batches.back());
    return Ok;
}

ResultCode rtbvh_gpu_batch_attach_triangles(RTBatch* batch, const float* vertices,
                                            size_t prim_count, void* stream) {
    if (vertices == nullptr || batch == nullptr) {
        set_error("batch and vertices must not be null");
        return Error;
    }
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    BatchEntry* be = find_batch(batch->id);
    if (be == nullptr) {
        set_error("unknown batch id %u", batch->id);
        return Error;
    }
    if (prim_count != be->bvh.prim_count) {
        set_error("batch has %u primitives, got vertices for %zu", be->bvh.prim_count,
                  prim_count);
        return InvalidArgument;
    }
    DeviceGuard g(be->device);
    if (!is_device_pointer(vertices)) {
        set_error("vertices (%p) is not device memory", static_cast<const void*>(vertices));
        return InvalidArgument;
    }
    // the traversal tree of the batch (a forest, scene-local ids)
    if (be->ttree.nodes || be->ttree.cnodes) {
        drain_device();  // traversals of the previous tree may be in flight
        release(be->ttree);
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    TraceTreePlan* plan = nullptr;
    ResultCode rc = trace_tree_plan_create(be->bvh, kDefaultMaxLeaf, kBatchTreeWidth, &plan,
                                           &be->ttree, true);
    if (rc != Ok) return rc;
    rc = trace_tree_build(plan, be->bvh, vertices, uint32_t(prim_count), be->scene_off,
                          be->scenes, &be->ttree, st);
    trace_tree_plan_destroy(plan);
    if (rc != Ok) {
        release(be->ttree);
        return rc;
    }
    *batch = handle_of(batch->id, *be);
    return Ok;
}

static ResultCode create_trace_tree(RTBvh bvh, const float* vertices, size_t prim_count,
                                    uint32_t max_leaf, uint32_t width, void* stream,
                                    RTTraceTree* result) {
    if (vertices == nullptr || result == nullptr) {
        set_error("vertices and result must not be null");
        return Error;
    }
    if (prim_count > 0xFFFFFFFFull) {
        set_error("prim_count %zu exceeds 2^32-1", prim_count);
        return InvalidArgument;
    }
    if (max_leaf == 0) max_leaf = kDefaultMaxLeaf;
    std::unique_ptr<TraceTreeEntry> te(new TraceTreeEntry());
    {
        Registry& r = reg();
        std::shared_lock<std::shared_mutex> lk(r.mu);
        BvhEntry* be = find_bvh(bvh.id);
        if (be == nullptr) {
            set_error("unknown bvh id %u", bvh.id);
            return Error;
        }
        DeviceGuard g(be->device);
        if (!is_device_pointer(vertices)) {
            set_error("vertices (%p) is not device memory", static_cast<const void*>(vertices));
            return InvalidArgument;
        }
        te->device = be->device;
        hipStream_t st = static_cast<hipStream_t>(stream);
        TraceTreePlan* plan = nullptr;
        ResultCode rc = trace_tree_plan_create(be->dev, max_leaf, width, &plan, &te->t, true);
        if (rc != Ok) return rc;
        rc = trace_tree_build(plan, be->dev, vertices, uint32_t(prim_count), nullptr, 0, &te->t,
                              st);
        trace_tree_plan_destroy(plan);
        if (rc != Ok) {
            release(te->t);
            return rc;
        }
    }
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    r.ttrees.push_back(std::move(te));
    result->id = uint32_t(r.ttrees.size() - 1);
    result->node_count = r.ttrees.back()->t.node_count;
    result->prim_count = r.ttrees.back()->t.prim_count;
    result->max_leaf = max_leaf;
    result->width = width;
    result->pad_ = 0;
    result->nodes = width == 8 ? static_cast<const void*>(r.ttrees.back()->t.cnodes)
                               : static_cast<const void*>(r.ttrees.back()->t.nodes);
    result->triangles = reinterpret_cast<const float*>(r.ttrees.back()->t.tris);
    return Ok;
}

ResultCode rtbvh_gpu_create_trace_tree(RTBvh bvh, const float* vertices, size_t prim_count,
                                       uint32_t max_leaf, void* stream, RTTraceTree* result) {
    return create_trace_tree(bvh, vertices, prim_count, max_leaf, 4, stream, result);
}

ResultCode rtbvh_gpu_create_trace_tree8(RTBvh bvh, const float* vertices, size_t prim_count,
                                        void* stream, RTTraceTree* result) {
    return create_trace_tree(bvh, vertices, prim_count, 1, 8, stream, result);
}

void rtbvh_gpu_free_trace_tree(RTTraceTree tree) {
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    if (tree.id >= r.ttrees.size() || !r.ttrees[tree.id]) {
        fprintf(stderr, "Could not free trace tree with id: %u\n", tree.id);
        return;
    }
    DeviceGuard g(r.ttrees[tree.id]->device);
    drain_device();
    release(r.ttrees[tree.id]->t);
    r.ttrees[tree.id].reset();
}

void rtbvh_gpu_free_batch(RTBatch batch) {
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    if (batch.id >= r.batches.size() || !r.batches[batch.id]) {
        fprintf(stderr, "Could not free batch with id: %u\n", batch.id);
        return;
    }
    DeviceGuard g(r.batches[batch.id]->device);
    drain_device();
    release(*r.batches[batch.id]);
    r.batches[batch.id].reset();
}

void rtbvh_gpu_free_bvh(RTBvh bvh) { free_bvh(bvh); }
void rtbvh_gpu_free_mbvh(RTMbvh mbvh) { free_mbvh(mbvh); }

// ---------------------------------------------------- rtbvh_ffi entry points

ResultCode create_bvh(const RTAabb* aabbs, size_t prim_count, const float* centers,
                      size_t center_stride, size_t prims_per_leaf, BvhType bvh_type,
                      RTBvh* result) {
    if (centers == nullptr || result == nullptr) {
        set_error("centers and result must not be null");
        return Error;  // rtbvh_ffi/src/lib.rs:437-439
    }
    ResultCode rc = validate_build(prim_count, center_stride, bvh_type);
    if (rc != Ok) return rc;
    std::unique_ptr<BvhEntry> e(new BvhEntry());
    rc = current_device(&e->device);
    if (rc != Ok) return rc;
    // stage the caller's host arrays on the device
    void* d_centers = nullptr;
    void* d_aabbs = nullptr;
    hipError_t err = hipMalloc(&d_centers, prim_count * center_stride);
    if (err == hipSuccess)
        err = hipMemcpy(d_centers, centers, prim_count * center_stride, hipMemcpyHostToDevice);
    if (err == hipSuccess && aabbs) {
        err = hipMalloc(&d_aabbs, prim_count * sizeof(RTAabb));
        if (err == hipSuccess)
            err = hipMemcpy(d_aabbs, aabbs, prim_count * sizeof(RTAabb), hipMemcpyHostToDevice);
    }
    if (err != hipSuccess) {
        if (d_centers) (void)hipFree(d_centers);
        if (d_aabbs) (void)hipFree(d_aabbs);
        return hip_fail(err, "stage inputs");
    }
    rc = build_device(static_cast<const RTAabb*>(d_aabbs), prim_count,
                      static_cast<const float*>(d_centers), center_stride, prims_per_leaf,
                      bvh_type, nullptr, &e->dev);
    (void)hipFree(d_centers);
    if (d_aabbs) (void)hipFree(d_aabbs);
    if (rc != Ok) return rc;
    rc = download_bvh(*e);
    if (rc != Ok) {
        release(e->dev);
        return rc;
    }
    BvhEntry* raw = e.get();
    const uint32_t id = register_bvh(std::move(e));
    *result = handle_of(id, *raw, true);
    return Ok;
}

ResultCode create_mbvh(RTBvh bvh, RTMbvh* mbvh) {
    if (mbvh == nullptr || bvh.nodes == nullptr || bvh.indices == nullptr) {
        set_error("null mbvh, nodes or indices");
        return Error;  // rtbvh_ffi/src/lib.rs:500-502
    }
    return collapse_entry(bvh.id, nullptr, true, mbvh);
}

ResultCode refit(const RTAabb* aabbs, RTBvh bvh) {
    if (aabbs == nullptr || bvh.nodes == nullptr || bvh.indices == nullptr) {
        set_error("null aabbs, nodes or indices");
        return Error;  // rtbvh_ffi/src/lib.rs:520-522
    }
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    BvhEntry* be = find_bvh(bvh.id);
    if (be == nullptr) {
        set_error("unknown bvh id %u", bvh.id);
        return Error;
    }
    DeviceGuard g(be->device);
    // Bvh::prim_count (src/bvh.rs:172-174): one box per primitive, also for
    // a spatial-split tree whose index array holds more entries
    const size_t n = be->dev.source_prims ? be->dev.source_prims : be->dev.prim_count;
    void* d_aabbs = nullptr;
    hipError_t err = hipMalloc(&d_aabbs, n * sizeof(RTAabb) + 16);
    if (err == hipSuccess && n)
        err = hipMemcpy(d_aabbs, aabbs, n * sizeof(RTAabb), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        if (d_aabbs) (void)hipFree(d_aabbs);
        return hip_fail(err, "stage aabbs");
    }
    ResultCode rc = gpu_refit(be->dev, static_cast<const RTAabb*>(d_aabbs), nullptr);
    (void)hipFree(d_aabbs);
    if (rc != Ok) return rc;
    if (!be->host_nodes.empty()) {
        // in place: handles keep pointing at the same host array
        err = hipMemcpy(be->host_nodes.data(), be->dev.nodes, be->host_nodes.size() * 32,
                        hipMemcpyDeviceToHost);
        if (err != hipSuccess) return hip_fail(err, "download refit");
    }
    return Ok;
}

void free_bvh(RTBvh bvh) {
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    if (bvh.id >= r.bvhs.size()) {
        fprintf(stderr, "Could not free bvh with id: %u\n", bvh.id);  // lib.rs:838-842
        return;
    }
    if (r.bvhs[bvh.id]) {
        DeviceGuard g(r.bvhs[bvh.id]->device);
        drain_device();
        release(r.bvhs[bvh.id]->dev);
        r.bvhs[bvh.id].reset();
    }
}

void free_mbvh(RTMbvh bvh) {
    Registry& r = reg();
    std::unique_lock<std::shared_mutex> lk(r.mu);
    if (bvh.id >= r.mbvhs.size()) {
        fprintf(stderr, "Could not free bvh with id: %u\n", bvh.id);  // lib.rs:845-849
        return;
    }
    if (r.mbvhs[bvh.id]) {
        DeviceGuard g(r.mbvhs[bvh.id]->device);
        drain_device();
        release(r.mbvhs[bvh.id]->dev);
        r.mbvhs[bvh.id].reset();
    }
}

}  // extern "C"
