// Allocation-free PLOC build (rtbvh_gpu_create_bvh_noalloc): the workspace of
// the build is `scratch`, a static device array of RTBVH_NOALLOC_SCRATCH_BYTES
// that every GPU the library runs on holds until the process ends.  Host code
// uses the array only through its device address, looked up once per device
// with hipGetSymbolAddress.
#include <mutex>

#include "internal.hpp"

static __device__ __align__(256) unsigned char scratch[RTBVH_NOALLOC_SCRATCH_BYTES];

namespace rtbvh {
namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device (looked up on first use
// per device and kept).
ResultCode scratch_address(void** out) {
    static std::mutex mu;
    static void* ptrs[kMaxDevices] = {nullptr};
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d out of range", dev);
        return NoDevice;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (ptrs[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        ptrs[dev] = p;
    }
    *out = ptrs[dev];
    return Ok;
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

}  // namespace
}  // namespace rtbvh

using namespace rtbvh;

extern "C" {

ResultCode rtbvh_gpu_create_bvh_noalloc(const RTAabb* aabbs, size_t prim_count,
                                        const float* centers, size_t center_stride,
                                        RTBvhNode* nodes, uint32_t* indices, void* stream) {
    if (centers == nullptr || nodes == nullptr || indices == nullptr) {
        set_error("centers, nodes and indices must not be null");
        return Error;
    }
    if (prim_count == 0) {
        set_error("no primitives");
        return NoPrimitives;
    }
    if (center_stride != 12 && center_stride != 16) {
        set_error("center stride must be 12 or 16 bytes (got %zu)", center_stride);
        return InvalidArgument;
    }
    if (prim_count > RTBVH_NOALLOC_MAX_PRIMS) {
        set_error("the allocation-free build takes at most %u primitives (got %zu)",
                  RTBVH_NOALLOC_MAX_PRIMS, prim_count);
        return InvalidArgument;
    }
    if (!on_device(centers) || !on_device(nodes) || !on_device(indices) ||
        (aabbs && !on_device(aabbs))) {
        set_error("aabbs, centers, nodes and indices must be device memory");
        return InvalidArgument;
    }
    const size_t need = ploc_workspace_bytes(prim_count, aabbs == nullptr);
    if (need > RTBVH_NOALLOC_SCRATCH_BYTES) {
        set_error("the build of %zu primitives needs %zu bytes of workspace, more than the %u "
                  "bytes held", prim_count, need, RTBVH_NOALLOC_SCRATCH_BYTES);
        return InvalidArgument;
    }
    void* ws = nullptr;
    ResultCode rc = scratch_address(&ws);
    if (rc != Ok) return rc;
    return gpu_build_ploc_into(aabbs, centers, center_stride / 4, uint32_t(prim_count), ws,
                               RTBVH_NOALLOC_SCRATCH_BYTES, nodes, indices,
                               static_cast<hipStream_t>(stream));
}

}  // extern "C"
