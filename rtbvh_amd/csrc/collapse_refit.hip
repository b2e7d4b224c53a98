// GPU MBVH collapse (Mbvh::construct, src/bvh.rs:381-404 with
// MbvhNode::merge_nodes, src/mbvh_node.rs:297-411) and GPU refit
// (Bvh::refit, src/bvh.rs:176-205).
//
// merge_nodes makes one MBVH node for the root and one for every interior
// binary node whose parent is an interior child of an MBVH node — i.e. for
// the interior nodes at even depth.  Its four slots come from the node's two
// children: a leaf child fills one slot and keeps the parent's bounds
// (mbvh_node.rs:305,319-321, reproduced as is), an interior child contributes
// its two children as slots with their own bounds.  The reference numbers
// MBVH nodes in depth-first pre-order; here they are numbered in binary-index
// order (one scan), which describes the same tree.
#include "internal.hpp"

namespace rtbvh {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxDepth = 1u << 20;

__device__ __forceinline__ void ld(const DevAabb* p, DevAabb& o) {
    const float4* q = reinterpret_cast<const float4*>(p);
    const float4 a = q[0], b = q[1];
    o.min[0] = a.x;
    o.min[1] = a.y;
    o.min[2] = a.z;
    o.extra1 = __float_as_int(a.w);
    o.max[0] = b.x;
    o.max[1] = b.y;
    o.max[2] = b.z;
    o.extra2 = __float_as_int(b.w);
}

__device__ __forceinline__ void st(DevAabb* p, const DevAabb& n) {
    float4* q = reinterpret_cast<float4*>(p);
    q[0] = make_float4(n.min[0], n.min[1], n.min[2], __int_as_float(n.extra1));
    q[1] = make_float4(n.max[0], n.max[1], n.max[2], __int_as_float(n.extra2));
}

__global__ void fill_i32_kernel(int32_t* p, uint32_t n, int32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// parent[child] for every interior node (count < 0, left_first >= 0).
__global__ void parent_kernel(const DevAabb* nodes, uint32_t count, int32_t* parent) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    DevAabb n;
    ld(nodes + i, n);
    if (n.extra1 < 0 && n.extra2 >= 0 && uint32_t(n.extra2) + 1 < count) {
        parent[n.extra2] = int32_t(i);
        parent[n.extra2 + 1] = int32_t(i);
    }
}

// flag[i] = 1 for the roots (nodes [0, roots)) and for interior nodes at
// even depth below them.
__global__ void mflag_kernel(const DevAabb* nodes, uint32_t count, uint32_t roots,
                             const int32_t* parent, uint32_t* flag, uint32_t* status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t f = 0;
    if (i < roots) {
        f = 1;
    } else {
        DevAabb n;
        ld(nodes + i, n);
        if (n.extra1 < 0 && n.extra2 >= 0) {
            uint32_t depth = 0;
            int32_t p = int32_t(i);
            while (p >= int32_t(roots) && depth < kMaxDepth) {
                p = parent[p];
                ++depth;
            }
            if (p >= 0 && p < int32_t(roots))
                f = (depth & 1u) ? 0u : 1u;
            else if (depth >= kMaxDepth)
                atomicOr(status, 2u);  // cycle or absurd depth
        }
    }
    flag[i] = f;
}

__device__ __forceinline__ void set_slot_bounds(DevMbvhNode& m, int s, const DevAabb& b) {
    float* mnx = reinterpret_cast<float*>(&m.min_x);
    float* mxx = reinterpret_cast<float*>(&m.max_x);
    float* mny = reinterpret_cast<float*>(&m.min_y);
    float* mxy = reinterpret_cast<float*>(&m.max_y);
    float* mnz = reinterpret_cast<float*>(&m.min_z);
    float* mxz = reinterpret_cast<float*>(&m.max_z);
    mnx[s] = b.min[0];
    mny[s] = b.min[1];
    mnz[s] = b.min[2];
    mxx[s] = b.max[0];
    mxy[s] = b.max[1];
    mxz[s] = b.max[2];
}

__global__ void mfill_kernel(const DevAabb* nodes, uint32_t count, const uint32_t* flag,
                             const uint32_t* mindex, DevMbvhNode* mnodes, uint32_t mcap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || flag[i] == 0 || mindex[i] >= mcap) return;
    DevAabb cur;
    ld(nodes + i, cur);
    DevMbvhNode m;
    int32_t* ch = reinterpret_cast<int32_t*>(&m.children);
    int32_t* ct = reinterpret_cast<int32_t*>(&m.counts);
    for (int s = 0; s < 4; ++s) {
        set_slot_bounds(m, s, cur);  // mbvh_node.rs:304-306
        ch[s] = -1;
        ct[s] = -1;
    }
    const int32_t lf = cur.extra2;
    if (lf >= 0) {
        for (int side = 0; side < 2; ++side) {
            const uint32_t c = uint32_t(lf) + uint32_t(side);
            if (c >= count) continue;
            DevAabb cn;
            ld(nodes + c, cn);
            if (cn.extra2 < 0) continue;
            const int s0 = 2 * side;
            if (cn.extra1 >= 0) {
                // leaf child: one slot, parent's bounds kept
                ch[s0] = cn.extra2;
                ct[s0] = cn.extra1;
                continue;
            }
            for (int g = 0; g < 2; ++g) {
                const uint32_t gi = uint32_t(cn.extra2) + uint32_t(g);
                if (gi >= count) continue;
                DevAabb gn;
                ld(nodes + gi, gn);
                set_slot_bounds(m, s0 + g, gn);
                if (gn.extra1 >= 0) {
                    if (gn.extra2 >= 0) {
                        ch[s0 + g] = gn.extra2;
                        ct[s0 + g] = gn.extra1;
                    }
                } else {
                    ch[s0 + g] = int32_t(mindex[gi]);
                }
            }
        }
    }
    mnodes[mindex[i]] = m;
}

// ------------------------------------------------------------------ refit --
// Leaves: grow from Aabb::new over their primitives' aabbs in order, then
// +1e-4; interior: grow by left then right child, then +1e-4; a node whose
// left_first is negative gets the empty box.  count and left_first are kept.
// Bottom-up by breadth-first level, deepest level first (the reference walks
// the node array in reverse, src/bvh.rs:176-205; children always come after
// their parent, so both orders finish every child before its parent).
__device__ __forceinline__ DevAabb refit_leaf(const DevAabb& n, const uint32_t* indices,
                                              uint32_t index_count, const DevAabb* aabbs) {
    DevAabb a;
    for (int k = 0; k < 3; ++k) {
        a.min[k] = 1e34f;
        a.max[k] = -1e34f;
    }
    if (n.extra2 >= 0) {
        for (int32_t j = 0; j < n.extra1; ++j) {
            const uint32_t pos = uint32_t(n.extra2) + uint32_t(j);
            if (pos >= index_count) break;
            DevAabb p;
            ld(aabbs + indices[pos], p);
            for (int k = 0; k < 3; ++k) {
                a.min[k] = rs_min(a.min[k], p.min[k]);
                a.max[k] = rs_max(a.max[k], p.max[k]);
            }
        }
        for (int k = 0; k < 3; ++k) {
            a.min[k] = a.min[k] - 0.0001f;
            a.max[k] = a.max[k] + 0.0001f;
        }
    }
    a.extra1 = n.extra1;
    a.extra2 = n.extra2;
    return a;
}

// Breadth-first order, one level per launch: the children of
// order[begin, end) are appended from *next_end on.
__global__ void bfs_level_kernel(const DevAabb* nodes, uint32_t count, uint32_t* order,
                                 uint32_t begin, uint32_t end, uint32_t* next_end) {
    const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    DevAabb n;
    ld(nodes + order[i], n);
    if (n.extra1 < 0 && n.extra2 >= 0 && uint32_t(n.extra2) + 1 < count) {
        const uint32_t pos = atomicAdd(next_end, 2u);
        if (pos + 1 < count) {
            order[pos] = uint32_t(n.extra2);
            order[pos + 1] = uint32_t(n.extra2) + 1;
        }
    }
}

__global__ void iota_roots_kernel(uint32_t* order, uint32_t roots, uint32_t* end) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < roots) order[i] = i;
    if (i == 0) *end = roots;
}

// One depth level of the refit, deepest first: every child is final before
// its parent's launch, so no thread waits for another.
__global__ void refit_level_kernel(DevAabb* nodes, const uint32_t* order, uint32_t begin,
                                   uint32_t end, const uint32_t* indices, uint32_t index_count,
                                   const DevAabb* aabbs) {
    const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const uint32_t k = order[i];
    DevAabb n;
    ld(nodes + k, n);
    if (!(n.extra1 < 0 && n.extra2 >= 0)) {
        st(nodes + k, refit_leaf(n, indices, index_count, aabbs));
        return;
    }
    DevAabb l, r;
    ld(nodes + n.extra2, l);
    ld(nodes + n.extra2 + 1, r);
    DevAabb a;
    for (int c = 0; c < 3; ++c) {
        a.min[c] = rs_min(rs_min(1e34f, l.min[c]), r.min[c]) - 0.0001f;
        a.max[c] = rs_max(rs_max(-1e34f, l.max[c]), r.max[c]) + 0.0001f;
    }
    a.extra1 = n.extra1;
    a.extra2 = n.extra2;
    st(nodes + k, a);
}

__global__ void copy_u32_kernel(const uint32_t* src, uint32_t* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

// Exclusive scan used by the collapse (defined in scan.hip).
hipError_t scan_exclusive_u32(uint32_t* data, uint32_t count, uint32_t* total, void* scratch,
                              size_t scratch_bytes, hipStream_t stream);
size_t scan_scratch_bytes(uint32_t count);

hipError_t launch_parent_links(const DevAabb* nodes, uint32_t count, int32_t* parents,
                               hipStream_t stream) {
    const uint32_t g = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(parent_kernel, dim3(g), dim3(kBlock), 0, stream, nodes, count, parents);
    return hipGetLastError();
}

ResultCode ensure_parents(DeviceBvh& bvh, hipStream_t stream) {
    if (bvh.parents) return Ok;
    hipError_t e = hipMalloc(&bvh.parents, size_t(bvh.node_count) * sizeof(int32_t));
    if (e != hipSuccess) {
        bvh.parents = nullptr;
        return hip_fail(e, "hipMalloc(parents)");
    }
    const uint32_t g = (bvh.node_count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(g), dim3(kBlock), 0, stream, bvh.parents,
                       bvh.node_count, -1);
    hipLaunchKernelGGL(parent_kernel, dim3(g), dim3(kBlock), 0, stream, bvh.nodes, bvh.node_count,
                       bvh.parents);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "parent pointers");
    return Ok;
}

ResultCode gpu_build_mbvh(DeviceBvh& bvh, hipStream_t stream, DeviceMbvh* out) {
    DeviceMbvh res;
    res.prim_count = bvh.prim_count;
    res.root_count = bvh.root_count;
    if (bvh.node_count == 0) {
        *out = res;
        return Ok;
    }
    ResultCode rc = ensure_parents(bvh, stream);
    if (rc != Ok) return rc;
    const uint32_t count = bvh.node_count;
    const uint32_t g = (count + kBlock - 1) / kBlock;
    hipError_t e;
    uint32_t* flag = nullptr;
    uint32_t* mindex = nullptr;
    uint32_t* total = nullptr;
    void* scratch = nullptr;
    const size_t sbytes = scan_scratch_bytes(count);
    auto cleanup = [&]() {
        if (flag) (void)hipFree(flag);
        if (mindex) (void)hipFree(mindex);
        if (total) (void)hipHostFree(total);
        if (scratch) (void)hipFree(scratch);
    };
    if ((e = hipMalloc(&flag, size_t(count) * 4)) != hipSuccess ||
        (e = hipMalloc(&mindex, size_t(count) * 4)) != hipSuccess ||
        (e = hipHostMalloc(&total, 4)) != hipSuccess ||
        (sbytes && (e = hipMalloc(&scratch, sbytes)) != hipSuccess)) {
        cleanup();
        return hip_fail(e, "collapse workspace");
    }
    uint32_t* status = device_status_ptr();
    hipLaunchKernelGGL(mflag_kernel, dim3(g), dim3(kBlock), 0, stream, bvh.nodes, count,
                       bvh.root_count, bvh.parents, flag, status);
    hipLaunchKernelGGL(copy_u32_kernel, dim3(g), dim3(kBlock), 0, stream, flag, mindex, count);
    if ((e = scan_exclusive_u32(mindex, count, total, scratch, sbytes, stream)) != hipSuccess) {
        cleanup();
        return hip_fail(e, "collapse scan");
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) {
        cleanup();
        return hip_fail(e, "collapse scan");
    }
    const uint32_t mcount = *total;
    if ((e = hipMalloc(&res.nodes, size_t(mcount) * sizeof(DevMbvhNode))) != hipSuccess ||
        (e = hipMalloc(&res.indices, size_t(bvh.prim_count) * 4 + 4)) != hipSuccess) {
        if (res.nodes) (void)hipFree(res.nodes);
        cleanup();
        return hip_fail(e, "hipMalloc(mbvh)");
    }
    res.node_count = mcount;
    hipLaunchKernelGGL(mfill_kernel, dim3(g), dim3(kBlock), 0, stream, bvh.nodes, count, flag,
                       mindex, res.nodes, mcount);
    const uint32_t gi = (bvh.prim_count + kBlock - 1) / kBlock;
    if (bvh.prim_count)
        hipLaunchKernelGGL(copy_u32_kernel, dim3(gi), dim3(kBlock), 0, stream, bvh.indices,
                           res.indices, bvh.prim_count);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    cleanup();
    if (e != hipSuccess) {
        (void)hipFree(res.nodes);
        (void)hipFree(res.indices);
        return hip_fail(e, "collapse");
    }
    *out = res;
    return Ok;
}

struct CollapsePlanImpl {
    uint32_t count = 0, mcap = 0, roots = 1;
    int32_t* parents = nullptr;
    uint32_t* flag = nullptr;
    uint32_t* mindex = nullptr;
    void* scratch = nullptr;
    size_t sbytes = 0;
    DevMbvhNode* mnodes = nullptr;
};

void collapse_plan_destroy(CollapsePlanImpl* c) {
    if (!c) return;
    if (c->parents) (void)hipFree(c->parents);
    if (c->flag) (void)hipFree(c->flag);
    if (c->mindex) (void)hipFree(c->mindex);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->mnodes) (void)hipFree(c->mnodes);
    delete c;
}

// A tree of n primitives and r roots has n - r interior nodes, which bounds
// the MBVH node count (the roots and the interior nodes at even depth).
ResultCode collapse_plan_create(const DeviceBvh& tree, CollapsePlanImpl** out, DeviceMbvh* view) {
    CollapsePlanImpl* c = new CollapsePlanImpl();
    c->count = tree.node_count;
    c->roots = tree.root_count;
    c->mcap = tree.prim_count > tree.root_count ? tree.prim_count - tree.root_count : 1u;
    c->sbytes = scan_scratch_bytes(c->count);
    hipError_t e;
    if ((e = hipMalloc(&c->parents, size_t(c->count) * 4)) != hipSuccess ||
        (e = hipMalloc(&c->flag, size_t(c->count) * 4)) != hipSuccess ||
        (e = hipMalloc(&c->mindex, size_t(c->count) * 4)) != hipSuccess ||
        (c->sbytes && (e = hipMalloc(&c->scratch, c->sbytes)) != hipSuccess) ||
        (e = hipMalloc(&c->mnodes, size_t(c->mcap) * sizeof(DevMbvhNode))) != hipSuccess) {
        collapse_plan_destroy(c);
        return hip_fail(e, "collapse plan");
    }
    view->nodes = c->mnodes;
    view->indices = tree.indices;  // shared with the tree
    view->tris = nullptr;
    view->node_count = c->mcap;
    view->prim_count = tree.prim_count;
    view->root_count = tree.root_count;
    *out = c;
    return Ok;
}

hipError_t collapse_plan_enqueue(const CollapsePlanImpl* c, const DeviceBvh& tree,
                                 hipStream_t stream) {
    const uint32_t count = c->count;
    const uint32_t g = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(g), dim3(kBlock), 0, stream, c->parents, count, -1);
    hipLaunchKernelGGL(parent_kernel, dim3(g), dim3(kBlock), 0, stream, tree.nodes, count,
                       c->parents);
    hipLaunchKernelGGL(mflag_kernel, dim3(g), dim3(kBlock), 0, stream, tree.nodes, count, c->roots,
                       c->parents, c->flag, device_status_ptr());
    hipLaunchKernelGGL(copy_u32_kernel, dim3(g), dim3(kBlock), 0, stream, c->flag, c->mindex,
                       count);
    hipError_t e = scan_exclusive_u32(c->mindex, count, nullptr, c->scratch, c->sbytes, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mfill_kernel, dim3(g), dim3(kBlock), 0, stream, tree.nodes, count, c->flag,
                       c->mindex, c->mnodes, c->mcap);
    return hipGetLastError();
}

// Breadth-first levels of the tree (topology only, so computed once).
ResultCode ensure_order(DeviceBvh& bvh, hipStream_t stream) {
    if (bvh.order) return Ok;
    const uint32_t count = bvh.node_count;
    uint32_t* word = nullptr;
    hipError_t e = hipMalloc(&bvh.order, size_t(count) * 4);
    if (e == hipSuccess) e = hipMalloc(&word, 4);
    if (e != hipSuccess) {
        if (bvh.order) (void)hipFree(bvh.order);
        bvh.order = nullptr;
        return hip_fail(e, "hipMalloc(order)");
    }
    std::vector<uint32_t> off{0};
    const uint32_t roots = bvh.root_count < count ? bvh.root_count : count;
    hipLaunchKernelGGL(iota_roots_kernel, dim3((roots + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       stream, bvh.order, roots, word);
    uint32_t begin = 0, end = roots;
    off.push_back(end);
    while (end > begin && off.size() <= count + 1) {
        const uint32_t g = (end - begin + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(bfs_level_kernel, dim3(g), dim3(kBlock), 0, stream, bvh.nodes, count,
                           bvh.order, begin, end, word);
        uint32_t next = 0;
        if ((e = hipMemcpyAsync(&next, word, 4, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
            break;
        if (next > count) next = count;
        begin = end;
        end = next;
        if (end > begin) off.push_back(end);
    }
    (void)hipFree(word);
    if (e != hipSuccess) {
        (void)hipFree(bvh.order);
        bvh.order = nullptr;
        return hip_fail(e, "breadth-first order");
    }
    bvh.level_off = off;
    return Ok;
}

ResultCode gpu_refit(DeviceBvh& bvh, const RTAabb* aabbs, hipStream_t stream) {
    if (bvh.node_count == 0) return Ok;
    ResultCode rc = ensure_order(bvh, stream);
    if (rc != Ok) return rc;
    for (size_t L = bvh.level_off.size() - 1; L-- > 0;) {
        const uint32_t b = bvh.level_off[L], e = bvh.level_off[L + 1];
        hipLaunchKernelGGL(refit_level_kernel, dim3((e - b + kBlock - 1) / kBlock), dim3(kBlock),
                           0, stream, bvh.nodes, bvh.order, b, e, bvh.indices, bvh.prim_count,
                           reinterpret_cast<const DevAabb*>(aabbs));
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "refit");
    return Ok;
}

}  // namespace rtbvh
