// Traversal tree: a 4-wide tree made for the lowest-id traversal from any
// binary BVH (no counterpart in the reference, whose MBVH collapse,
// src/mbvh_node.rs:297-411, always merges exactly two binary levels and gives
// a leaf slot its parent's bounds).
//
// Each wide node starts from a binary node's two children and repeatedly
// opens the child with the largest surface area, until it has four slots or
// nothing is left to open.  A subtree of at most `max_leaf` primitives is not
// opened: it becomes one leaf slot with its own (tight) bounds, and its
// triangles are stored contiguously, in leaf order, as one 64-byte record each
// (v0 with the primitive id, e1, e2).  The closest hit under the lowest
// (t, id) rule does not depend on the tree, so tracing this tree gives the
// same hit records as tracing the reference-layout MBVH.
//
// Construction is breadth-first: one kernel per level over a queue of
// (binary node, wide node) pairs held in device memory; the level kernels
// read the queue length on the device, so a fixed number of them can also be
// recorded into a graph.
#include <vector>

#include "block_scan.hpp"
#include "internal.hpp"

namespace rtbvh {

hipError_t scan_exclusive_u32(uint32_t* data, uint32_t count, uint32_t* total, void* scratch,
                              size_t scratch_bytes, hipStream_t stream);

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kLevelGrid = 2048;  // persistent blocks per level kernel
constexpr int kMaxLeafLimit = 8;
constexpr int kWide = 8;  // slots of a compressed node

struct QItem {
    uint32_t b;  // binary node
    uint32_t w;  // wide node it becomes
};

__device__ __forceinline__ void ldn(const DevAabb* p, DevAabb& o) {
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

__device__ __forceinline__ float half_area(const DevAabb& a) {
    const float dx = a.max[0] - a.min[0], dy = a.max[1] - a.min[1], dz = a.max[2] - a.min[2];
    return (dx + dy) * dz + dx * dy;
}

// Primitives below every node, capped at cap + 1: a depth-first count of
// the node's subtree that stops once it exceeds the cap (only "at most
// max_leaf, and how many" matters).  Each node is counted on its own, so no
// thread waits for another and no cross-workgroup ordering is needed.
__global__ void capped_counts_kernel(const DevAabb* nodes, uint32_t count, uint32_t cap,
                                     uint32_t* cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int32_t st[2 * kMaxLeafLimit + 4];
    int sp = 0;
    st[sp++] = int32_t(i);
    uint32_t sum = 0;
    uint32_t steps = 0;
    const uint32_t max_steps = 4 * cap + 8;  // bounds a chain of empty leaves too
    while (sp > 0 && sum <= cap) {
        const int32_t k = st[--sp];
        if (k < 0 || uint32_t(k) >= count || ++steps > max_steps) {
            sum = cap + 1;
            break;
        }
        DevAabb n;
        ldn(nodes + k, n);
        if (n.extra1 >= 0) {
            sum += uint32_t(n.extra1);
        } else if (n.extra2 >= 0) {
            if (sp + 2 > 2 * kMaxLeafLimit + 4) {
                sum = cap + 1;
                break;
            }
            st[sp++] = n.extra2 + 1;
            st[sp++] = n.extra2;
        }
    }
    cnt[i] = sum <= cap ? sum : cap + 1;
}

__device__ __forceinline__ void write_tri(float4* tris, uint32_t pos, const float* verts,
                                          uint32_t prim_count, uint32_t prim,
                                          const uint32_t* scene_off, uint32_t scenes) {
    float4 a = make_float4(0.f, 0.f, 0.f, __uint_as_float(RTBVH_MISS));
    float4 e1 = make_float4(0.f, 0.f, 0.f, 0.f), e2 = e1;
    if (prim < prim_count) {
        const float* v = verts + size_t(prim) * 9;
        uint32_t id = prim;
        if (scene_off) {
            uint32_t lo = 0, hi = scenes;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (scene_off[mid] <= prim)
                    lo = mid;
                else
                    hi = mid;
            }
            id = prim - scene_off[lo];
        }
        a = make_float4(v[0], v[1], v[2], __uint_as_float(id));
        e1 = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], 0.f);
        e2 = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.f);
    }
    tris[size_t(pos) * kTriRecord] = a;
    tris[size_t(pos) * kTriRecord + 1] = e1;
    tris[size_t(pos) * kTriRecord + 2] = e2;
}

struct LevelArgs {
    const DevAabb* nodes;
    uint32_t node_count;
    const uint32_t* cnt;
    const uint32_t* indices;
    uint32_t index_count;
    const float* verts;
    uint32_t prim_count;
    const uint32_t* scene_off;  // batches: scene-local ids
    uint32_t scenes;
    uint32_t max_leaf;
    DevMbvhNode* wnodes;
    DevCw8Node* cnodes;  // width 8: compressed nodes instead of wnodes
    const uint32_t* dpd;  // width 8: slot-distribution decisions per binary node
    uint32_t wcap;
    float4* tris;
    uint32_t* counters;  // [0] wide nodes, [1] triangles, [2] fail
};

// Writes the triangles of the binary subtree at `root` from position pos.
__device__ bool gather_leaf(const LevelArgs& a, int32_t root, uint32_t pos, uint32_t end) {
    int32_t st[2 * kMaxLeafLimit + 2];
    int sp = 0;
    st[sp++] = root;
    while (sp > 0) {
        const int32_t i = st[--sp];
        if (i < 0 || uint32_t(i) >= a.node_count) return false;
        DevAabb n;
        ldn(a.nodes + i, n);
        if (n.extra1 >= 0) {
            for (int32_t j = 0; j < n.extra1; ++j) {
                const uint32_t ip = uint32_t(n.extra2) + uint32_t(j);
                if (pos >= end || n.extra2 < 0 || ip >= a.index_count) return false;
                write_tri(a.tris, pos++, a.verts, a.prim_count, a.indices[ip], a.scene_off,
                          a.scenes);
            }
        } else {
            if (sp + 2 > 2 * kMaxLeafLimit + 2 || n.extra2 < 0) return false;
            st[sp++] = n.extra2 + 1;
            st[sp++] = n.extra2;
        }
    }
    return pos == end;
}

__device__ __forceinline__ bool opens(const LevelArgs& a, const DevAabb& n, uint32_t i) {
    return n.extra1 < 0 && n.extra2 >= 0 && a.cnt[i] > a.max_leaf;
}

// One wide node from queue item (b, w): pick its slots, allocate wide
// children and triangle ranges, queue the wide children.
__device__ void make_node(const LevelArgs& a, uint32_t b, uint32_t w, QItem* qout,
                          uint32_t* qout_count) {
    int32_t slot[4];
    int ns = 0;
    DevAabb bn;
    ldn(a.nodes + b, bn);
    const auto bad = [&](int32_t i) { return i < 0 || uint32_t(i) >= a.node_count; };
    if (bn.extra1 >= 0) {  // a root that is a leaf: one leaf slot
        slot[ns++] = int32_t(b);
    } else {
        slot[ns++] = bn.extra2;
        slot[ns++] = bn.extra2 + 1;
        if (bad(slot[0]) || bad(slot[1])) {  // only in a tree left incomplete
            atomicOr(&a.counters[2], 1u);
            return;
        }
        while (ns < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int k = 0; k < ns; ++k) {
                DevAabb sn;
                ldn(a.nodes + slot[k], sn);
                if (!opens(a, sn, uint32_t(slot[k]))) continue;
                const float ar = half_area(sn);
                if (!(ar <= best_area)) {  // larger, or the first (NaN areas still open)
                    best_area = ar;
                    best = k;
                }
            }
            if (best < 0) break;
            DevAabb en;
            ldn(a.nodes + slot[best], en);
            if (bad(en.extra2) || bad(en.extra2 + 1)) {
                atomicOr(&a.counters[2], 1u);
                return;
            }
            slot[best] = en.extra2;
            slot[ns++] = en.extra2 + 1;
        }
    }
    DevAabb sn[4];
    bool inner[4];
    uint32_t n_inner = 0, n_tris = 0;
    for (int k = 0; k < ns; ++k) {
        ldn(a.nodes + slot[k], sn[k]);
        inner[k] = bn.extra1 < 0 && opens(a, sn[k], uint32_t(slot[k]));
        if (inner[k])
            ++n_inner;
        else
            n_tris += sn[k].extra1 >= 0 ? uint32_t(sn[k].extra1) : a.cnt[slot[k]];
    }
    const uint32_t wbase = n_inner ? atomicAdd(&a.counters[0], n_inner) : 0u;
    const uint32_t tbase = n_tris ? atomicAdd(&a.counters[1], n_tris) : 0u;
    const uint32_t qbase = n_inner ? atomicAdd(qout_count, n_inner) : 0u;
    if (wbase + n_inner > a.wcap || tbase + n_tris > a.index_count) {
        atomicOr(&a.counters[2], 1u);
        return;
    }
    DevMbvhNode m;
    float* mnx = reinterpret_cast<float*>(&m.min_x);
    float* mxx = reinterpret_cast<float*>(&m.max_x);
    float* mny = reinterpret_cast<float*>(&m.min_y);
    float* mxy = reinterpret_cast<float*>(&m.max_y);
    float* mnz = reinterpret_cast<float*>(&m.min_z);
    float* mxz = reinterpret_cast<float*>(&m.max_z);
    int32_t* ch = reinterpret_cast<int32_t*>(&m.children);
    int32_t* ct = reinterpret_cast<int32_t*>(&m.counts);
    uint32_t wi = wbase, ti = tbase, qi = qbase;
    for (int k = 0; k < 4; ++k) {
        if (k >= ns) {
            // an unused slot: a point box beyond any ray's reach, never entered
            mnx[k] = mxx[k] = mny[k] = mxy[k] = mnz[k] = mxz[k] = 3.0e38f;
            ch[k] = -1;
            ct[k] = -1;
            continue;
        }
        mnx[k] = sn[k].min[0];
        mny[k] = sn[k].min[1];
        mnz[k] = sn[k].min[2];
        mxx[k] = sn[k].max[0];
        mxy[k] = sn[k].max[1];
        mxz[k] = sn[k].max[2];
        if (inner[k]) {
            ch[k] = int32_t(wi);
            ct[k] = -1;
            qout[qi].b = uint32_t(slot[k]);
            qout[qi].w = wi;
            ++wi;
            ++qi;
        } else {
            const uint32_t c = sn[k].extra1 >= 0 ? uint32_t(sn[k].extra1) : a.cnt[slot[k]];
            ch[k] = int32_t(ti);
            ct[k] = int32_t(c);
            if (!gather_leaf(a, slot[k], ti, ti + c)) atomicOr(&a.counters[2], 1u);
            ti += c;
        }
    }
    a.wnodes[w] = m;
}

// ---- compressed 8-wide nodes -------------------------------------------

// A slot candidate: a binary interior node, a binary leaf of several
// primitives (opened into them when they fit), or one primitive (an index
// position, with the bounds of its triangle).
struct Cand {
    int32_t id;
    int32_t kind;  // 0 interior, 1 leaf of n > 1 primitives, 2 primitive
    uint32_t n;
    float lo[3], hi[3];
};

// A triangle slot: the box of its three vertices cut to its binary leaf's
// box.  The cut changes nothing where the leaf box holds the whole triangle
// (every tree without spatial splits); after a spatial split the leaf holds
// the part of the triangle inside it, and the cut keeps the slot as tight.
__device__ __forceinline__ bool cand_prim(const LevelArgs& a, uint32_t pos, Cand& c,
                                          const DevAabb& leaf) {
    if (pos >= a.index_count) return false;
    c.id = int32_t(pos);
    c.kind = 2;
    c.n = 1;
    const uint32_t prim = a.indices[pos];
    if (prim < a.prim_count) {
        const float* v = a.verts + size_t(prim) * 9;
        for (int k = 0; k < 3; ++k) {
            c.lo[k] = fminf(fminf(v[k], v[3 + k]), v[6 + k]);
            c.hi[k] = fmaxf(fmaxf(v[k], v[3 + k]), v[6 + k]);
        }
        float lo[3], hi[3];
        bool cut = true;
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmaxf(c.lo[k], leaf.min[k]);
            hi[k] = fminf(c.hi[k], leaf.max[k]);
            cut = cut && lo[k] <= hi[k];
        }
        if (cut)
            for (int k = 0; k < 3; ++k) {
                c.lo[k] = lo[k];
                c.hi[k] = hi[k];
            }
    } else {  // stored as a triangle that no ray hits
        for (int k = 0; k < 3; ++k) c.lo[k] = c.hi[k] = 0.0f;
    }
    return true;
}

// Appends binary node i as a candidate (an empty leaf adds nothing).
__device__ bool cand_binary(const LevelArgs& a, int32_t i, Cand* c, int& ns) {
    if (i < 0 || uint32_t(i) >= a.node_count) return false;
    DevAabb n;
    ldn(a.nodes + i, n);
    if (n.extra1 == 0) return true;
    if (ns >= kWide) return false;
    if (n.extra1 == 1) return cand_prim(a, uint32_t(n.extra2), c[ns++], n);
    Cand& d = c[ns++];
    d.id = i;
    d.kind = n.extra1 > 1 ? 1 : 0;
    d.n = n.extra1 > 1 ? uint32_t(n.extra1) : 0u;
    for (int k = 0; k < 3; ++k) {
        d.lo[k] = n.min[k];
        d.hi[k] = n.max[k];
    }
    return n.extra2 >= 0;
}

// The candidates of one compressed node held in registers: an append
// shifts every entry down one place and writes the new one at the end, so
// after ns appends they sit at [kWide - ns, kWide) in order, and every write
// is by a compile-time index (an access by a run-time index would place
// the arrays in scratch memory).
struct CandRegs {
    int32_t id[kWide], kind[kWide];
    uint32_t n[kWide];
    float lo0[kWide], lo1[kWide], lo2[kWide], hi0[kWide], hi1[kWide], hi2[kWide];

    __device__ __forceinline__ void push(const Cand& t) {
#pragma unroll
        for (int q = 0; q + 1 < kWide; ++q) {
            id[q] = id[q + 1];
            kind[q] = kind[q + 1];
            n[q] = n[q + 1];
            lo0[q] = lo0[q + 1];
            lo1[q] = lo1[q + 1];
            lo2[q] = lo2[q + 1];
            hi0[q] = hi0[q + 1];
            hi1[q] = hi1[q + 1];
            hi2[q] = hi2[q + 1];
        }
        id[kWide - 1] = t.id;
        kind[kWide - 1] = t.kind;
        n[kWide - 1] = t.n;
        lo0[kWide - 1] = t.lo[0];
        lo1[kWide - 1] = t.lo[1];
        lo2[kWide - 1] = t.lo[2];
        hi0[kWide - 1] = t.hi[0];
        hi1[kWide - 1] = t.hi[1];
        hi2[kWide - 1] = t.hi[2];
    }
};

// Binary node i as one candidate t (add: whether it gives one; an empty
// leaf gives none): cand_binary for a caller that stores the candidate.
__device__ __forceinline__ bool cand_binary1(const LevelArgs& a, int32_t i, Cand& t, bool& add) {
    add = false;
    if (i < 0 || uint32_t(i) >= a.node_count) return false;
    DevAabb n;
    ldn(a.nodes + i, n);
    if (n.extra1 == 0) return true;
    add = true;
    if (n.extra1 == 1) return cand_prim(a, uint32_t(n.extra2), t, n);
    t.id = i;
    t.kind = n.extra1 > 1 ? 1 : 0;
    t.n = n.extra1 > 1 ? uint32_t(n.extra1) : 0u;
    for (int k = 0; k < 3; ++k) {
        t.lo[k] = n.min[k];
        t.hi[k] = n.max[k];
    }
    return n.extra2 >= 0;
}

// Grid of one axis: origin p at least one step below lo, step 2^(e-127) a
// power of two with (hi - lo) / step <= 252 and step >= 2^-22 * max(|lo|,
// |hi|) (so the spare step outlasts float rounding in the traversal).
__device__ bool grid_axis(float lo, float hi, float& p, uint32_t& ebyte, double& step) {
    const double ext = double(hi) - double(lo);
    const double mag = fmax(fabs(double(lo)), fabs(double(hi)));
    double need = fmax(ext / 252.0, mag * 0x1p-22);
    need = fmax(need, 0x1p-100);
    int k = 0;
    const double m = frexp(need, &k);  // need = m 2^k, m in [0.5, 1)
    if (m == 0.5) --k;                 // need is itself a power of two
    if (k + 127 < 1 || k + 127 > 254) return false;
    ebyte = uint32_t(k + 127);
    step = ldexp(1.0, k);
    const double pd = double(lo) - step;
    float pf = float(pd);
    if (double(pf) > pd) pf = nextafterf(pf, -3.4028235e38f);
    p = pf;
    return true;
}

// inv_step: 1 / step, exact (a power of two), so the product equals the
// quotient
__device__ __forceinline__ uint32_t quant_lo(float v, float p, double inv_step) {
    const double q = floor((double(v) - double(p)) * inv_step) - 1.0;
    return q <= 0.0 ? 0u : (q >= 255.0 ? 255u : uint32_t(q));
}

__device__ __forceinline__ uint32_t quant_hi(float v, float p, double inv_step) {
    const double q = ceil((double(v) - double(p)) * inv_step) + 1.0;
    return q <= 0.0 ? 0u : (q >= 255.0 ? 255u : uint32_t(q));
}

// Decision nibble j (1..8) of binary node k: 0 the node takes one slot (an
// interior wide node, or a triangle for a one-primitive leaf), kDissolve a
// leaf of several primitives spreads them over their own slots, otherwise k
// in 1..j-1 gives its left child k slots and its right child j - k.
constexpr uint32_t kDissolve = 15u;

__device__ __forceinline__ uint32_t dp_decision(const uint32_t* dpd, uint32_t k, uint32_t j) {
    return (dpd[k] >> (4u * (j - 1u))) & 15u;
}

// One compressed node from queue item (b, w).  Its slots come from the
// SAH-optimal distribution of b's subtree over eight slots (dp_level_kernel):
// b's children get k and 8 - k slots, and each part either stays one slot
// or spreads further.  The slots are then placed so that slot s lies towards
// the octant s (bit k of s set: the + side of axis k) of the node's centre,
// greedily by alignment; a ray whose direction has signs O visits keys
// s ^ O in increasing order, the near side first.  Called by all 64 lanes
// of a wave (`live` false: no item): the wide-node, triangle and queue
// counters take one atomic per wave, the lanes' offsets from wave scans.
__device__ void make_node8(const LevelArgs& a, bool live, uint32_t b, uint32_t w, QItem* qout,
                           uint32_t* qout_count) {
    CandRegs cr;
    int ns = 0;
    const auto append = [&](const Cand& t) {
        cr.push(t);
        ++ns;
    };
    bool ok = live;
    if (ok && (b >= a.node_count || w >= a.wcap)) {
        atomicOr(&a.counters[2], 1u);
        ok = false;
    }
    DevAabb bn;
    if (ok) ldn(a.nodes + b, bn);
    if (!ok) {
        // nothing to gather
    } else if (bn.extra1 >= 0) {  // a root that is a leaf, or a leaf of several primitives
        if (bn.extra1 > kWide || bn.extra2 < 0) ok = false;
        for (int32_t j = 0; ok && j < bn.extra1; ++j)
        {
            Cand t;
            ok = cand_prim(a, uint32_t(bn.extra2 + j), t, bn);
            append(t);
        }
    } else {
        // (binary node, slots) pairs still to place; at most 8 are pending
        uint32_t st_node[kWide], st_slots[kWide];
        int sp = 0;
        const uint32_t k8 = dp_decision(a.dpd, b, 8u);
        if (bn.extra2 < 0 || k8 == 0u || k8 >= 8u) {
            ok = false;
        } else {
            st_node[sp] = uint32_t(bn.extra2) + 1u;
            st_slots[sp++] = 8u - k8;
            st_node[sp] = uint32_t(bn.extra2);
            st_slots[sp++] = k8;
        }
        while (ok && sp > 0) {
            --sp;
            const uint32_t n = st_node[sp], j = st_slots[sp];
            if (n >= a.node_count || j < 1u || j > 8u) {
                ok = false;
                break;
            }
            DevAabb nd;
            ldn(a.nodes + n, nd);
            const uint32_t d = dp_decision(a.dpd, n, j);
            if (nd.extra1 >= 0 && d == kDissolve) {  // its primitives, one slot each
                if (nd.extra2 < 0 || uint32_t(nd.extra1) > j) {
                    ok = false;
                    break;
                }
                for (int32_t q = 0; ok && q < nd.extra1; ++q) {
                    Cand t;
                    ok = ns < kWide && cand_prim(a, uint32_t(nd.extra2 + q), t, nd);
                    if (ok) append(t);
                }
            } else if (nd.extra1 >= 0 || d == 0u) {  // one slot
                Cand t;
                bool add = false;
                ok = cand_binary1(a, int32_t(n), t, add);
                if (ok && add) {
                    if (ns >= kWide)
                        ok = false;
                    else
                        append(t);
                }
            } else {  // spread over j slots: k to the left child, j - k to the right
                if (nd.extra2 < 0 || d >= j || sp + 2 > kWide) {
                    ok = false;
                    break;
                }
                st_node[sp] = uint32_t(nd.extra2) + 1u;
                st_slots[sp++] = j - d;
                st_node[sp] = uint32_t(nd.extra2);
                st_slots[sp++] = d;
            }
        }
    }
    if (live && (!ok || ns == 0)) atomicOr(&a.counters[2], 1u);
    ok = ok && ns > 0;
    // bounds of the node and the centre of every candidate
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    if (ok) {
        // the candidates sit at [kWide - ns, kWide); fminf / fmaxf skip NaN as
        // a fold from the first candidate would
        for (int k = 0; k < 3; ++k) {
            lo[k] = __builtin_inff();
            hi[k] = -__builtin_inff();
        }
#pragma unroll
        for (int j = 0; j < kWide; ++j)
            if (j >= kWide - ns) {
                lo[0] = fminf(lo[0], cr.lo0[j]);
                lo[1] = fminf(lo[1], cr.lo1[j]);
                lo[2] = fminf(lo[2], cr.lo2[j]);
                hi[0] = fmaxf(hi[0], cr.hi0[j]);
                hi[1] = fmaxf(hi[1], cr.hi1[j]);
                hi[2] = fmaxf(hi[2], cr.hi2[j]);
            }
        for (int k = 0; k < 3; ++k)
            if (!(isfinite(lo[k]) && isfinite(hi[k]) && lo[k] <= hi[k])) ok = false;
        if (!ok) atomicOr(&a.counters[2], 2u);
    }
    // octant placement: each round takes the first (candidate, free slot)
    // pair of the highest alignment score.  Every loop over candidates and
    // slots is unrolled, so nothing is indexed at run time.
    float off0[kWide], off1[kWide], off2[kWide];
#pragma unroll
    for (int k = 0; k < kWide; ++k) {
        const bool v = ok && k >= kWide - ns;
        off0[k] = v ? (cr.lo0[k] + cr.hi0[k]) - (lo[0] + hi[0]) : 0.0f;
        off1[k] = v ? (cr.lo1[k] + cr.hi1[k]) - (lo[1] + hi[1]) : 0.0f;
        off2[k] = v ? (cr.lo2[k] + cr.hi2[k]) - (lo[2] + hi[2]) : 0.0f;
    }
    uint32_t cand_slot[kWide];  // the slot of each candidate
#pragma unroll
    for (int k = 0; k < kWide; ++k) cand_slot[k] = 0;
    uint32_t placed = 0, taken = 0;  // candidates placed, slots taken
    for (int it = 0; ok && it < ns; ++it) {
        float best = -3.4e38f;
        int bk = -1, bs = -1;
#pragma unroll
        for (int k = 0; k < kWide; ++k) {
            if (k < kWide - ns || (placed & (1u << k))) continue;
#pragma unroll
            for (int s = 0; s < kWide; ++s) {
                if (taken & (1u << s)) continue;
                const float sc = ((s & 1) ? off0[k] : -off0[k]) +
                                 ((s & 2) ? off1[k] : -off1[k]) +
                                 ((s & 4) ? off2[k] : -off2[k]);
                if (bk < 0 || sc > best) {
                    best = sc;
                    bk = k;
                    bs = s;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kWide; ++k)
            if (k == bk) cand_slot[k] = uint32_t(bs);
        placed |= 1u << bk;
        taken |= 1u << bs;
    }
    // grid, children and triangles
    float p[3] = {0.0f, 0.0f, 0.0f};
    uint32_t eb[3] = {0u, 0u, 0u};
    double step[3] = {1.0, 1.0, 1.0};
    for (int k = 0; ok && k < 3; ++k)
        if (!grid_axis(lo[k], hi[k], p[k], eb[k], step[k])) {
            atomicOr(&a.counters[2], 2u);
            ok = false;
        }
    double inv_step[3];
    for (int k = 0; k < 3; ++k) inv_step[k] = 1.0 / step[k];
    uint32_t n_inner = 0, n_tris = 0;
#pragma unroll
    for (int j = 0; j < kWide; ++j) {
        if (!(ok && j >= kWide - ns)) continue;
        if (cr.kind[j] == 2)
            ++n_tris;
        else
            ++n_inner;
    }
    // offsets: wave scans of the lanes' counts, one atomic per counter and
    // wave (the last lane's inclusive sums are the wave's totals)
    const uint32_t inc_i = wave_inclusive(n_inner), inc_t = wave_inclusive(n_tris);
    uint32_t wb = 0, tb = 0, qb = 0;
    if ((threadIdx.x & 63u) == 63u) {
        wb = inc_i ? atomicAdd(&a.counters[0], inc_i) : 0u;
        tb = inc_t ? atomicAdd(&a.counters[1], inc_t) : 0u;
        qb = inc_i ? atomicAdd(qout_count, inc_i) : 0u;
    }
    wb = __shfl(wb, 63, 64);
    tb = __shfl(tb, 63, 64);
    qb = __shfl(qb, 63, 64);
    if (!ok) return;
    const uint32_t wbase = wb + inc_i - n_inner, tbase = tb + inc_t - n_tris;
    const uint32_t qbase = qb + inc_i - n_inner;
    if (wbase + n_inner > a.wcap || tbase + n_tris > a.index_count) {
        atomicOr(&a.counters[2], 1u);
        return;
    }
    DevCw8Node m;
    m.px = p[0];
    m.py = p[1];
    m.pz = p[2];
    // slot masks first: the triangles and the inner children are stored in
    // slot order, each at its rank among the slots of its kind
    uint32_t imask = 0, lmask = 0;
#pragma unroll
    for (int k = 0; k < kWide; ++k) {
        if (k < kWide - ns) continue;
        if (cr.kind[k] == 2)
            lmask |= 1u << cand_slot[k];
        else
            imask |= 1u << cand_slot[k];
    }
    uint32_t qlo[3][2] = {{0u, 0u}, {0u, 0u}, {0u, 0u}};
    uint32_t qhi[3][2] = {{0u, 0u}, {0u, 0u}, {0u, 0u}};
#pragma unroll
    for (int s = 0; s < kWide; ++s)  // empty slots: an inverted box, never entered
        if (!((imask | lmask) & (1u << s)))
            for (int ax = 0; ax < 3; ++ax) qlo[ax][s >> 2] |= 255u << ((s & 3) * 8);
#pragma unroll
    for (int k = 0; k < kWide; ++k) {
        if (k < kWide - ns) continue;
        const uint32_t sl = cand_slot[k], sh = (sl & 3u) * 8u, below = (1u << sl) - 1u;
        const float clo[3] = {cr.lo0[k], cr.lo1[k], cr.lo2[k]};
        const float chi[3] = {cr.hi0[k], cr.hi1[k], cr.hi2[k]};
        for (int ax = 0; ax < 3; ++ax) {
            const uint32_t ql = quant_lo(clo[ax], p[ax], inv_step[ax]) << sh;
            const uint32_t qh = quant_hi(chi[ax], p[ax], inv_step[ax]) << sh;
            if (sl >= 4u) {
                qlo[ax][1] |= ql;
                qhi[ax][1] |= qh;
            } else {
                qlo[ax][0] |= ql;
                qhi[ax][0] |= qh;
            }
        }
        if (cr.kind[k] == 2) {
            write_tri(a.tris, tbase + uint32_t(__popc(lmask & below)), a.verts, a.prim_count,
                      a.indices[cr.id[k]], a.scene_off, a.scenes);
        } else {
            const uint32_t r = uint32_t(__popc(imask & below));
            qout[qbase + r].b = uint32_t(cr.id[k]);
            qout[qbase + r].w = wbase + r;
        }
    }
    m.emask = eb[0] | (eb[1] << 8) | (eb[2] << 16) | (imask << 24);
    m.child_base = wbase;
    m.tri_base = tbase;
    m.lmask = lmask;
    m.pad = 0;
    for (int word = 0; word < 2; ++word) {
        m.qlo_x[word] = qlo[0][word];
        m.qlo_y[word] = qlo[1][word];
        m.qlo_z[word] = qlo[2][word];
        m.qhi_x[word] = qhi[0][word];
        m.qhi_y[word] = qhi[1][word];
        m.qhi_z[word] = qhi[2][word];
    }
    a.cnodes[w] = m;
}

// ---- SAH-optimal distribution of binary subtrees over compressed slots ---
//
// cost(n, j): expected cost (surface area of the compressed nodes a ray may
// enter) of binary node n's subtree when it may occupy up to j slots of its
// parent's node:
//   leaf of c primitives: j >= c: 0, the c triangles each take a slot of the
//   parent; j < c (c > 1): one slot holding a node of its own, its area;
//   interior: min(one slot holding a node of its own: area + dist(n, 8);
//   dist(n, j)), dist(n, j) = min over k of cost(left, k) +
//   cost(right, j - k).
// Every triangle ends in a slot of its own, entered through its own box, so
// the triangle tests cost the same for every distribution and only the nodes
// count.  Computed level by level from the deepest binary level up, so that
// every child is final before its parent.
constexpr uint32_t kMaxBinLevels = 8192;

// The binary tree breadth-first: level 0 holds the roots; level L is
// order[off[L] .. off[L] + cnt[L]).  Cleared before each build.
__global__ void bin_init_kernel(uint32_t roots, uint32_t* order, uint32_t* off, uint32_t* cnt,
                                uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kMaxBinLevels + 2;
         i += gridDim.x * blockDim.x) {
        cnt[i] = i == 0 ? roots : 0u;
        off[i] = 0u;
        if (i < roots && i < count) order[i] = i;
    }
    for (uint32_t i = kMaxBinLevels + 2 + blockIdx.x * blockDim.x + threadIdx.x; i < roots;
         i += gridDim.x * blockDim.x)
        if (i < count) order[i] = i;
}

// Children of level `level` appended as level + 1.  Whole blocks iterate
// together, so that the next level's entries are claimed with one atomic per
// block and tile: same-address atomics serialize, and one per wave was the
// kernel's bottleneck on wide levels.
constexpr int kBinBlock = 1024;
constexpr int kBinWaves = kBinBlock / 64;

__global__ void __launch_bounds__(kBinBlock)
    bin_level_kernel(const DevAabb* nodes, uint32_t count, uint32_t* order, uint32_t* off,
                     uint32_t* cnt, uint32_t level, uint32_t* fail) {
    __shared__ uint32_t s_cnt[kBinWaves];
    __shared__ uint32_t s_base[kBinWaves];
    const uint32_t begin = off[level];
    const uint32_t n = cnt[level];
    const uint32_t end = begin + n;
    if (blockIdx.x == 0 && threadIdx.x == 0) off[level + 1] = end;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t t = blockIdx.x * kBinBlock; t < n; t += gridDim.x * kBinBlock) {
        const uint32_t idx = t + threadIdx.x;
        const uint32_t i = begin + idx;
        bool interior = false;
        DevAabb nd;
        if (idx < n) {
            if (i >= count) {
                atomicOr(fail, 1u);
            } else {
                ldn(nodes + order[i], nd);
                interior = nd.extra1 < 0;
            }
        }
        const uint64_t mask = __ballot(interior);
        if (lane == 0) s_cnt[wave] = 2u * uint32_t(__popcll(mask));
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t total = 0;
            for (int k = 0; k < kBinWaves; ++k) {
                s_base[k] = total;
                total += s_cnt[k];
            }
            const uint32_t base = total ? atomicAdd(&cnt[level + 1], total) : 0u;
            for (int k = 0; k < kBinWaves; ++k) s_base[k] += base;
        }
        __syncthreads();
        if (!interior) continue;
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
            uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
        const uint32_t pos = end + s_base[wave] + 2u * below;
        if (nd.extra2 < 0 || uint32_t(nd.extra2) + 1u >= count || pos + 1u >= count) {
            atomicOr(fail, 1u);
            continue;
        }
        order[pos] = uint32_t(nd.extra2);
        order[pos + 1] = uint32_t(nd.extra2) + 1u;
    }
}

__global__ void __launch_bounds__(kBlock)
    dp_level_kernel(LevelArgs a, const uint32_t* order, const uint32_t* off, const uint32_t* cnt,
                    uint32_t level, float* cost, uint32_t* dec) {
    const uint32_t begin = off[level];
    const uint32_t end = begin + cnt[level];
    for (uint32_t i = begin + blockIdx.x * kBlock + threadIdx.x; i < end;
         i += gridDim.x * kBlock) {
        const uint32_t k = order[i];
        if (k >= a.node_count) {
            atomicOr(&a.counters[2], 1u);
            continue;
        }
        DevAabb nd;
        ldn(a.nodes + k, nd);
        float c[9];
        uint32_t d = 0;
        if (nd.extra1 >= 0) {
            const uint32_t np = uint32_t(nd.extra1);
            if (np > uint32_t(kWide)) atomicOr(&a.counters[2], 1u);
            const float one = np <= 1u ? 0.0f : half_area(nd);
            for (uint32_t j = 1; j <= 8u; ++j) {
                c[j] = j >= np ? 0.0f : one;
                if (j >= np && np > 1u) d |= kDissolve << (4u * (j - 1u));
            }
        } else {
            const uint32_t l = uint32_t(nd.extra2), r = l + 1u;
            if (nd.extra2 < 0 || r >= a.node_count) {
                atomicOr(&a.counters[2], 1u);
                continue;
            }
            float cl[9], cr[9];
            for (uint32_t j = 1; j <= 8u; ++j) {
                cl[j] = cost[size_t(l) * 8 + j - 1];
                cr[j] = cost[size_t(r) * 8 + j - 1];
            }
            float dist[9];
            uint32_t dk[9];
            for (uint32_t j = 2; j <= 8u; ++j) {
                float best = 3.0e38f;
                uint32_t bk = 1;
                for (uint32_t q = 1; q < j; ++q) {
                    const float v = cl[q] + cr[j - q];
                    if (v < best) {
                        best = v;
                        bk = q;
                    }
                }
                dist[j] = best;
                dk[j] = bk;
            }
            c[1] = half_area(nd) + dist[8];
            for (uint32_t j = 2; j <= 8u; ++j) {
                const bool spread = j == 8u || dist[j] < c[1];
                c[j] = spread ? dist[j] : c[1];
                if (spread) d |= dk[j] << (4u * (j - 1u));
            }
        }
        for (uint32_t j = 1; j <= 8u; ++j) cost[size_t(k) * 8 + j - 1] = c[j];
        dec[k] = d;
    }
}

// clear: the queue count two levels ahead (the level before's input, read
// by nobody now) is zeroed for the level after this one to count into.
__global__ void __launch_bounds__(kBlock)
    level8_kernel(LevelArgs a, const QItem* qin, const uint32_t* qin_count, QItem* qout,
                  uint32_t* qout_count, uint32_t* clear, uint32_t level) {
    const uint32_t q = *qin_count;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *clear = 0;
        if (q > 0) atomicMax(&a.counters[4], level + 1);  // levels with work
    }
    // whole blocks iterate together: make_node8 needs every lane of a wave
    for (uint32_t base = blockIdx.x * kBlock; base < q; base += gridDim.x * kBlock) {
        const uint32_t i = base + threadIdx.x;
        QItem it;
        it.b = it.w = 0;
        if (i < q) it = qin[i];
        make_node8(a, i < q, it.b, it.w, qout, qout_count);
    }
}

__global__ void __launch_bounds__(kBlock)
    level_kernel(LevelArgs a, const QItem* qin, const uint32_t* qin_count, QItem* qout,
                 uint32_t* qout_count, uint32_t* clear, uint32_t level) {
    const uint32_t q = *qin_count;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *clear = 0;
        if (q > 0) atomicMax(&a.counters[4], level + 1);  // levels with work
    }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < q; i += gridDim.x * kBlock) {
        const QItem it = qin[i];
        if (it.b >= a.node_count || it.w >= a.wcap) {
            atomicOr(&a.counters[2], 1u);
            continue;
        }
        make_node(a, it.b, it.w, qout, qout_count);
    }
}

// Roots [0, roots) become wide nodes [0, roots); counters start after them.
// q_count: level 0's input count, q_next: level 0's output count.
__global__ void roots_kernel(uint32_t roots, QItem* q, uint32_t* q_count, uint32_t* q_next,
                             uint32_t* counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *q_count = roots;
        *q_next = 0;
        counters[0] = roots;
        counters[1] = 0;
        counters[2] = 0;
        counters[4] = 0;
    }
    if (i < roots) {
        q[i].b = i;
        q[i].w = i;
    }
}

// Queue counts rotate over three words (level L counts its input in
// qword(L) and its output in qword(L + 1), and clears qword(L + 2)):
// words[0], words[1] and words[5]; words[2..4] are the counters (wide
// nodes, triangles, failure) and words[6] the number of levels with work.
__host__ __device__ constexpr uint32_t qword(uint32_t level) {
    return level % 3 == 2 ? 5u : level % 3;
}

// fail when the last level still queued work (the level budget ran out) or
// the triangles do not add up to every index entry.
__global__ void finish_kernel(const uint32_t* last_q, uint32_t* counters, uint32_t index_count,
                              uint32_t* fail_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (*last_q != 0 || counters[1] != index_count) counters[2] |= 1u;
    if (fail_out) *fail_out |= counters[2];
}

}  // namespace

struct TraceTreePlan {
    uint32_t count = 0, roots = 1, n = 0, wcap = 0, max_leaf = 1, levels = 0, width = 4;
    uint32_t bin_levels = 0;    // width 8: binary-tree depth seen by the eager build
    uint32_t* cnt = nullptr;
    QItem* q[2] = {nullptr, nullptr};
    uint32_t* words = nullptr;  // [0, 1, 5] queue counts (qword), [2..4] counters, [6] depth
    // width 8: breadth-first binary order, its level offsets and counts, and
    // the distribution costs (8 per node) and decisions (8 nibbles per node)
    uint32_t* bord = nullptr;
    uint32_t* blev = nullptr;
    float* dpc = nullptr;
    uint32_t* dpd = nullptr;
    void* ws = nullptr;  // one allocation holding all of the above
    Workspace wsh;       // ws from the workspace cache
    hipStream_t last_stream = nullptr;  // of the last build
};

void trace_tree_plan_destroy(TraceTreePlan* t) {
    if (!t) return;
    if (t->wsh.ptr)
        workspace_release(&t->wsh, t->last_stream, true);  // after the build's work
    else if (t->ws)
        (void)hipFree(t->ws);
    delete t;
}

namespace {

LevelArgs level_args(const TraceTreePlan* t, const DeviceBvh& tree, const float* verts,
                     uint32_t prim_count, const uint32_t* scene_off, uint32_t scenes,
                     DeviceMbvh* out) {
    LevelArgs a;
    a.nodes = tree.nodes;
    a.node_count = tree.node_count;
    a.cnt = t->cnt;
    a.indices = tree.indices;
    a.index_count = tree.prim_count;
    a.verts = verts;
    a.prim_count = prim_count;
    a.scene_off = scene_off;
    a.scenes = scenes;
    a.max_leaf = t->max_leaf;
    a.wnodes = out->nodes;
    a.cnodes = out->cnodes;
    a.dpd = t->dpd;
    a.wcap = t->wcap;
    a.tris = out->tris;
    a.counters = t->words + 2;
    return a;
}

// Capped subtree counts and the root queue.
hipError_t enqueue_setup(const TraceTreePlan* t, const DeviceBvh& tree, hipStream_t stream) {
    const uint32_t count = t->count;
    const uint32_t g = (count + kBlock - 1) / kBlock;
    if (t->width == 4)  // the compressed tree never collapses subtrees into leaves
        hipLaunchKernelGGL(capped_counts_kernel, dim3(g), dim3(kBlock), 0, stream, tree.nodes,
                           count, t->max_leaf, t->cnt);
    else
        hipLaunchKernelGGL(bin_init_kernel, dim3(64), dim3(kBlock), 0, stream, t->roots, t->bord,
                           t->blev, t->blev + kMaxBinLevels + 2, count);
    const uint32_t gr = (t->roots + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(roots_kernel, dim3(gr), dim3(kBlock), 0, stream, t->roots, t->q[0],
                       t->words + qword(0), t->words + qword(1), t->words + 2);
    return hipGetLastError();
}

constexpr uint32_t kBinGrid = 2048;   // persistent blocks per distribution-level kernel
constexpr uint32_t kBinLevelGrid = 512;  // blocks of kBinBlock per breadth-first level

void enqueue_bin_level(const TraceTreePlan* t, const DeviceBvh& tree, uint32_t level,
                       hipStream_t stream) {
    hipLaunchKernelGGL(bin_level_kernel, dim3(kBinLevelGrid), dim3(kBinBlock), 0, stream,
                       tree.nodes, t->count, t->bord, t->blev, t->blev + kMaxBinLevels + 2, level,
                       t->words + 4);
}

void enqueue_dp_levels(const TraceTreePlan* t, const LevelArgs& a, uint32_t levels,
                       hipStream_t stream) {
    for (uint32_t l = levels; l-- > 0;)
        hipLaunchKernelGGL(dp_level_kernel, dim3(kBinGrid), dim3(kBlock), 0, stream, a, t->bord,
                           t->blev, t->blev + kMaxBinLevels + 2, l, t->dpc, t->dpd);
}

__global__ void bin_depth_check_kernel(const uint32_t* cnt, uint32_t levels, uint32_t* fail) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && cnt[levels] != 0) *fail |= 1u;
}

void enqueue_level(const TraceTreePlan* t, const LevelArgs& a, uint32_t level, uint32_t grid,
                   hipStream_t stream) {
    const int in = level & 1, outq = in ^ 1;  // queue buffers alternate
    uint32_t* w = t->words;
    if (t->width == 8)
        hipLaunchKernelGGL(level8_kernel, dim3(grid), dim3(kBlock), 0, stream, a, t->q[in],
                           w + qword(level), t->q[outq], w + qword(level + 1),
                           w + qword(level + 2), level);
    else
        hipLaunchKernelGGL(level_kernel, dim3(grid), dim3(kBlock), 0, stream, a, t->q[in],
                           w + qword(level), t->q[outq], w + qword(level + 1),
                           w + qword(level + 2), level);
}

}  // namespace

ResultCode trace_tree_plan_create(const DeviceBvh& tree, uint32_t max_leaf, uint32_t width,
                                  TraceTreePlan** out_plan, DeviceMbvh* out, bool cached) {
    if (width != 4 && width != 8) {
        set_error("trace tree width must be 4 or 8");
        return InvalidArgument;
    }
    if (max_leaf < 1 || max_leaf > uint32_t(kMaxLeafLimit) || (width == 8 && max_leaf != 1)) {
        set_error("max_leaf must be 1..%d (1 for the 8-wide tree)", kMaxLeafLimit);
        return InvalidArgument;
    }
    TraceTreePlan* t = new TraceTreePlan();
    t->count = tree.node_count;
    t->roots = tree.root_count;
    t->n = tree.prim_count;
    t->max_leaf = max_leaf;
    t->width = width;
    t->wcap = tree.prim_count + tree.root_count;
    // one workspace allocation (allocations and frees cost far more than
    // the kernels of a small tree), carved at 256-byte offsets
    size_t bytes = 0;
    const auto carve = [&bytes](size_t n) {
        const size_t at = bytes;
        bytes += (n + 255) & ~size_t(255);
        return at;
    };
    const size_t o_cnt = carve(size_t(t->count) * 4);
    const size_t o_q0 = carve(size_t(t->wcap) * sizeof(QItem));
    const size_t o_q1 = carve(size_t(t->wcap) * sizeof(QItem));
    const size_t o_words = carve(8 * 4);
    size_t o_bord = 0, o_blev = 0, o_dpc = 0, o_dpd = 0;
    if (width == 8) {
        o_bord = carve(size_t(t->count) * 4);
        o_blev = carve(size_t(kMaxBinLevels + 2) * 2 * 4);
        o_dpc = carve(size_t(t->count) * 8 * sizeof(float));
        o_dpd = carve(size_t(t->count) * 4);
    }
    hipError_t e =
        cached ? workspace_acquire(kWsTree, bytes, &t->wsh) : hipMalloc(&t->ws, bytes);
    if (e != hipSuccess) {
        t->ws = nullptr;
        trace_tree_plan_destroy(t);
        return hip_fail(e, "trace tree workspace");
    }
    if (cached) t->ws = t->wsh.ptr;
    char* base = static_cast<char*>(t->ws);
    t->cnt = reinterpret_cast<uint32_t*>(base + o_cnt);
    t->q[0] = reinterpret_cast<QItem*>(base + o_q0);
    t->q[1] = reinterpret_cast<QItem*>(base + o_q1);
    t->words = reinterpret_cast<uint32_t*>(base + o_words);
    if (width == 8) {
        t->bord = reinterpret_cast<uint32_t*>(base + o_bord);
        t->blev = reinterpret_cast<uint32_t*>(base + o_blev);
        t->dpc = reinterpret_cast<float*>(base + o_dpc);
        t->dpd = reinterpret_cast<uint32_t*>(base + o_dpd);
    }
    DeviceMbvh res;
    if ((width == 4 &&
         (e = hipMalloc(&res.nodes, size_t(t->wcap) * sizeof(DevMbvhNode))) != hipSuccess) ||
        (width == 8 &&
         (e = hipMalloc(&res.cnodes, size_t(t->wcap) * sizeof(DevCw8Node))) != hipSuccess) ||
        (e = hipMalloc(&res.tris, size_t(t->n) * kTriRecord * sizeof(float4))) != hipSuccess) {
        if (res.nodes) (void)hipFree(res.nodes);
        if (res.cnodes) (void)hipFree(res.cnodes);
        trace_tree_plan_destroy(t);
        return hip_fail(e, "trace tree");
    }
    res.indices = nullptr;
    res.node_count = t->wcap;
    res.prim_count = t->n;
    res.root_count = t->roots;
    *out = res;
    *out_plan = t;
    return Ok;
}

// Eager: levels until the queue is empty (host check every few levels).
ResultCode trace_tree_build(TraceTreePlan* t, const DeviceBvh& tree, const float* verts,
                            uint32_t prim_count, const uint32_t* scene_off, uint32_t scenes,
                            DeviceMbvh* out, hipStream_t stream) {
    t->last_stream = stream;
    hipError_t e = enqueue_setup(t, tree, stream);
    if (e != hipSuccess) return hip_fail(e, "trace tree setup");
    const LevelArgs a = level_args(t, tree, verts, prim_count, scene_off, scenes, out);
    uint32_t first_wide = 4;  // wide levels in the first batch
    if (t->width == 8) {  // binary levels until one is empty, then the distribution
        // first batch: 3 lg n + 4 levels (a PLOC tree of 1M is about 61
        // deep), then 24 at a time; levels past the last return at once
        uint32_t lg = 0;
        while (lg < 31 && (1u << lg) < t->count) ++lg;
        uint32_t chunk = 3u * lg + 4u;
        uint32_t bl = 0;
        std::vector<uint32_t> counts;
        for (;; chunk = 24u) {
            if (bl + chunk > kMaxBinLevels) {
                set_error("binary tree deeper than %u levels", kMaxBinLevels);
                return Error;
            }
            for (uint32_t k = 0; k < chunk; ++k, ++bl) enqueue_bin_level(t, tree, bl, stream);
            counts.resize(bl + 1);
            if ((e = hipMemcpyAsync(counts.data(), t->blev + kMaxBinLevels + 2, (bl + 1) * 4,
                                    hipMemcpyDeviceToHost, stream)) != hipSuccess ||
                (e = hipStreamSynchronize(stream)) != hipSuccess)
                return hip_fail(e, "binary levels");
            if (counts[bl] == 0) break;
        }
        uint32_t depth = bl;
        while (depth > 0 && counts[depth - 1] == 0) --depth;
        t->bin_levels = depth;
        enqueue_dp_levels(t, a, depth, stream);
        // a compressed node spans about three binary levels
        first_wide = depth / 3 + 2 > 4 ? depth / 3 + 2 : 4;
    }
    uint32_t level = 0;
    uint32_t host_words[8];
    for (uint32_t chunk = first_wide;; chunk = 4) {
        for (uint32_t k = 0; k < chunk; ++k, ++level) enqueue_level(t, a, level, kLevelGrid, stream);
        if ((e = hipMemcpyAsync(host_words, t->words, sizeof(host_words), hipMemcpyDeviceToHost,
                                stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
            return hip_fail(e, "trace tree level");
        if (host_words[qword(level)] == 0) break;  // nothing queued for the next level
        if (level > 4 * 4096) {
            set_error("trace tree deeper than expected");
            return Error;
        }
    }
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(1), 0, stream, t->words + qword(level),
                       t->words + 2, t->n, nullptr);
    if ((e = hipMemcpyAsync(host_words, t->words, sizeof(host_words), hipMemcpyDeviceToHost,
                            stream)) != hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
        return hip_fail(e, "trace tree finish");
    if (host_words[4] != 0) {
        set_error(host_words[4] & 2u ? "trace tree construction failed (non-finite bounds)"
                                     : "trace tree construction failed (inconsistent binary tree?)");
        return Error;
    }
    t->levels = host_words[6];  // the levels with work (the batches enqueue more)
    out->node_count = host_words[2];
    return Ok;
}

// Capturable: a fixed number of levels (the eager build's depth plus
// headroom), failure reported in *fail.
hipError_t trace_tree_enqueue(const TraceTreePlan* t, const DeviceBvh& tree, const float* verts,
                              uint32_t prim_count, const uint32_t* scene_off, uint32_t scenes,
                              DeviceMbvh* out, uint32_t levels, uint32_t* fail,
                              hipStream_t stream) {
    hipError_t e = enqueue_setup(t, tree, stream);
    if (e != hipSuccess) return e;
    const LevelArgs a = level_args(t, tree, verts, prim_count, scene_off, scenes, out);
    if (t->width == 8) {  // the eager build's binary depth plus headroom
        uint32_t bl = t->bin_levels + (t->bin_levels / 4 > 8 ? t->bin_levels / 4 : 8);
        if (bl > kMaxBinLevels) bl = kMaxBinLevels;
        for (uint32_t l = 0; l < bl; ++l) enqueue_bin_level(t, tree, l, stream);
        hipLaunchKernelGGL(bin_depth_check_kernel, dim3(1), dim3(1), 0, stream,
                           t->blev + kMaxBinLevels + 2, bl, t->words + 4);
        enqueue_dp_levels(t, a, bl, stream);
    }
    for (uint32_t level = 0; level < levels; ++level) enqueue_level(t, a, level, kLevelGrid, stream);
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(1), 0, stream, t->words + qword(levels),
                       t->words + 2, t->n, fail);
    return hipGetLastError();
}

uint32_t trace_tree_levels(const TraceTreePlan* t) { return t->levels; }

}  // namespace rtbvh
