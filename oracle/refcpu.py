"""ctypes front end of the CPU oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg use this module as the checker.  The product package
``rtbvh_amd`` never imports it.

Each wrapper names the reference function it restates (see refcpu.cpp for the
file:line citations).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefcpu.so")

# rtbvh_ffi RTAabb / RTBvhNode (rtbvh_ffi/src/lib.rs:144-151)
AABB = np.dtype([("min", "<f4", 3), ("count", "<i4"), ("max", "<f4", 3), ("left_first", "<i4")])
# rtbvh_ffi RTMbvhNode (rtbvh_ffi/src/lib.rs:197-208)
MBVH = np.dtype([("min_x", "<f4", 4), ("max_x", "<f4", 4), ("min_y", "<f4", 4),
                 ("max_y", "<f4", 4), ("min_z", "<f4", 4), ("max_z", "<f4", 4),
                 ("children", "<i4", 4), ("counts", "<i4", 4)])
RAY = np.dtype([("o", "<f4", 3), ("t_min", "<f4"), ("d", "<f4", 3), ("t_max", "<f4")])
HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4")])
assert AABB.itemsize == 32 and MBVH.itemsize == 128 and RAY.itemsize == 32 and HIT.itemsize == 16

MODE_LOWEST_ID = 0
MODE_VISIT_ORDER = 1

_lib = None


def build() -> str:
    """Compile the oracle with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _declare(L):
    vp, sz, u32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int
    sig = {
        "ref_morton_split": (u32, [u32]),
        "ref_world_bounds": (None, [vp, sz, vp]),
        "ref_morton_codes": (None, [vp, vp, sz, sz, vp]),
        "ref_sorted_indices": (None, [vp, sz, vp]),
        "ref_prefix_sum_u32": (u32, [vp, sz, vp]),
        "ref_half_area": (ctypes.c_float, [vp]),
        "ref_triangle_aabbs_centers": (None, [vp, sz, vp, vp]),
        "ref_build_locb": (sz, [vp, vp, sz, sz, vp, vp]),
        "ref_build_locb_radius": (sz, [vp, vp, sz, sz, sz, vp, vp, vp]),
        "ref_build_binned_sah": (sz, [vp, vp, sz, sz, sz, vp, vp]),
        "ref_build_binned_sah_contained": (sz, [vp, vp, sz, sz, sz, vp, vp]),
        "ref_build_mbvh": (sz, [vp, sz, vp]),
        "ref_refit": (None, [vp, sz, vp, sz, vp]),
        "ref_mt": (i32, [vp, vp, vp, ctypes.c_float, ctypes.c_float, vp, vp, vp]),
        "ref_trace_mbvh": (None, [vp, sz, vp, sz, vp, sz, vp, sz, vp, i32, i32, vp]),
        "ref_trace_bvh": (None, [vp, sz, vp, sz, vp, sz, vp, sz, vp, i32, i32, vp]),
        "ref_trace_brute": (None, [vp, sz, vp, sz, vp, i32]),
        "ref_occluded_mbvh": (None, [vp, sz, vp, sz, vp, sz, vp, sz, vp, i32]),
        "ref_occluded_bvh": (None, [vp, sz, vp, sz, vp, sz, vp, sz, vp, i32]),
        "ref_occluded_brute": (None, [vp, sz, vp, sz, vp, i32]),
        "ref_check_bvh": (i32, [vp, sz, vp, sz, sz, vp, vp]),
        "ref_check_mbvh": (i32, [vp, sz, vp, sz, sz, vp]),
        "ref_sah_cost": (ctypes.c_double, [vp, sz]),
        "ref_compare_bvh": (i32, [vp, sz, vp, vp, sz, vp, i32]),
        "ref_compare_mbvh": (i32, [vp, sz, vp, vp, sz, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------- geometry ----
def morton_split(v: int) -> int:
    """morton_split, src/morton.rs:10-25."""
    return int(lib().ref_morton_split(v))


def world_bounds(aabbs):
    """Aabb::union_of_list (+1e-4), src/aabb.rs:125-131."""
    aabbs = _c(aabbs, AABB)
    out = np.zeros(1, AABB)
    lib().ref_world_bounds(_p(aabbs), len(aabbs), _p(out))
    return out[0]


def morton_codes(aabbs, centers):
    """MortonEncoder::new + encode over Aabb::union_of_list(aabbs), morton.rs:36-61."""
    aabbs = _c(aabbs, AABB)
    centers = _c(centers, np.float32).reshape(len(aabbs), -1) if len(aabbs) else np.zeros((0, 3), np.float32)
    w = np.zeros(1, AABB)
    lib().ref_world_bounds(_p(aabbs), len(aabbs), _p(w))
    codes = np.zeros(len(aabbs), np.uint32)
    lib().ref_morton_codes(_p(w), _p(centers), centers.shape[1], len(aabbs), _p(codes))
    return codes


def sorted_indices(codes):
    """Stable sort of 0..n by code, morton.rs:63-102."""
    codes = _c(codes, np.uint32)
    idx = np.zeros(len(codes), np.uint32)
    lib().ref_sorted_indices(_p(codes), len(codes), _p(idx))
    return idx


def prefix_sum(values, count=None):
    """prefix_sum, src/utils.rs:42-57.  Returns (output, total)."""
    values = _c(values, np.uint32)
    count = len(values) if count is None else count
    out = np.zeros(max(1, len(values)), np.uint32)
    total = lib().ref_prefix_sum_u32(_p(values), count, _p(out))
    return out[:count], int(total)


def half_area(aabb):
    a = _c(np.array([aabb], dtype=AABB) if not isinstance(aabb, np.ndarray) else aabb, AABB)
    return float(lib().ref_half_area(_p(a)))


def triangle_aabbs_centers(verts):
    """shared/src/lib.rs:27-38: per-triangle aabb and (v0+v1+v2)*(1/3)."""
    verts = _c(verts, np.float32).reshape(-1, 9)
    n = len(verts)
    aabbs = np.zeros(n, AABB)
    centers = np.zeros((n, 3), np.float32)
    lib().ref_triangle_aabbs_centers(_p(verts), n, _p(aabbs), _p(centers))
    return aabbs, centers


# ------------------------------------------------------------- builders ----
def build_locb(aabbs, centers, radius=14, return_iterations=False):
    """LocallyOrderedClusteringBuilder::build, src/builders/locb.rs:249-329."""
    aabbs = _c(aabbs, AABB)
    centers = _c(centers, np.float32).reshape(len(aabbs), -1) if len(aabbs) else np.zeros((0, 3), np.float32)
    n = len(aabbs)
    nodes = np.zeros(max(1, 2 * n - 1), AABB)
    idx = np.zeros(max(1, n), np.uint32)
    iters = np.zeros(1, np.uint32)
    cnt = lib().ref_build_locb_radius(_p(aabbs), _p(centers), centers.shape[1], n, radius,
                                      _p(nodes), _p(idx), _p(iters))
    res = (nodes[:cnt].copy(), idx[:n].copy())
    return res + (int(iters[0]),) if return_iterations else res


def build_binned_sah(aabbs, centers, prims_per_leaf=1, contained=False):
    """BinnedSahBuilder::build, src/builders/binned_sah.rs:346-400.
    contained=True fixes the reference's two box defects (stale left-box
    count, binned_sah.rs:188-235; double leaf pad, binned_sah.rs:133,135)."""
    aabbs = _c(aabbs, AABB)
    centers = _c(centers, np.float32).reshape(len(aabbs), -1) if len(aabbs) else np.zeros((0, 3), np.float32)
    n = len(aabbs)
    nodes = np.zeros(max(1, 2 * n - 1), AABB)
    idx = np.zeros(max(1, n), np.uint32)
    fn = lib().ref_build_binned_sah_contained if contained else lib().ref_build_binned_sah
    cnt = fn(_p(aabbs), _p(centers), centers.shape[1], n, prims_per_leaf, _p(nodes), _p(idx))
    return nodes[:cnt].copy(), idx[:n].copy()


def build_mbvh(nodes):
    """Mbvh::construct, src/bvh.rs:381-404."""
    nodes = _c(nodes, AABB)
    m = np.zeros(max(1, len(nodes)), MBVH)
    cnt = lib().ref_build_mbvh(_p(nodes), len(nodes), _p(m))
    return m[:cnt].copy()


def refit(nodes, indices, aabbs):
    """Bvh::refit, src/bvh.rs:176-205 (count/left_first kept).  Returns a copy."""
    nodes = _c(nodes, AABB).copy()
    indices = _c(indices, np.uint32)
    aabbs = _c(aabbs, AABB)
    lib().ref_refit(_p(nodes), len(nodes), _p(indices), len(indices), _p(aabbs))
    return nodes


# ------------------------------------------------------------ traversal ----
def mt(v9, o, d, t_min=1e-4, t_cur=1e34):
    v9 = _c(v9, np.float32)
    o = _c(o, np.float32)
    d = _c(d, np.float32)
    out = np.zeros(3, np.float32)
    ok = lib().ref_mt(_p(v9), _p(o), _p(d), t_min, t_cur, _p(out[0:1]), _p(out[1:2]),
                      _p(out[2:3]))
    return (bool(ok), float(out[0]), float(out[1]), float(out[2]))


def trace_mbvh(m, indices, verts, rays, mode=MODE_LOWEST_ID, threads=1, return_stats=False):
    """MbvhIterator closest hit (iter.rs:308-354) with Moller-Trumbore."""
    m = _c(m, MBVH)
    indices = _c(indices, np.uint32)
    verts = _c(verts, np.float32).reshape(-1)
    rays = _c(rays, RAY)
    hits = np.zeros(len(rays), HIT)
    stats = np.zeros(2, np.uint64)
    lib().ref_trace_mbvh(_p(m), len(m), _p(indices), len(indices), _p(verts), len(verts) // 9,
                         _p(rays), len(rays), _p(hits), mode, threads, _p(stats))
    return (hits, stats) if return_stats else hits


def trace_bvh(nodes, indices, verts, rays, mode=MODE_LOWEST_ID, threads=1, return_stats=False):
    """BvhIterator closest hit (iter.rs:88-127) with Moller-Trumbore."""
    nodes = _c(nodes, AABB)
    indices = _c(indices, np.uint32)
    verts = _c(verts, np.float32).reshape(-1)
    rays = _c(rays, RAY)
    hits = np.zeros(len(rays), HIT)
    stats = np.zeros(2, np.uint64)
    lib().ref_trace_bvh(_p(nodes), len(nodes), _p(indices), len(indices), _p(verts),
                        len(verts) // 9, _p(rays), len(rays), _p(hits), mode, threads,
                        _p(stats))
    return (hits, stats) if return_stats else hits


def trace_brute(verts, rays, threads=1):
    verts = _c(verts, np.float32).reshape(-1)
    rays = _c(rays, RAY)
    hits = np.zeros(len(rays), HIT)
    lib().ref_trace_brute(_p(verts), len(verts) // 9, _p(rays), len(rays), _p(hits), threads)
    return hits


def occluded_mbvh(m, indices, verts, rays, threads=1):
    m = _c(m, MBVH)
    indices = _c(indices, np.uint32)
    verts = _c(verts, np.float32).reshape(-1)
    rays = _c(rays, RAY)
    out = np.zeros(len(rays), np.uint8)
    lib().ref_occluded_mbvh(_p(m), len(m), _p(indices), len(indices), _p(verts),
                            len(verts) // 9, _p(rays), len(rays), _p(out), threads)
    return out


def occluded_bvh(nodes, indices, verts, rays, threads=1):
    nodes = _c(nodes, AABB)
    indices = _c(indices, np.uint32)
    verts = _c(verts, np.float32).reshape(-1)
    rays = _c(rays, RAY)
    out = np.zeros(len(rays), np.uint8)
    lib().ref_occluded_bvh(_p(nodes), len(nodes), _p(indices), len(indices), _p(verts),
                           len(verts) // 9, _p(rays), len(rays), _p(out), threads)
    return out


def occluded_brute(verts, rays, threads=1):
    verts = _c(verts, np.float32).reshape(-1)
    rays = _c(rays, RAY)
    out = np.zeros(len(rays), np.uint8)
    lib().ref_occluded_brute(_p(verts), len(verts) // 9, _p(rays), len(rays), _p(out), threads)
    return out


# ------------------------------------------------------------- checkers ----
def check_bvh(nodes, indices, prim_count):
    """Returns (code, containment_violations, max_depth); code 0 = valid tree
    referencing every primitive exactly once."""
    nodes = _c(nodes, AABB)
    indices = _c(indices, np.uint32)
    viol = np.zeros(1, np.uint64)
    depth = np.zeros(1, np.uint32)
    code = lib().ref_check_bvh(_p(nodes), len(nodes), _p(indices), len(indices), prim_count,
                               _p(viol), _p(depth))
    return int(code), int(viol[0]), int(depth[0])


def check_mbvh(m, indices, prim_count):
    m = _c(m, MBVH)
    indices = _c(indices, np.uint32)
    depth = np.zeros(1, np.uint32)
    code = lib().ref_check_mbvh(_p(m), len(m), _p(indices), len(indices), prim_count, _p(depth))
    return int(code), int(depth[0])


def sah_cost(nodes):
    nodes = _c(nodes, AABB)
    return float(lib().ref_sah_cost(_p(nodes), len(nodes)))


def compare_bvh(a, ia, b, ib, as_sets=True):
    a, b = _c(a, AABB), _c(b, AABB)
    ia, ib = _c(ia, np.uint32), _c(ib, np.uint32)
    return int(lib().ref_compare_bvh(_p(a), len(a), _p(ia), _p(b), len(b), _p(ib),
                                     1 if as_sets else 0))


def compare_mbvh(a, ia, b, ib):
    a, b = _c(a, MBVH), _c(b, MBVH)
    ia, ib = _c(ia, np.uint32), _c(ib, np.uint32)
    return int(lib().ref_compare_mbvh(_p(a), len(a), _p(ia), _p(b), len(b), _p(ib)))
