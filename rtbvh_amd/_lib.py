"""ctypes binding of librtbvh_amd.so (declarations of include/rtbvh.h).

The library is built in-tree by rtbvh_amd/Makefile.  Loading fails loudly when
it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtbvh_amd.so")

# ResultCode, include/rtbvh.h (rtbvh_ffi/src/lib.rs:17-25 + appended codes)
RESULT_NAMES = {
    0: "Ok", 1: "Error", 2: "NoPrimitives", 3: "InequalAabbsAndPrimitives", 4: "Nan",
    5: "DeviceError", 6: "OutOfMemory", 7: "InvalidArgument", 8: "StackOverflow",
    9: "NoDevice",
}
OK, ERROR, NO_PRIMITIVES, INEQUAL, NAN_CODE = 0, 1, 2, 3, 4
DEVICE_ERROR, OUT_OF_MEMORY, INVALID_ARGUMENT, STACK_OVERFLOW, NO_DEVICE = 5, 6, 7, 8, 9

LOCALLY_ORDERED_CLUSTERED = 0
BINNED_SAH = 1
TRACE_REFERENCE_ORDER = 0x1
TRACE_ORDER_RAYS = 0x2
MISS = 0xFFFFFFFF
MAX_STACK = 256
MAX_BATCH_SCENES = 4096
NOALLOC_MAX_PRIMS = 262144
NOALLOC_SCRATCH_BYTES = 40 << 20


class RTAabb(ctypes.Structure):
    _fields_ = [("min", ctypes.c_float * 3), ("count", ctypes.c_int32),
                ("max", ctypes.c_float * 3), ("left_first", ctypes.c_int32)]


class RTBvh(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
                ("nodes", ctypes.c_void_p), ("index_count", ctypes.c_uint32),
                ("indices", ctypes.c_void_p)]


class RTMbvh(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
                ("nodes", ctypes.c_void_p), ("index_count", ctypes.c_uint32),
                ("indices", ctypes.c_void_p)]


class RTBatch(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("scene_count", ctypes.c_uint32),
                ("bvh_node_count", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
                ("bvh_nodes", ctypes.c_void_p), ("nodes", ctypes.c_void_p),
                ("index_count", ctypes.c_uint32), ("pad_", ctypes.c_uint32),
                ("indices", ctypes.c_void_p)]


class RTBatchPlan(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("scene_count", ctypes.c_uint32),
                ("prim_count", ctypes.c_uint32), ("iterations", ctypes.c_uint32)]


class RTTraceConfig(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("lds_stack", ctypes.c_uint32),
                ("occupancy_pct", ctypes.c_uint32), ("kernel", ctypes.c_uint32),
                ("refill", ctypes.c_uint32), ("schedule", ctypes.c_uint32),
                ("postpone", ctypes.c_uint32)]


class RTTraceTree(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
                ("prim_count", ctypes.c_uint32), ("max_leaf", ctypes.c_uint32),
                ("width", ctypes.c_uint32), ("pad_", ctypes.c_uint32),
                ("nodes", ctypes.c_void_p), ("triangles", ctypes.c_void_p)]


CALLBACK = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                            ctypes.c_void_p)

# symbol -> (restype, argtypes); every function declared in include/rtbvh.h
_vp, _sz, _u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32
_rc = ctypes.c_int
SIGNATURES = {
    "create_bvh": (_rc, [_vp, _sz, _vp, _sz, _sz, _u32, ctypes.POINTER(RTBvh)]),
    "create_mbvh": (_rc, [RTBvh, ctypes.POINTER(RTMbvh)]),
    "refit": (_rc, [_vp, RTBvh]),
    "intersect": (_rc, [RTBvh, _vp, _vp, _vp, _vp, CALLBACK]),
    "intersect_packet": (_rc, [RTBvh, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, CALLBACK]),
    "intersect_mbvh": (_rc, [RTMbvh, _vp, _vp, _vp, _vp, CALLBACK]),
    "intersect_mbvh_packet": (_rc, [RTMbvh, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, CALLBACK]),
    "free_bvh": (None, [RTBvh]),
    "free_mbvh": (None, [RTMbvh]),
    "rtbvh_gpu_create_bvh": (_rc, [_vp, _sz, _vp, _sz, _sz, _u32, _vp, ctypes.POINTER(RTBvh)]),
    "rtbvh_gpu_create_spatial_bvh": (_rc, [_vp, _sz, _vp, _sz, _vp, _sz, _vp,
                                          ctypes.POINTER(RTBvh)]),
    "rtbvh_gpu_create_mbvh": (_rc, [RTBvh, _vp, ctypes.POINTER(RTMbvh)]),
    "rtbvh_gpu_refit": (_rc, [_vp, RTBvh, _vp]),
    "rtbvh_gpu_intersect": (_rc, [RTBvh, _vp, _sz, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_intersect_mbvh": (_rc, [RTMbvh, _vp, _sz, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_intersect_mbvh_stats": (_rc, [RTMbvh, _vp, _sz, _vp, _sz, _vp, _u32, _vp, _vp]),
    "rtbvh_gpu_mbvh_attach_triangles": (_rc, [RTMbvh, _vp, _sz, _vp]),
    "rtbvh_gpu_intersect_mbvh_prepared": (_rc, [RTMbvh, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_occluded_mbvh_prepared": (_rc, [RTMbvh, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_create_trace_tree": (_rc, [RTBvh, _vp, _sz, _u32, _vp, ctypes.POINTER(RTTraceTree)]),
    "rtbvh_gpu_create_trace_tree8": (_rc, [RTBvh, _vp, _sz, _vp, ctypes.POINTER(RTTraceTree)]),
    "rtbvh_gpu_trace_tree_intersect": (_rc, [RTTraceTree, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_trace_tree_intersect_stats": (_rc, [RTTraceTree, _vp, _sz, _vp, _vp, _vp]),
    "rtbvh_gpu_trace_tree_occluded": (_rc, [RTTraceTree, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_free_trace_tree": (None, [RTTraceTree]),
    "rtbvh_gpu_rays_from_packets": (_rc, [_vp, _sz, _vp, _vp]),
    "rtbvh_gpu_set_trace_config": (_rc, [ctypes.POINTER(RTTraceConfig)]),
    "rtbvh_gpu_get_trace_config": (_rc, [ctypes.POINTER(RTTraceConfig)]),
    "rtbvh_gpu_autotune_mbvh": (_rc, [RTMbvh, _vp, _sz, _vp, _vp, ctypes.POINTER(RTTraceConfig)]),
    "rtbvh_gpu_autotune_trace_tree": (_rc, [RTTraceTree, _vp, _sz, _vp, _vp,
                                           ctypes.POINTER(RTTraceConfig)]),
    "rtbvh_gpu_autotune_batch": (_rc, [RTBatch, _vp, _vp, _sz, _vp, _vp,
                                      ctypes.POINTER(RTTraceConfig)]),
    "rtbvh_gpu_adjoint": (_rc, [_vp, _sz, _vp, _vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "rtbvh_gpu_batch_plan_create": (_rc, [_vp, _vp, _sz, _vp, _vp, _sz, _vp,
                                         ctypes.POINTER(RTBatchPlan)]),
    "rtbvh_gpu_batch_plan_enqueue": (_rc, [RTBatchPlan, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp,
                                          _vp]),
    "rtbvh_gpu_batch_plan_launch": (_rc, [RTBatchPlan, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp,
                                         _vp]),
    "rtbvh_gpu_batch_plan_status": (_rc, [RTBatchPlan, _vp]),
    "rtbvh_gpu_batch_plan_order_rays": (_rc, [RTBatchPlan, _sz, _vp]),
    "rtbvh_gpu_batch_plan_free": (None, [RTBatchPlan]),
    "rtbvh_gpu_create_batch": (_rc, [_vp, _vp, _sz, _vp, _sz, _vp, ctypes.POINTER(RTBatch)]),
    "rtbvh_gpu_batch_attach_triangles": (_rc, [ctypes.POINTER(RTBatch), _vp, _sz, _vp]),
    "rtbvh_gpu_intersect_batch": (_rc, [RTBatch, _vp, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_occluded_batch": (_rc, [RTBatch, _vp, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_free_batch": (None, [RTBatch]),
    "rtbvh_gpu_occluded": (_rc, [RTBvh, _vp, _sz, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_occluded_mbvh": (_rc, [RTMbvh, _vp, _sz, _vp, _sz, _vp, _u32, _vp]),
    "rtbvh_gpu_create_bvh_noalloc": (_rc, [_vp, _sz, _vp, _sz, _vp, _vp, _vp]),
    "rtbvh_gpu_status": (_rc, [_vp]),
    "rtbvh_gpu_free_bvh": (None, [RTBvh]),
    "rtbvh_gpu_free_mbvh": (None, [RTMbvh]),
    "rtbvh_last_error": (ctypes.c_char_p, []),
}

_lib = None


class RTBvhError(RuntimeError):
    """A ResultCode other than Ok from the C ABI."""

    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        name = RESULT_NAMES.get(code, f"code {code}")
        super().__init__(f"{where}: {name}" + (f" ({detail})" if detail else ""))


def lib():
    """Load librtbvh_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C rtbvh_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue  # reported by missing_symbols()
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def missing_symbols():
    """Symbols declared in include/rtbvh.h that the built library lacks."""
    L = lib()
    return [name for name in SIGNATURES if getattr(L, name, None) is None]


def check(code: int, where: str):
    if code != OK:
        detail = lib().rtbvh_last_error().decode(errors="replace")
        raise RTBvhError(code, where, detail)
