"""Host-side mirror of rtbvh's Builder / Bvh / Mbvh API over the C ABI.

Reference interface (src/bvh.rs):
  Builder { aabbs: Option<&[Aabb]>, primitives: &[T], primitives_per_leaf }   :48-56
  Builder::construct_binned_sah / construct_locally_ordered_clustered         :87-137
      -> Result<Bvh, BuildError>; BuildError::{NoPrimitives,
         InequalAabbsAndPrimitives(aabbs, prims)}                              :26-46
  Bvh::refit(&[Aabb])                                                          :176-205
  Mbvh::from(Bvh) / Mbvh::construct(&Bvh)                                      :381-404,446-450
  Bvh/Mbvh::traverse_iter* (candidate iterators)                               :258-283,418-443

On the GPU the candidate iterator and the caller's triangle test are fused:
``intersect`` returns the closest hit per ray and ``occluded`` the any-hit
flag.  Tensors are torch tensors on a ROCm device:
  primitives  (N, 3) or (N, 4) float32 centres (center_stride 12 / 16 bytes)
  aabbs       (N, 8) float32 RTAabb records, or None
  vertices    (N, 9) float32 triangles v0, v1, v2
  rays        (R, 8) float32: origin.xyz, t_min, direction.xyz, t_max
  hits        (R, 4) float32: t, u, v, prim id bits (RTBVH_MISS on a miss)
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib as L

RAY_FLOATS = 8
HIT_FLOATS = 4
AABB_FLOATS = 8
MBVH_FLOATS = 32


class BuildError(Exception):
    """rtbvh::BuildError (src/bvh.rs:26-46)."""

    def __init__(self, kind: str, aabbs: int = 0, prims: int = 0):
        self.kind = kind
        self.aabbs = aabbs
        self.prims = prims
        if kind == "NoPrimitives":
            msg = "No primitives supplied"
        else:
            msg = f"#Aabbs({aabbs}) != #Primitives({prims})"
        super().__init__(msg)

    def __eq__(self, other):
        return isinstance(other, BuildError) and (self.kind, self.aabbs, self.prims) == (
            other.kind, other.aabbs, other.prims)

    def __hash__(self):
        return hash((self.kind, self.aabbs, self.prims))


def _stream(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _ptr(t: Optional[torch.Tensor]) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _require_device(t: torch.Tensor, name: str, dtype=torch.float32):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a torch tensor on a ROCm device")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def rays_tensor(origins: torch.Tensor, directions: torch.Tensor, t_min: float = 1e-4,
                t_max: float = 1e34) -> torch.Tensor:
    """Pack origins/directions into (R, 8) ray records (Ray::new defaults,
    src/ray.rs:163-182)."""
    n = origins.shape[0]
    r = torch.empty((n, RAY_FLOATS), dtype=torch.float32, device=origins.device)
    r[:, 0:3] = origins
    r[:, 3] = t_min
    r[:, 4:7] = directions
    r[:, 7] = t_max
    return r


def rays_from_packets(packets: torch.Tensor) -> torch.Tensor:
    """(P, 40) float32 RayPacket4 records (src/ray.rs:44-59) -> (4P, 8) ray
    records on the device; ray 4p + i is lane i of packet p."""
    _require_device(packets, "packets")
    if packets.dim() != 2 or packets.shape[1] != 40:
        raise ValueError("packets must be (P, 40) float32 RayPacket4 records")
    rays = torch.empty((packets.shape[0] * 4, RAY_FLOATS), dtype=torch.float32,
                       device=packets.device)
    L.check(L.lib().rtbvh_gpu_rays_from_packets(_ptr(packets), packets.shape[0], _ptr(rays),
                                                _stream(packets)),
            "rtbvh_gpu_rays_from_packets")
    return rays


def hits_prim(hits: torch.Tensor) -> torch.Tensor:
    """Primitive ids of (R, 4) hit records as int64 (-1 on a miss)."""
    p = hits[:, 3].contiguous().view(torch.int32).to(torch.int64)
    return torch.where(p < 0, p & 0xFFFFFFFF, p).masked_fill(p == -1, -1)


def _hip():
    """libamdhip64 for device-to-device copies of library-owned arrays."""
    global _HIP
    try:
        return _HIP
    except NameError:
        pass
    _HIP = ctypes.CDLL("libamdhip64.so")
    _HIP.hipMemcpyAsync.restype = ctypes.c_int
    _HIP.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_int, ctypes.c_void_p]
    return _HIP


def _copy_out(ptr: int, count: int, floats_per: int, dtype, device) -> torch.Tensor:
    out = torch.empty((count, floats_per), dtype=dtype, device=device)
    if count:
        nbytes = count * floats_per * 4
        e = _hip().hipMemcpyAsync(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr),
                                  nbytes, 3, _stream(out))  # hipMemcpyDeviceToDevice
        if e != 0:
            raise RuntimeError(f"hipMemcpyAsync failed ({e})")
    return out


class _Tree:
    _kind = "bvh"

    def __init__(self, handle, device, owned: bool, keep=None):
        self._h = handle
        self.device = device
        self._owned = owned
        self._keep = keep  # tensors that back a non-owned handle

    @property
    def node_count(self) -> int:
        return int(self._h.node_count)

    @property
    def index_count(self) -> int:
        return int(self._h.index_count)

    def prim_count(self) -> int:
        return self.index_count

    def indices(self) -> torch.Tensor:
        """prim_indices as an int32 device tensor (copy)."""
        if self._keep is not None:
            return self._keep[1].clone()
        return _copy_out(self._h.indices, self.index_count, 1, torch.int32, self.device).view(-1)

    def _trace(self, fn, vertices, rays, out, flags):
        _require_device(rays, "rays")
        if rays.dim() != 2 or rays.shape[1] != RAY_FLOATS:
            raise ValueError("rays must be (R, 8) float32")
        _require_device(vertices, "vertices")
        verts = vertices.reshape(-1)
        if verts.numel() % 9:
            raise ValueError("vertices must hold 9 floats per triangle")
        code = fn(self._h, _ptr(verts), verts.numel() // 9, _ptr(rays), rays.shape[0],
                  _ptr(out), flags, _stream(rays))
        L.check(code, fn.__name__)
        return out

    def status(self):
        """Synchronise and report whether a ray overflowed its stack."""
        L.check(L.lib().rtbvh_gpu_status(_stream(torch.empty(0, device=self.device))),
                "rtbvh_gpu_status")


class Bvh(_Tree):
    """rtbvh::Bvh (src/bvh.rs:143-147) resident in device memory."""

    @classmethod
    def from_arrays(cls, nodes: torch.Tensor, indices: torch.Tensor) -> "Bvh":
        """Wrap caller-owned device arrays (nodes (M, 8) float32 RTBvhNode
        records, indices int32) without copying, like the FFI's from_slices."""
        _require_device(nodes, "nodes")
        _require_device(indices, "indices", torch.int32)
        h = L.RTBvh(0xFFFFFFFF, nodes.shape[0], nodes.data_ptr(), indices.numel(),
                    indices.data_ptr())
        return cls(h, nodes.device, owned=False, keep=(nodes, indices))

    def nodes(self) -> torch.Tensor:
        """Nodes as (M, 8) float32 RTBvhNode records (copy)."""
        if self._keep is not None:
            return self._keep[0].clone()
        return _copy_out(self._h.nodes, self.node_count, AABB_FLOATS, torch.float32,
                         self.device)

    def intersect(self, vertices, rays, reference_order: bool = False, out=None):
        """Closest hit per ray -> (R, 4) hit records."""
        out = torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32,
                          device=rays.device) if out is None else out
        flags = L.TRACE_REFERENCE_ORDER if reference_order else 0
        return self._trace(L.lib().rtbvh_gpu_intersect, vertices, rays, out, flags)

    def occluded(self, vertices, rays, reference_order: bool = False, out=None):
        """Any hit per ray -> uint8 (R,)."""
        out = torch.empty((rays.shape[0],), dtype=torch.uint8,
                          device=rays.device) if out is None else out
        flags = L.TRACE_REFERENCE_ORDER if reference_order else 0
        return self._trace(L.lib().rtbvh_gpu_occluded, vertices, rays, out, flags)

    @property
    def primitive_count(self) -> int:
        """Primitives the tree was built over (index_count counts references,
        more than the primitives after spatial splits)."""
        return int(getattr(self, "_prims", self.index_count))

    def refit(self, aabbs: torch.Tensor):
        """Bvh::refit (src/bvh.rs:176-205) on the GPU, in place."""
        if not self._owned:
            raise ValueError("refit needs a tree built by this library")
        _require_device(aabbs, "aabbs")
        if aabbs.shape != (self.primitive_count, AABB_FLOATS):
            raise ValueError("aabbs must be (prim_count, 8)")
        L.check(L.lib().rtbvh_gpu_refit(_ptr(aabbs), self._h, _stream(aabbs)), "rtbvh_gpu_refit")

    def trace_tree(self, vertices: torch.Tensor, max_leaf: int = 1,
                   width: int = 8) -> "TraceTree":
        """Traversal tree over `vertices` (see TraceTree)."""
        return TraceTree(self, vertices, max_leaf, width)

    def __del__(self):
        if getattr(self, "_owned", False) and L._lib is not None:
            L._lib.rtbvh_gpu_free_bvh(self._h)
            self._owned = False


class _DeviceSpan:
    """A device byte range presented through __cuda_array_interface__."""

    def __init__(self, ptr: int, nbytes: int):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1",
                                         "data": (ptr, False), "version": 2}


def _device_bytes(ptr, nbytes: int, device) -> torch.Tensor:
    if not ptr or nbytes == 0:
        return torch.empty(0, dtype=torch.uint8, device=device)
    torch.cuda.synchronize(device)
    return torch.as_tensor(_DeviceSpan(int(ptr), nbytes), device=device).clone()


class TraceTree:
    """Traversal tree with leaf-ordered triangles, built from a Bvh for the
    lowest-(t, id) traversal; hits equal Mbvh.intersect's default mode on the
    same triangles.  width 4: rtbvh_gpu_create_trace_tree (leaves of up to
    max_leaf triangles); width 8: rtbvh_gpu_create_trace_tree8 (compressed
    8-wide nodes, one triangle per leaf slot)."""

    def __init__(self, bvh: "Bvh", vertices: torch.Tensor, max_leaf: int = 1, width: int = 8):
        if not bvh._owned:
            raise ValueError("a trace tree needs a Bvh built by this library")
        if width not in (4, 8):
            raise ValueError("width must be 4 or 8")
        _require_device(vertices, "vertices")
        verts = vertices.reshape(-1)
        self.device = vertices.device
        # the tree holds a snapshot of these vertices: remember which tensor
        # and which version of it (in-place writes bump _version), and keep it
        # alive so that its memory cannot be handed to another tensor
        self._source = vertices
        self._source_version = vertices._version
        self._owned = False
        self._h = L.RTTraceTree()
        if width == 8:
            L.check(L.lib().rtbvh_gpu_create_trace_tree8(bvh._h, _ptr(verts), verts.numel() // 9,
                                                         _stream(verts), ctypes.byref(self._h)),
                    "rtbvh_gpu_create_trace_tree8")
        else:
            L.check(L.lib().rtbvh_gpu_create_trace_tree(bvh._h, _ptr(verts), verts.numel() // 9,
                                                        max_leaf, _stream(verts),
                                                        ctypes.byref(self._h)),
                    "rtbvh_gpu_create_trace_tree")
        self._owned = True

    def built_from(self, vertices: torch.Tensor) -> bool:
        """True when `vertices` is the tensor (or a view of the same memory
        and shape, e.g. a .detach()) this tree was built from, unmodified
        since."""
        src = self._source
        return (vertices.data_ptr() == src.data_ptr() and vertices.numel() == src.numel()
                and vertices.device == src.device and vertices._version == self._source_version)

    @property
    def node_count(self) -> int:
        return int(self._h.node_count)

    @property
    def width(self) -> int:
        return int(self._h.width)

    def node_bytes(self) -> torch.Tensor:
        """Copy of the node array as bytes (node_count x 128 for width 4,
        node_count x 80 for width 8)."""
        rec = 80 if self.width == 8 else 128
        return _device_bytes(self._h.nodes, self.node_count * rec, self.device)

    def triangles(self) -> torch.Tensor:
        """Copy of the leaf-ordered triangle records, (prim_count, 12) float32
        (v0 with the primitive id, e1, e2; the records' padding dropped)."""
        b = _device_bytes(self._h.triangles, int(self._h.prim_count) * 64, self.device)
        return b.view(torch.float32).reshape(-1, 16)[:, :12].contiguous()

    def intersect_stats(self, rays: torch.Tensor, out=None):
        """intersect() plus (nodes visited, triangles tested) over all rays."""
        _require_device(rays, "rays")
        out = torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32,
                          device=rays.device) if out is None else out
        counters = torch.zeros(2, dtype=torch.int64, device=rays.device)
        L.check(L.lib().rtbvh_gpu_trace_tree_intersect_stats(self._h, _ptr(rays), rays.shape[0],
                                                             _ptr(out), _ptr(counters),
                                                             _stream(rays)),
                "rtbvh_gpu_trace_tree_intersect_stats")
        c = counters.cpu()
        return out, int(c[0]), int(c[1])

    def intersect(self, rays: torch.Tensor, out=None, order_rays: bool = False):
        """Closest hits; order_rays traces the rays grouped by the cell of
        their origin (RTBVH_TRACE_ORDER_RAYS: same results, for large sets of
        unrelated rays)."""
        _require_device(rays, "rays")
        out = torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32,
                          device=rays.device) if out is None else out
        flags = L.TRACE_ORDER_RAYS if order_rays else 0
        L.check(L.lib().rtbvh_gpu_trace_tree_intersect(self._h, _ptr(rays), rays.shape[0],
                                                       _ptr(out), flags, _stream(rays)),
                "rtbvh_gpu_trace_tree_intersect")
        return out

    def occluded(self, rays: torch.Tensor, out=None, order_rays: bool = False):
        """Any hit per ray (1 occluded, 0 not); order_rays as for intersect."""
        _require_device(rays, "rays")
        out = torch.empty((rays.shape[0],), dtype=torch.uint8,
                          device=rays.device) if out is None else out
        flags = L.TRACE_ORDER_RAYS if order_rays else 0
        L.check(L.lib().rtbvh_gpu_trace_tree_occluded(self._h, _ptr(rays), rays.shape[0],
                                                      _ptr(out), flags, _stream(rays)),
                "rtbvh_gpu_trace_tree_occluded")
        return out

    def status(self):
        L.check(L.lib().rtbvh_gpu_status(_stream(torch.empty(0, device=self.device))),
                "rtbvh_gpu_status")

    def __del__(self):
        if getattr(self, "_owned", False) and L._lib is not None:
            L._lib.rtbvh_gpu_free_trace_tree(self._h)
            self._owned = False


class Mbvh(_Tree):
    """rtbvh::Mbvh (src/bvh.rs:320-324) resident in device memory."""

    @classmethod
    def from_bvh(cls, bvh: Bvh) -> "Mbvh":
        """From<Bvh> for Mbvh (src/bvh.rs:446-450): GPU collapse."""
        if not bvh._owned:
            raise ValueError("Mbvh.from_bvh needs a tree built by this library")
        h = L.RTMbvh()
        s = _stream(torch.empty(0, device=bvh.device))
        L.check(L.lib().rtbvh_gpu_create_mbvh(bvh._h, s, ctypes.byref(h)),
                "rtbvh_gpu_create_mbvh")
        return cls(h, bvh.device, owned=True)

    @classmethod
    def from_arrays(cls, nodes: torch.Tensor, indices: torch.Tensor) -> "Mbvh":
        """Wrap caller-owned device arrays (nodes (M, 32) float32 RTMbvhNode
        records, indices int32)."""
        _require_device(nodes, "nodes")
        _require_device(indices, "indices", torch.int32)
        h = L.RTMbvh(0xFFFFFFFF, nodes.shape[0], nodes.data_ptr(), indices.numel(),
                     indices.data_ptr())
        return cls(h, nodes.device, owned=False, keep=(nodes, indices))

    def quad_nodes(self) -> torch.Tensor:
        """MBVH nodes as (M, 32) float32 RTMbvhNode records (copy)."""
        if self._keep is not None:
            return self._keep[0].clone()
        return _copy_out(self._h.nodes, self.node_count, MBVH_FLOATS, torch.float32,
                         self.device)

    def intersect(self, vertices, rays, reference_order: bool = False, out=None):
        out = torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32,
                          device=rays.device) if out is None else out
        flags = L.TRACE_REFERENCE_ORDER if reference_order else 0
        return self._trace(L.lib().rtbvh_gpu_intersect_mbvh, vertices, rays, out, flags)

    def occluded(self, vertices, rays, reference_order: bool = False, out=None):
        out = torch.empty((rays.shape[0],), dtype=torch.uint8,
                          device=rays.device) if out is None else out
        flags = L.TRACE_REFERENCE_ORDER if reference_order else 0
        return self._trace(L.lib().rtbvh_gpu_occluded_mbvh, vertices, rays, out, flags)

    def attach_triangles(self, vertices):
        """Device copy of the triangles in leaf order, used by the
        *_prepared traversals (a snapshot of `vertices`)."""
        if not self._owned:
            raise ValueError("attach_triangles needs an MBVH built by this library")
        _require_device(vertices, "vertices")
        verts = vertices.reshape(-1)
        L.check(L.lib().rtbvh_gpu_mbvh_attach_triangles(self._h, _ptr(verts), verts.numel() // 9,
                                                        _stream(verts)),
                 "rtbvh_gpu_mbvh_attach_triangles")

    def intersect_prepared(self, rays, reference_order: bool = False, out=None):
        """intersect() over the attached triangles."""
        _require_device(rays, "rays")
        out = torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32,
                          device=rays.device) if out is None else out
        flags = L.TRACE_REFERENCE_ORDER if reference_order else 0
        L.check(L.lib().rtbvh_gpu_intersect_mbvh_prepared(self._h, _ptr(rays), rays.shape[0],
                                                          _ptr(out), flags, _stream(rays)),
                "rtbvh_gpu_intersect_mbvh_prepared")
        return out

    def occluded_prepared(self, rays, reference_order: bool = False, out=None):
        """occluded() over the attached triangles."""
        _require_device(rays, "rays")
        out = torch.empty((rays.shape[0],), dtype=torch.uint8,
                          device=rays.device) if out is None else out
        flags = L.TRACE_REFERENCE_ORDER if reference_order else 0
        L.check(L.lib().rtbvh_gpu_occluded_mbvh_prepared(self._h, _ptr(rays), rays.shape[0],
                                                         _ptr(out), flags, _stream(rays)),
                "rtbvh_gpu_occluded_mbvh_prepared")
        return out

    def __del__(self):
        if getattr(self, "_owned", False) and L._lib is not None:
            L._lib.rtbvh_gpu_free_mbvh(self._h)
            self._owned = False


class SceneBatch:
    """Many scenes of uneven size built and traced together (no counterpart in
    the reference, which builds one scene per Builder).  The scenes'
    primitives are concatenated; scene s owns [scene_offsets[s],
    scene_offsets[s+1]).  Each scene's tree is the PLOC tree of that scene on
    its own, and intersect() returns the hits of tracing each scene on its
    own, with scene-local primitive ids."""

    def __init__(self, primitives: torch.Tensor, scene_offsets, aabbs: Optional[torch.Tensor] = None):
        _require_device(primitives, "primitives")
        if primitives.dim() != 2 or primitives.shape[1] not in (3, 4):
            raise ValueError("primitives must be (N, 3) or (N, 4) float32 centres")
        if aabbs is not None:
            _require_device(aabbs, "aabbs")
            if aabbs.shape != (primitives.shape[0], AABB_FLOATS):
                raise ValueError("aabbs must be (N, 8) RTAabb records")
        offs = [int(x) for x in scene_offsets]
        if len(offs) < 2 or offs[-1] != primitives.shape[0]:
            raise ValueError("scene_offsets must have scene_count + 1 entries ending at N")
        self.scene_offsets = offs
        host = (ctypes.c_uint32 * len(offs))(*offs)
        self._h = L.RTBatch()
        self.device = primitives.device
        self._owned = False
        L.check(L.lib().rtbvh_gpu_create_batch(_ptr(aabbs), _ptr(primitives),
                                               4 * int(primitives.shape[1]), host, len(offs) - 1,
                                               _stream(primitives), ctypes.byref(self._h)),
                "rtbvh_gpu_create_batch")
        self._owned = True

    @property
    def scene_count(self) -> int:
        return int(self._h.scene_count)

    @property
    def node_count(self) -> int:
        return int(self._h.node_count)

    @property
    def bvh_node_count(self) -> int:
        return int(self._h.bvh_node_count)

    @property
    def index_count(self) -> int:
        return int(self._h.index_count)

    def bvh_nodes(self) -> torch.Tensor:
        """Binary forest as (M, 8) float32 RTBvhNode records (copy); node s is
        the root of scene s."""
        return _copy_out(self._h.bvh_nodes, self.bvh_node_count, AABB_FLOATS, torch.float32,
                         self.device)

    def quad_nodes(self) -> torch.Tensor:
        """Traversal-tree forest as (M, 32) float32 records (copy, after
        attach_triangles); node s is the root of scene s."""
        return _copy_out(self._h.nodes, self.node_count, MBVH_FLOATS, torch.float32, self.device)

    def indices(self) -> torch.Tensor:
        """Batch-wide primitive ids in leaf order (int32 copy)."""
        return _copy_out(self._h.indices, self.index_count, 1, torch.int32, self.device).view(-1)

    def attach_triangles(self, vertices: torch.Tensor):
        _require_device(vertices, "vertices")
        verts = vertices.reshape(-1)
        L.check(L.lib().rtbvh_gpu_batch_attach_triangles(ctypes.byref(self._h), _ptr(verts),
                                                         verts.numel() // 9, _stream(verts)),
                "rtbvh_gpu_batch_attach_triangles")

    @staticmethod
    def ray_offsets(counts, device) -> torch.Tensor:
        """Device offsets (scene_count + 1 uint32 words, held in an int32
        tensor) from per-scene ray counts."""
        import numpy as np

        off = np.zeros(len(counts) + 1, dtype=np.uint64)
        off[1:] = np.cumsum(np.asarray(list(counts), dtype=np.uint64))
        if int(off[-1]) > 0xFFFFFFFF:
            raise ValueError("a batch traces at most 2^32-1 rays")
        return torch.from_numpy(off.astype(np.uint32).view(np.int32)).to(device)

    def _trace(self, fn, rays, ray_offsets, out):
        _require_device(rays, "rays")
        if rays.dim() != 2 or rays.shape[1] != RAY_FLOATS:
            raise ValueError("rays must be (R, 8) float32")
        if not isinstance(ray_offsets, torch.Tensor) or ray_offsets.device != rays.device \
                or ray_offsets.element_size() != 4 or ray_offsets.numel() != self.scene_count + 1:
            raise ValueError("ray_offsets must be a 32-bit device tensor of scene_count + 1 entries")
        L.check(fn(self._h, _ptr(rays), _ptr(ray_offsets.contiguous()), rays.shape[0], _ptr(out),
                   0, _stream(rays)), fn.__name__)
        return out

    def intersect(self, rays: torch.Tensor, ray_offsets: torch.Tensor, out=None):
        """Closest hit per ray (lowest (t, id) rule), scene-local prim ids."""
        out = torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32,
                          device=rays.device) if out is None else out
        return self._trace(L.lib().rtbvh_gpu_intersect_batch, rays, ray_offsets, out)

    def occluded(self, rays: torch.Tensor, ray_offsets: torch.Tensor, out=None):
        out = torch.empty((rays.shape[0],), dtype=torch.uint8,
                          device=rays.device) if out is None else out
        return self._trace(L.lib().rtbvh_gpu_occluded_batch, rays, ray_offsets, out)

    def __del__(self):
        if getattr(self, "_owned", False) and L._lib is not None:
            L._lib.rtbvh_gpu_free_batch(self._h)
            self._owned = False


class BatchPlan:
    """A batch shape (scene offsets fixed) whose whole frame — PLOC build,
    MBVH collapse, leaf-ordered triangles, traversal — is enqueued without
    allocation or host synchronisation (rtbvh_gpu_batch_plan_*), so that it
    can run as one hipGraph.  No counterpart in the reference."""

    def __init__(self, primitives: torch.Tensor, scene_offsets, vertices: torch.Tensor,
                 aabbs: Optional[torch.Tensor] = None):
        _require_device(primitives, "primitives")
        _require_device(vertices, "vertices")
        if aabbs is not None:
            _require_device(aabbs, "aabbs")
        offs = [int(x) for x in scene_offsets]
        if len(offs) < 2 or offs[-1] != primitives.shape[0]:
            raise ValueError("scene_offsets must have scene_count + 1 entries ending at N")
        self.scene_offsets = offs
        self.device = primitives.device
        self._h = L.RTBatchPlan()
        self._owned = False
        host = (ctypes.c_uint32 * len(offs))(*offs)
        L.check(L.lib().rtbvh_gpu_batch_plan_create(_ptr(aabbs), _ptr(primitives),
                                                    4 * int(primitives.shape[1]),
                                                    _ptr(vertices.reshape(-1)), host,
                                                    len(offs) - 1, _stream(primitives),
                                                    ctypes.byref(self._h)),
                "rtbvh_gpu_batch_plan_create")
        self._owned = True

    @property
    def iterations(self) -> int:
        return int(self._h.iterations)

    def _run(self, fn, primitives, aabbs, vertices, rays, ray_offsets, out, any_hit):
        _require_device(primitives, "primitives")
        _require_device(vertices, "vertices")
        _require_device(rays, "rays")
        if aabbs is not None:
            _require_device(aabbs, "aabbs")
        if ray_offsets is not None and (ray_offsets.element_size() != 4 or
                                        ray_offsets.numel() != len(self.scene_offsets)):
            raise ValueError("ray_offsets must be a 32-bit device tensor of scene_count + 1 entries")
        hits = None if any_hit else out
        occ = out if any_hit else None
        L.check(fn(self._h, _ptr(aabbs), _ptr(primitives), _ptr(vertices.reshape(-1)), _ptr(rays),
                   _ptr(ray_offsets), rays.shape[0], _ptr(hits), _ptr(occ), _stream(rays)),
                fn.__name__)
        return out

    def _out(self, rays, any_hit, out):
        if out is not None:
            return out
        if any_hit:
            return torch.empty((rays.shape[0],), dtype=torch.uint8, device=rays.device)
        return torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32, device=rays.device)

    def enqueue(self, primitives, vertices, rays, ray_offsets=None, aabbs=None, any_hit=False,
                out=None):
        """Build + trace on the current stream (stream-ordered, eager launches)."""
        out = self._out(rays, any_hit, out)
        return self._run(L.lib().rtbvh_gpu_batch_plan_enqueue, primitives, aabbs, vertices, rays,
                         ray_offsets, out, any_hit)

    def launch(self, primitives, vertices, rays, ray_offsets=None, aabbs=None, any_hit=False,
               out=None):
        """The same frame as one hipGraph (recorded on first use of these
        tensors, replayed afterwards; keep passing the same tensors)."""
        out = self._out(rays, any_hit, out)
        return self._run(L.lib().rtbvh_gpu_batch_plan_launch, primitives, aabbs, vertices, rays,
                         ray_offsets, out, any_hit)

    def order_rays(self, max_rays: int):
        """Trace later frames (one scene, at most max_rays rays) grouped by
        the cells of the ray origins, from plan-owned buffers of about 16
        bytes per ray (rtbvh_gpu_batch_plan_order_rays); 0 returns to the caller's
        order."""
        L.check(L.lib().rtbvh_gpu_batch_plan_order_rays(
            self._h, int(max_rays), _stream(torch.empty(0, device=self.device))),
            "rtbvh_gpu_batch_plan_order_rays")

    def status(self):
        L.check(L.lib().rtbvh_gpu_batch_plan_status(self._h, _stream(torch.empty(0, device=self.device))),
                "rtbvh_gpu_batch_plan_status")

    def __del__(self):
        if getattr(self, "_owned", False) and L._lib is not None:
            L._lib.rtbvh_gpu_batch_plan_free(self._h)
            self._owned = False


class Builder:
    """rtbvh::Builder (src/bvh.rs:48-56): primitives are the centres."""

    def __init__(self, primitives: torch.Tensor, aabbs: Optional[torch.Tensor] = None,
                 primitives_per_leaf: Optional[int] = None):
        self.primitives = primitives
        self.aabbs = aabbs
        self.primitives_per_leaf = primitives_per_leaf

    def _construct(self, bvh_type: int) -> Bvh:
        n = 0 if self.primitives is None else int(self.primitives.shape[0])
        if n == 0:  # bvh.rs:88-90,114-116
            raise BuildError("NoPrimitives")
        if self.aabbs is not None and int(self.aabbs.shape[0]) != n:  # bvh.rs:92-98,118-124
            raise BuildError("InequalAabbsAndPrimitives", int(self.aabbs.shape[0]), n)
        _require_device(self.primitives, "primitives")
        if self.primitives.dim() != 2 or self.primitives.shape[1] not in (3, 4):
            raise ValueError("primitives must be (N, 3) or (N, 4) float32 centres")
        if self.aabbs is not None:
            _require_device(self.aabbs, "aabbs")
            if self.aabbs.shape[1] != AABB_FLOATS:
                raise ValueError("aabbs must be (N, 8) RTAabb records")
        h = L.RTBvh()
        stride = 4 * int(self.primitives.shape[1])
        ppl = self.primitives_per_leaf or 0
        code = L.lib().rtbvh_gpu_create_bvh(_ptr(self.aabbs), n, _ptr(self.primitives), stride,
                                            ppl, bvh_type, _stream(self.primitives),
                                            ctypes.byref(h))
        if code == L.NO_PRIMITIVES:
            raise BuildError("NoPrimitives")
        L.check(code, "rtbvh_gpu_create_bvh")
        return Bvh(h, self.primitives.device, owned=True)

    def construct_locally_ordered_clustered(self) -> Bvh:
        """src/bvh.rs:113-137 — PLOC (search radius 14) on the GPU."""
        return self._construct(L.LOCALLY_ORDERED_CLUSTERED)

    def construct_binned_sah(self) -> Bvh:
        """src/bvh.rs:87-111 — 16-bin SAH on the GPU."""
        return self._construct(L.BINNED_SAH)

    def construct_spatial_sah(self, vertices: torch.Tensor) -> Bvh:
        """Builder::construct_spatial_sah (src/bvh.rs, SpatialSahBuilder,
        src/builders/spatial_sah.rs:247-1034) on the GPU: object and spatial
        splits over `vertices` ((N, 9) or (N, 3, 3) float32, device), which
        the straddling references are clipped against.  The tree's
        index_count counts the references (a triangle may sit in several
        leaves); primitive_count is N."""
        n = 0 if self.primitives is None else int(self.primitives.shape[0])
        if n == 0:
            raise BuildError("NoPrimitives")
        if self.aabbs is not None and int(self.aabbs.shape[0]) != n:
            raise BuildError("InequalAabbsAndPrimitives", int(self.aabbs.shape[0]), n)
        _require_device(self.primitives, "primitives")
        _require_device(vertices, "vertices")
        if self.primitives.dim() != 2 or self.primitives.shape[1] not in (3, 4):
            raise ValueError("primitives must be (N, 3) or (N, 4) float32 centres")
        verts = vertices.reshape(-1).contiguous()
        if verts.numel() != 9 * n or verts.dtype != torch.float32:
            raise ValueError("vertices must be (N, 9) float32 triangles")
        if self.aabbs is not None:
            _require_device(self.aabbs, "aabbs")
            if self.aabbs.shape[1] != AABB_FLOATS:
                raise ValueError("aabbs must be (N, 8) RTAabb records")
        h = L.RTBvh()
        stride = 4 * int(self.primitives.shape[1])
        code = L.lib().rtbvh_gpu_create_spatial_bvh(
            _ptr(self.aabbs), n, _ptr(self.primitives), stride, _ptr(verts),
            self.primitives_per_leaf or 0, _stream(self.primitives), ctypes.byref(h))
        if code == L.NO_PRIMITIVES:
            raise BuildError("NoPrimitives")
        L.check(code, "rtbvh_gpu_create_spatial_bvh")
        bvh = Bvh(h, self.primitives.device, owned=True)
        bvh._prims = n
        return bvh


def get_trace_config() -> dict:
    """Launch shape of the lowest-id MBVH traversal on the current device."""
    c = L.RTTraceConfig()
    L.check(L.lib().rtbvh_gpu_get_trace_config(ctypes.byref(c)), "rtbvh_gpu_get_trace_config")
    return {"block": c.block, "lds_stack": c.lds_stack, "occupancy_pct": c.occupancy_pct,
            "kernel": c.kernel, "refill": c.refill, "schedule": c.schedule,
            "postpone": c.postpone}


def set_trace_config(block: Optional[int] = None, lds_stack: int = 8, occupancy_pct: int = 800,
                     kernel: int = 0, refill: int = 16, schedule: int = 1, postpone: int = 0):
    """Sets the traversal launch shape (block=None restores the default)."""
    if block is None:
        L.check(L.lib().rtbvh_gpu_set_trace_config(None), "rtbvh_gpu_set_trace_config")
        return
    c = L.RTTraceConfig(block, lds_stack, occupancy_pct, kernel, refill, schedule, postpone)
    L.check(L.lib().rtbvh_gpu_set_trace_config(ctypes.byref(c)), "rtbvh_gpu_set_trace_config")


def autotune(tree, rays: torch.Tensor, ray_offsets: Optional[torch.Tensor] = None) -> dict:
    """Times the candidate launch shapes on `rays` against `tree` (an Mbvh with
    attached triangles, a TraceTree, or a SceneBatch with its per-scene
    `ray_offsets`) and keeps the fastest for the current device."""
    _require_device(rays, "rays")
    hits = torch.empty((rays.shape[0], HIT_FLOATS), dtype=torch.float32, device=rays.device)
    c = L.RTTraceConfig()
    if isinstance(tree, SceneBatch):
        if ray_offsets is None or ray_offsets.device != rays.device \
                or ray_offsets.element_size() != 4 or ray_offsets.numel() != tree.scene_count + 1:
            raise ValueError("a SceneBatch needs ray_offsets: a 32-bit device tensor of "
                             "scene_count + 1 entries")
        fn = L.lib().rtbvh_gpu_autotune_batch
        L.check(fn(tree._h, _ptr(rays), _ptr(ray_offsets.contiguous()), rays.shape[0],
                   _ptr(hits), _stream(rays), ctypes.byref(c)), fn.__name__)
    else:
        fn = (L.lib().rtbvh_gpu_autotune_trace_tree if isinstance(tree, TraceTree)
              else L.lib().rtbvh_gpu_autotune_mbvh)
        L.check(fn(tree._h, _ptr(rays), rays.shape[0], _ptr(hits), _stream(rays),
                   ctypes.byref(c)), fn.__name__)
    return {"block": c.block, "lds_stack": c.lds_stack, "occupancy_pct": c.occupancy_pct,
            "kernel": c.kernel, "refill": c.refill, "schedule": c.schedule,
            "postpone": c.postpone}
