"""rtbvh_amd — MI355X (gfx950) implementation of meirbon/rtbvh's hot path.

Host-side mirror of the reference's Rust API (src/bvh.rs):

    Builder(aabbs, primitives, primitives_per_leaf)
        .construct_locally_ordered_clustered() -> Bvh     (GPU PLOC)
        .construct_binned_sah()                -> Bvh     (GPU binned SAH)
    Mbvh.from_bvh(bvh)                         -> Mbvh    (GPU collapse)
    bvh.refit(aabbs)
    bvh / mbvh .intersect(vertices, rays)      closest hit on the GPU
    bvh / mbvh .occluded(vertices, rays)       any hit on the GPU
    mbvh.attach_triangles(vertices); mbvh.intersect_prepared(rays)
                                               leaf-ordered triangle copy
    bvh.trace_tree(vertices)                   compressed 8-wide traversal tree (fastest path)
    SceneBatch(centres, scene_offsets, aabbs)  many scenes, one build + one launch
    BatchPlan(centres, scene_offsets, aabbs)   build + trace as one hipGraph

Data lives in device memory (torch tensors on a ROCm device); the work runs
in librtbvh_amd.so through the C ABI of include/rtbvh.h.
"""
from . import _lib
from ._lib import RTBvhError
from .api import (BatchPlan, Bvh, BuildError, Builder, Mbvh, SceneBatch, TraceTree, RAY_FLOATS, HIT_FLOATS,
                  autotune, get_trace_config, hits_prim, rays_from_packets, rays_tensor,
                  set_trace_config)

__all__ = ["Builder", "Bvh", "Mbvh", "SceneBatch", "BatchPlan", "TraceTree", "BuildError", "RAY_FLOATS", "HIT_FLOATS", "hits_prim",
           "rays_tensor", "rays_from_packets", "RTBvhError", "autotune", "get_trace_config",
           "set_trace_config", "_lib"]
