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
    SceneBatch(centres, scene_offsets, aabbs)  many scenes, one build + one launch

Data lives in device memory (torch tensors on a ROCm device); the work runs
in librtbvh_amd.so through the C ABI of include/rtbvh.h.
"""
from . import _lib
from ._lib import RTBvhError
from .api import (Bvh, BuildError, Builder, Mbvh, SceneBatch, RAY_FLOATS, HIT_FLOATS,
                  hits_prim, rays_tensor)

__all__ = ["Builder", "Bvh", "Mbvh", "SceneBatch", "BuildError", "RAY_FLOATS", "HIT_FLOATS", "hits_prim",
           "rays_tensor", "RTBvhError", "_lib"]
