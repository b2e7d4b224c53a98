"""rtbvh_amd — MI355X (gfx950) implementation of meirbon/rtbvh's hot path.

Host-side mirror of the reference's Rust API (src/bvh.rs):

    Builder(aabbs, primitives, primitives_per_leaf)
        .construct_locally_ordered_clustered() -> Bvh     (GPU PLOC)
        .construct_binned_sah()                -> Bvh     (GPU binned SAH)
    Mbvh.from_bvh(bvh)                         -> Mbvh    (GPU collapse)
    bvh.refit(aabbs)
    bvh / mbvh .intersect(vertices, rays)      closest hit on the GPU
    bvh / mbvh .occluded(vertices, rays)       any hit on the GPU

Data lives in device memory (torch tensors on a ROCm device); the work runs
in librtbvh_amd.so through the C ABI of include/rtbvh.h.
"""
from . import _lib
from .api import (Bvh, BuildError, Builder, Mbvh, RAY_FLOATS, HIT_FLOATS, hits_prim,
                  rays_tensor)

__all__ = ["Builder", "Bvh", "Mbvh", "BuildError", "RAY_FLOATS", "HIT_FLOATS", "hits_prim",
           "rays_tensor", "_lib"]
