"""Shared test helpers: fixtures as arrays in the oracle's record layouts and
host <-> device conversion."""
from __future__ import annotations

import os

import numpy as np

from oracle import refcpu as ref
from rtbvh_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def teapot():
    """objects/teapot.obj, the reference's own fixture (6,320 triangles)."""
    return scenes.load_obj(os.path.join(GOLDEN, "teapot.obj"))


def as_aabb(a8: np.ndarray) -> np.ndarray:
    """(N, 8) float32 -> RTAabb structured records (same bytes)."""
    return np.ascontiguousarray(a8, dtype=np.float32).view(ref.AABB).reshape(-1)


def as_rays(r8: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(r8, dtype=np.float32).view(ref.RAY).reshape(-1)


def flat(records: np.ndarray) -> np.ndarray:
    """Structured records -> (N, k) float32 view of the same bytes."""
    n = len(records)
    return np.ascontiguousarray(records).view(np.float32).reshape(n, -1)


def prep(verts):
    """Triangle aabbs and centres exactly as the oracle computes them."""
    return ref.triangle_aabbs_centers(verts)


def ray_set(verts, n, seed=0, kind="sphere"):
    v = verts.reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    if kind == "sphere":
        c = (lo + hi) / 2
        rad = float(np.linalg.norm(hi - lo)) * 1.5
        r = scenes.rays_sphere_to_box(n, lo, hi, radius=rad, seed=seed)
        r[:, 0:3] += c.astype(np.float32)
        return r
    return scenes.rays_in_box(n, lo, hi, seed=seed)


def hits_equal(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-ray bitwise equality of hit records (t, u, v, prim)."""
    return np.all(flat(a).view(np.uint32) == flat(b).view(np.uint32), axis=1)


def to_device(arr, torch, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(arr))
    if dtype is not None:
        t = t.view(dtype)
    return t.to("cuda")
