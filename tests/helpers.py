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


def flat_grid(g=40, cell=0.05, offset=0.0):
    """2 g^2 triangles tiling [0, g cell]^2 in the plane z = offset, shifted by
    offset in x and y (a scene with one axis of zero extent)."""
    xs, ys = np.meshgrid(np.arange(g, dtype=np.float32), np.arange(g, dtype=np.float32))
    xs, ys = xs.reshape(-1) * np.float32(cell), ys.reshape(-1) * np.float32(cell)
    z = np.zeros_like(xs)
    d = np.float32(cell)
    t1 = np.stack([xs, ys, z, xs + d, ys, z, xs, ys + d, z], 1)
    t2 = np.stack([xs + d, ys, z, xs + d, ys + d, z, xs, ys + d, z], 1)
    verts = (np.concatenate([t1, t2]) + np.float32(offset)).astype(np.float32)
    verts[:, 2::3] = np.float32(offset)
    return verts


def flat_grid_rays(verts, n, seed=11):
    """Rays for flat_grid scenes: a quarter straight down or up (zero x, y
    direction components), a quarter parallel to the plane, the rest oblique
    (some with a zero x component); every 5th with t_min 0.5, every 7th with
    t_max 1.0 -- exactly the distance of the axis rays to the plane."""
    rng = np.random.default_rng(seed)
    lo, hi = verts.reshape(-1, 3).min(0), verts.reshape(-1, 3).max(0)
    plane = np.float32(verts[0, 2])
    o = rng.uniform(lo - 0.2, hi + 0.2, (n, 3)).astype(np.float32)
    o[:, 2] = plane + rng.choice([-1.0, 1.0], n).astype(np.float32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    dirs[0::4] = np.float32(0.0)
    dirs[0::4, 2] = -np.sign(o[0::4, 2] - plane)
    dirs[1::4, 2] = np.float32(0.0)
    dirs[2::8, 0] = np.float32(0.0)
    dirs[2::8, 2] = -np.sign(o[2::8, 2] - plane) * np.abs(dirs[2::8, 2])
    tmin = np.where(np.arange(n) % 5 == 0, np.float32(0.5), np.float32(1e-4)).astype(np.float32)
    tmax = np.where(np.arange(n) % 7 == 0, np.float32(1.0), np.float32(1e30)).astype(np.float32)
    return np.concatenate([o, tmin[:, None], dirs, tmax[:, None]], 1).astype(np.float32)


def geometric_chain(n=80, ratio=1.5, seed=21):
    """n triangles along x at ratio^k, each a tenth of its distance from the
    origin in size: binned SAH peels the largest off a few per level, so the
    tree (23 levels for the default 80) is deeper than the 2 lg n + 2 levels
    of a builder's first batch, and PLOC's is a chain of n levels."""
    rng = np.random.default_rng(seed)
    x = ratio ** np.arange(n, dtype=np.float64)
    c = np.stack([x, 0.01 * rng.random(n) * x, 0.01 * rng.random(n) * x], 1)
    s = 0.1 * x
    v0 = c + np.stack([-s, -s, 0 * s], 1)
    v1 = c + np.stack([s, -s, 0 * s], 1)
    v2 = c + np.stack([0 * s, s, s], 1)
    return np.ascontiguousarray(np.concatenate([v0, v1, v2], 1), dtype=np.float32)


def tree_depth(nodes) -> int:
    """Deepest leaf of a binary node array (count >= 0 marks a leaf)."""
    depth, stack = 0, [(0, 0)]
    while stack:
        i, d = stack.pop()
        depth = max(depth, d)
        if int(nodes[i]["count"]) < 0:
            lf = int(nodes[i]["left_first"])
            stack += [(lf, d + 1), (lf + 1, d + 1)]
    return depth
