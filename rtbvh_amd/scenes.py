"""Seeded synthetic scenes and ray sets (host side, numpy).

These play the role of the reference's benchmark harness (examples/
benchmark.rs, shared/src/lib.rs): they produce FP32 triangle soups (9 floats
per triangle: v0, v1, v2) and 32-byte ray records (origin, t_min, direction,
t_max) that feed both the GPU path and the CPU oracle unchanged.

Workloads (sizes as named in the benchmark configuration):
  torus(80, 40)                   6,400 tris, R=1, r=0.4
  load_obj(teapot.obj)            6,320 tris (the reference's fixture)
  noisy_sphere + random_soup      1,000,000 tris in the unit cube
"""
from __future__ import annotations

import numpy as np

RAY_T_MIN = np.float32(1e-4)  # Ray::DEFAULT_T_MIN, src/ray.rs:163
RAY_T_MAX = np.float32(1e34)  # Ray::DEFAULT_T_MAX, src/ray.rs:164


# ------------------------------------------------------------------ meshes --
def _quads_to_tris(p: np.ndarray) -> np.ndarray:
    """p: (nu+1, nv+1, 3) grid of points (the u axis wraps via the caller).
    Returns (2*nu*nv, 9) float32, two triangles per quad."""
    a = p[:-1, :-1]
    b = p[1:, :-1]
    c = p[1:, 1:]
    d = p[:-1, 1:]
    t0 = np.concatenate([a, b, c], axis=-1)
    t1 = np.concatenate([a, c, d], axis=-1)
    tris = np.stack([t0, t1], axis=2)  # (nu, nv, 2, 9)
    return np.ascontiguousarray(tris.reshape(-1, 9), dtype=np.float32)


def torus(nu: int = 80, nv: int = 40, R: float = 1.0, r: float = 0.4) -> np.ndarray:
    """Torus R, r with nu x nv quads split into 2*nu*nv triangles."""
    u = np.linspace(0.0, 2.0 * np.pi, nu + 1)
    v = np.linspace(0.0, 2.0 * np.pi, nv + 1)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    x = (R + r * np.cos(vv)) * np.cos(uu)
    y = (R + r * np.cos(vv)) * np.sin(uu)
    z = r * np.sin(vv)
    p = np.stack([x, y, z], axis=-1)
    p[-1] = p[0]
    p[:, -1] = p[:, 0]
    return _quads_to_tris(p)


def _hash3(ix, iy, iz, seed):
    """32-bit lattice hash -> [0, 1)."""
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iy.astype(np.uint32) * np.uint32(0xD8163841)) \
        ^ (iz.astype(np.uint32) * np.uint32(0xCB1AB31F)) ^ np.uint32(seed * 0x9E3779B1 & 0xFFFFFFFF)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h.astype(np.float64) * (1.0 / 4294967296.0)


def value_noise(p: np.ndarray, seed: int, octaves: int = 3, base_freq: float = 4.0,
                amplitude: float = 0.1) -> np.ndarray:
    """Lattice value noise with smoothstep trilinear interpolation, octaves
    doubling in frequency and halving in amplitude; the amplitudes sum to
    `amplitude`.  Returns values in [-amplitude, amplitude]."""
    with np.errstate(over="ignore"):
        total = np.zeros(p.shape[:-1])
        weights = [0.5 ** o for o in range(octaves)]
        norm = amplitude / sum(weights)
        for o in range(octaves):
            q = p * (base_freq * 2.0 ** o)
            i0 = np.floor(q).astype(np.int64)
            f = q - i0
            s = f * f * (3.0 - 2.0 * f)
            acc = np.zeros(p.shape[:-1])
            for dx in (0, 1):
                wx = s[..., 0] if dx else 1.0 - s[..., 0]
                for dy in (0, 1):
                    wy = s[..., 1] if dy else 1.0 - s[..., 1]
                    for dz in (0, 1):
                        wz = s[..., 2] if dz else 1.0 - s[..., 2]
                        h = _hash3(i0[..., 0] + dx, i0[..., 1] + dy, i0[..., 2] + dz, seed + o)
                        acc += wx * wy * wz * h
            total += norm * weights[o] * (2.0 * acc - 1.0)
    return total


def noisy_sphere(nu: int, nv: int, center=(0.5, 0.5, 0.5), radius: float = 0.35,
                 amplitude: float = 0.1, seed: int = 0) -> np.ndarray:
    """UV sphere of nu (longitude) x nv (latitude) quads, 2*nu*nv triangles,
    displaced along the normal by value_noise (3 octaves, total amplitude)."""
    u = np.linspace(0.0, 2.0 * np.pi, nu + 1)
    v = np.linspace(0.0, np.pi, nv + 1)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    n = np.stack([np.sin(vv) * np.cos(uu), np.sin(vv) * np.sin(uu), np.cos(vv)], axis=-1)
    n[-1] = n[0]
    c = np.asarray(center, dtype=np.float64)
    base = c + radius * n
    disp = value_noise(base, seed, amplitude=amplitude)
    disp[-1] = disp[0]
    p = c + (radius + disp)[..., None] * n
    return _quads_to_tris(p)


def random_soup(n: int, edge: float = 0.01, seed: int = 0, lo=(0.0, 0.0, 0.0),
                hi=(1.0, 1.0, 1.0)) -> np.ndarray:
    """n triangles: v0 uniform in the box, v1, v2 = v0 + edge * unit vectors."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo)
    hi = np.asarray(hi)
    v0 = lo + rng.random((n, 3)) * (hi - lo)
    e = rng.normal(size=(n, 2, 3))
    e /= np.linalg.norm(e, axis=-1, keepdims=True)
    v1 = v0 + edge * e[:, 0]
    v2 = v0 + edge * e[:, 1]
    return np.ascontiguousarray(np.concatenate([v0, v1, v2], axis=1), dtype=np.float32)


def scene_1m(seed: int = 2024) -> np.ndarray:
    """1,000,000 triangles: a 750x600-quad noisy sphere (900,000 tris, centre
    (0.5,0.5,0.5), radius 0.35, noise amplitude 0.1) plus 100,000 random
    triangles of edge 0.01 in the unit cube."""
    s = noisy_sphere(750, 600, seed=seed)
    r = random_soup(100_000, 0.01, seed=seed + 1)
    return np.ascontiguousarray(np.concatenate([s, r], axis=0))


def load_obj(path: str) -> np.ndarray:
    """Minimal OBJ reader: `v x y z` and triangular `f a b c` (1-based,
    optional /vt/vn suffixes), faces in file order."""
    verts = []
    faces = []
    with open(path, "r") as fh:
        for line in fh:
            if line.startswith("v "):
                _, x, y, z = line.split()[:4]
                verts.append((float(x), float(y), float(z)))
            elif line.startswith("f "):
                idx = [int(tok.split("/")[0]) for tok in line.split()[1:]]
                for k in range(1, len(idx) - 1):  # fan for polygons
                    faces.append((idx[0], idx[k], idx[k + 1]))
    v = np.asarray(verts, dtype=np.float32)
    f = np.asarray(faces, dtype=np.int64) - 1
    return np.ascontiguousarray(v[f].reshape(-1, 9))


def triangle_aabbs_centers(verts: np.ndarray):
    """Per-triangle aabb (src/aabb.rs:252-261 grow from empty) and centre
    (v0 + v1 + v2) * (1/3) (shared/src/lib.rs:28-30), float32.  Returns
    (aabbs as (N, 8) float32 view of RTAabb records, centers (N, 3))."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3, 3)
    third = np.float32(1.0) / np.float32(3.0)
    centers = ((v[:, 0] + v[:, 1]) + v[:, 2]) * third
    mn = np.minimum(np.minimum(v[:, 0], v[:, 1]), v[:, 2])
    mx = np.maximum(np.maximum(v[:, 0], v[:, 1]), v[:, 2])
    aabbs = np.zeros((len(v), 8), np.float32)
    aabbs[:, 0:3] = mn
    aabbs[:, 4:7] = mx
    return aabbs, np.ascontiguousarray(centers, dtype=np.float32)


# -------------------------------------------------------------------- rays --
def make_rays(origins: np.ndarray, dirs: np.ndarray, t_min=RAY_T_MIN, t_max=RAY_T_MAX):
    n = len(origins)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = origins
    r[:, 3] = t_min
    r[:, 4:7] = dirs
    r[:, 7] = t_max
    return r


def rays_sphere_to_box(n: int, lo, hi, radius: float = 3.0, seed: int = 0) -> np.ndarray:
    """Origins uniform on a sphere of `radius` around the origin, aimed at
    points uniform in the box [lo, hi]."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = radius * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = np.asarray(lo) + rng.random((n, 3)) * (np.asarray(hi) - np.asarray(lo))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(o.astype(np.float32), d.astype(np.float32))


def rays_in_box(n: int, lo, hi, seed: int = 0) -> np.ndarray:
    """Origins uniform in [lo, hi], directions uniform on the unit sphere."""
    rng = np.random.default_rng(seed)
    o = np.asarray(lo) + rng.random((n, 3)) * (np.asarray(hi) - np.asarray(lo))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(o.astype(np.float32), d.astype(np.float32))


def pinhole_rays(width: int, height: int, eye=(0.5, 0.5, -1.5), fov_deg: float = 40.0,
                 rows=None) -> np.ndarray:
    """Pinhole camera at `eye` looking along +z; the vertical field of view
    `fov_deg` (40 deg covers the unit cube from z = -1.5).  `rows` selects a
    (start, stop) band of image rows."""
    r0, r1 = (0, height) if rows is None else rows
    ys = (np.arange(r0, r1, dtype=np.float64) + 0.5) / height
    xs = (np.arange(width, dtype=np.float64) + 0.5) / width
    tan_half = np.tan(np.radians(fov_deg) * 0.5)
    aspect = width / height
    px = (2.0 * xs - 1.0) * tan_half * aspect
    py = (1.0 - 2.0 * ys) * tan_half
    gx, gy = np.meshgrid(px, py)
    d = np.stack([gx, gy, np.ones_like(gx)], axis=-1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(np.asarray(eye, dtype=np.float64), d.shape)
    return make_rays(o.astype(np.float32), d.astype(np.float32))


def reference_camera_rays(width: int = 1000, height: int = 1000) -> np.ndarray:
    """The camera of the reference's benchmark (examples/benchmark.rs:72-95,
    CameraView3D::generate_ray in shared/src/lib.rs:157-164), in float32: eye
    (0, 1.5, -100) looking along +z, pixel (x, y) through p1 + (x / width)
    right + (y / height) up.  Its field of view is converted to radians twice
    (tan(pi/4 / (180/pi)) = 0.0137 half-extent at distance 1), so these rays
    form a narrow, very coherent bundle around the teapot's middle."""
    f = np.float32
    fov = f(np.pi / 2)
    screen = f(np.tan(f(fov * f(0.5)) / f(180.0 / np.pi)))
    pos = np.array([0.0, 1.5, -100.0], f)
    center = pos + np.array([0.0, 0.0, 1.0], f)
    up = np.array([0.0, 1.0, 0.0], f)
    right = np.array([-1.0, 0.0, 0.0], f)  # Vec3::Z.cross(Vec3::Y)
    aspect = f(width / height)
    p1 = center - screen * right * aspect + screen * up
    p2 = center + screen * right * aspect + screen * up
    p3 = center - screen * right * aspect - screen * up
    r_vec, u_vec = p2 - p1, p3 - p1
    xs = np.arange(width, dtype=f) * f(1.0 / width)
    ys = np.arange(height, dtype=f) * f(1.0 / height)
    u, v = np.meshgrid(xs, ys)  # row-major: y outer, x inner
    pts = p1 + u.reshape(-1, 1) * r_vec + v.reshape(-1, 1) * u_vec
    d = (pts - pos).astype(f)
    d /= np.sqrt((d * d).sum(1, keepdims=True)).astype(f)
    return make_rays(np.broadcast_to(pos, d.shape).astype(f), d.astype(f))


# ------------------------------------------------------- device ray sets --
def pinhole_rays_torch(width: int, height: int, device, eye=(0.5, 0.5, -1.5),
                       fov_deg: float = 40.0, row0: int = 0, rows: int = None):
    """pinhole_rays() on a torch device: rows [row0, row0 + rows) of the image,
    as (rows * width, 8) ray records."""
    import torch

    rows = height - row0 if rows is None else rows
    ys = (torch.arange(row0, row0 + rows, device=device, dtype=torch.float64) + 0.5) / height
    xs = (torch.arange(width, device=device, dtype=torch.float64) + 0.5) / width
    tan_half = float(np.tan(np.radians(fov_deg) * 0.5))
    px = (2.0 * xs - 1.0) * tan_half * (width / height)
    py = (1.0 - 2.0 * ys) * tan_half
    gy, gx = torch.meshgrid(py, px, indexing="ij")
    d = torch.stack([gx, gy, torch.ones_like(gx)], dim=-1).reshape(-1, 3)
    d = d / torch.linalg.norm(d, dim=1, keepdim=True)
    r = torch.empty((d.shape[0], 8), dtype=torch.float32, device=device)
    r[:, 0:3] = torch.tensor(eye, dtype=torch.float32, device=device)
    r[:, 3] = float(RAY_T_MIN)
    r[:, 4:7] = d.to(torch.float32)
    r[:, 7] = float(RAY_T_MAX)
    return r


def _onb(n):
    """Orthonormal basis around unit normals n (Duff et al. 2017)."""
    import torch

    s = torch.where(n[:, 2] >= 0, torch.ones_like(n[:, 2]), -torch.ones_like(n[:, 2]))
    a = -1.0 / (s + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = torch.stack([1.0 + s * n[:, 0] * n[:, 0] * a, s * b, -s * n[:, 0]], dim=1)
    bb = torch.stack([b, s + n[:, 1] * n[:, 1] * a, -n[:, 1]], dim=1)
    return t, bb


def bounce_rays_torch(verts, primary, hits, generator, offset: float = 1e-4):
    """Cosine-weighted bounce rays.  For a primary hit the origin is the hit
    point moved `offset` along the geometric normal facing the incoming ray;
    for a primary miss it is a uniformly sampled surface point (area-weighted
    triangle, uniform barycentrics, random side).  verts: (N, 9) float32
    device tensor; primary (R, 8) and hits (R, 4) device tensors."""
    import torch

    dev = primary.device
    R = primary.shape[0]
    v = verts.view(-1, 3, 3).to(torch.float64)
    prim = hits[:, 3].contiguous().view(torch.int32).to(torch.int64)
    hit = prim >= 0
    e1 = v[:, 1] - v[:, 0]
    e2 = v[:, 2] - v[:, 0]
    nrm = torch.cross(e1, e2, dim=1)
    area = torch.linalg.norm(nrm, dim=1)
    nrm = nrm / area.clamp_min(1e-30)[:, None]
    cdf = torch.cumsum(area, 0)
    u = torch.rand(R, generator=generator, device=dev, dtype=torch.float64) * cdf[-1]
    rand_tri = torch.searchsorted(cdf, u).clamp_max(v.shape[0] - 1)
    tri = torch.where(hit, prim, rand_tri)
    n = nrm[tri]
    d_in = primary[:, 4:7].to(torch.float64)
    r1 = torch.rand(R, generator=generator, device=dev, dtype=torch.float64)
    r2 = torch.rand(R, generator=generator, device=dev, dtype=torch.float64)
    sq = torch.sqrt(r1)
    vt = v[tri]
    surf = (1 - sq)[:, None] * vt[:, 0] + (sq * (1 - r2))[:, None] * vt[:, 1] + (sq * r2)[:, None] * vt[:, 2]
    p_hit = primary[:, 0:3].to(torch.float64) + hits[:, 0:1].to(torch.float64) * d_in
    p = torch.where(hit[:, None], p_hit, surf)
    side = torch.where(hit, -torch.sign((n * d_in).sum(1)),
                       torch.where(torch.rand(R, generator=generator, device=dev) < 0.5, -1.0, 1.0)
                       .to(torch.float64))
    side = torch.where(side == 0, torch.ones_like(side), side)
    n = n * side[:, None]
    o = p + offset * n
    phi = 2 * np.pi * torch.rand(R, generator=generator, device=dev, dtype=torch.float64)
    r = torch.sqrt(torch.rand(R, generator=generator, device=dev, dtype=torch.float64))
    lx, ly = r * torch.cos(phi), r * torch.sin(phi)
    lz = torch.sqrt(torch.clamp(1 - r * r, min=0))
    t, b = _onb(n)
    d = lx[:, None] * t + ly[:, None] * b + lz[:, None] * n
    d = d / torch.linalg.norm(d, dim=1, keepdim=True)
    out = torch.empty((R, 8), dtype=torch.float32, device=dev)
    out[:, 0:3] = o.to(torch.float32)
    out[:, 3] = float(RAY_T_MIN)
    out[:, 4:7] = d.to(torch.float32)
    out[:, 7] = float(RAY_T_MAX)
    return out


def batch_scene_sizes(count: int = 256, seed: int = 7, lo: float = 1e3, hi: float = 5e5):
    """Triangle counts log-uniform in [lo, hi]: round(exp(U(ln lo, ln hi)))."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)


def batch_scenes(count: int = 256, seed: int = 7):
    """Scenes of uneven size for batched builds: even indices are noise-
    displaced spheres (about N triangles: 2*nu*nv with nu/nv = 1.25), odd
    ones random soups of exactly N triangles with edge 0.01, each in the unit
    cube and seeded by its index."""
    out = []
    for s, n in enumerate(batch_scene_sizes(count, seed)):
        if s % 2 == 0:
            nv = max(3, int(round(np.sqrt(n / 2.5))))
            nu = max(4, int(round(n / (2 * nv))))
            out.append(noisy_sphere(nu, nv, seed=seed * 1000 + s))
        else:
            out.append(random_soup(int(n), 0.01, seed=seed * 1000 + s))
    return out


def batch_rays_torch(lo, hi, counts, device, generator):
    """counts[s] rays for scene s: origin uniform in the scene's box [lo[s],
    hi[s]], direction uniform on the unit sphere; (R, 8) float32 records in
    scene order."""
    import torch

    counts_t = torch.as_tensor(np.asarray(counts, dtype=np.int64), device=device)
    scene = torch.repeat_interleave(torch.arange(len(counts), device=device), counts_t)
    lo_t = torch.as_tensor(np.asarray(lo, dtype=np.float32), device=device)
    hi_t = torch.as_tensor(np.asarray(hi, dtype=np.float32), device=device)
    n = scene.numel()
    out = torch.empty((n, 8), dtype=torch.float32, device=device)
    chunk = 1 << 24
    for b in range(0, n, chunk):
        e = min(n, b + chunk)
        sc = scene[b:e]
        u = torch.rand((e - b, 3), device=device, generator=generator)
        out[b:e, 0:3] = lo_t[sc] + u * (hi_t[sc] - lo_t[sc])
        d = torch.randn((e - b, 3), device=device, generator=generator)
        out[b:e, 4:7] = d / d.norm(dim=1, keepdim=True).clamp_min(1e-12)
        out[b:e, 3] = float(RAY_T_MIN)
        out[b:e, 7] = float(RAY_T_MAX)
    return out


def sphere_grid(nx: int = 64, nz: int = 64, nu: int = 61, nv: int = 40, seed: int = 11):
    """nx * nz noise-displaced spheres (2 * nu * nv triangles each, radius 0.35,
    amplitude 0.1) centred at (i + 0.5, 0.5, k + 0.5), each seeded by its
    index: 64 x 64 x 4,880 = 19,988,480 triangles by default."""
    out = np.empty((nx * nz, 2 * nu * nv, 9), dtype=np.float32)
    for i in range(nx):
        for k in range(nz):
            s = i * nz + k
            out[s] = noisy_sphere(nu, nv, center=(i + 0.5, 0.5, k + 0.5), seed=seed * 100003 + s)
    return out.reshape(-1, 9)


def animate_y_torch(base, frame: int, out):
    """Per-frame animation of a (N, 9) vertex tensor: y += 0.02 sin(2 pi (x +
    frame / 60)), written into `out` (same shape)."""
    import math

    import torch

    out.copy_(base)
    out[:, 1::3] += 0.02 * torch.sin(2.0 * math.pi * (base[:, 0::3] + frame / 60.0))
    return out


def triangle_aabbs_centers_torch(verts, aabbs_out, centers_out):
    """RTAabb records (N, 8) and centres (N, 3) of (N, 9) triangles on the
    device: boxes grown from the empty box by v0, v1, v2, centre
    ((v0 + v1) + v2) * (1 / 3)."""
    import torch

    v = verts.view(-1, 3, 3)
    aabbs_out[:, 0:3] = torch.minimum(torch.minimum(v[:, 0], v[:, 1]), v[:, 2])
    aabbs_out[:, 4:7] = torch.maximum(torch.maximum(v[:, 0], v[:, 1]), v[:, 2])
    aabbs_out[:, 3] = 0.0
    aabbs_out[:, 7] = 0.0
    centers_out.copy_(((v[:, 0] + v[:, 1]) + v[:, 2]) * np.float32(1.0 / 3.0))


def shadow_rays_torch(verts, n: int, light, device, generator):
    """n rays from uniformly chosen points on uniformly chosen triangles
    (offset 1e-4 along the normal facing the light) towards a point light,
    t_max = (1 - 1e-4) * distance."""
    import torch

    tri = torch.randint(0, verts.shape[0], (n,), device=device, generator=generator)
    out = torch.empty((n, 8), dtype=torch.float32, device=device)
    lt = torch.as_tensor(light, dtype=torch.float32, device=device)
    chunk = 1 << 24
    for b in range(0, n, chunk):
        e = min(n, b + chunk)
        v = verts[tri[b:e]].view(-1, 3, 3)
        r1 = torch.rand((e - b, 1), device=device, generator=generator).sqrt()
        r2 = torch.rand((e - b, 1), device=device, generator=generator)
        p = (1 - r1) * v[:, 0] + r1 * (1 - r2) * v[:, 1] + r1 * r2 * v[:, 2]
        nrm = torch.linalg.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
        nrm = nrm / nrm.norm(dim=1, keepdim=True).clamp_min(1e-20)
        to_l = lt - p
        nrm = torch.where(((nrm * to_l).sum(1, keepdim=True) < 0), -nrm, nrm)
        o = p + 1e-4 * nrm
        d = lt - o
        dist = d.norm(dim=1, keepdim=True).clamp_min(1e-20)
        out[b:e, 0:3] = o
        out[b:e, 3] = float(RAY_T_MIN)
        out[b:e, 4:7] = d / dist
        out[b:e, 7] = (1.0 - 1e-4) * dist[:, 0]
    return out
