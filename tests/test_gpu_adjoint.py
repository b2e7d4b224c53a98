"""Adjoint of the closest hit (rtbvh_gpu_adjoint, rtbvh_amd.diff.trace_tuv).

No reference counterpart (the reference's triangle test is forward only), so
the gradient is pinned by mathematics, not by reference outputs:
  * an FP64 CPU gradient computed a different way (numpy linear solves of
    M x = s and M^T lambda = g, M = [-d | e1 | e2]) against the kernel's
    closed form in FP32: relative error (norm of the difference over the norm
    of the FP64 gradient) <= 1e-4 for vertex, origin and direction
    gradients;
  * central finite differences of the FP64 forward with the hit triangles
    fixed, against the FP64 gradient: relative error <= 2e-2 at h = 1e-3
    (the truncation error is ~(h / edge)^2 with edges ~0.03) and <= 1e-5 at
    h = 1e-5;
  * central finite differences, h = 1e-3, of the GPU forward itself (full
    re-trace, shared vertex positions moved together so the mesh stays
    watertight), against the GPU gradient of L = sum of w_t t: relative
    error <= 5e-2 (FP32 forward, and rays near an edge may change triangle;
    t is continuous there, u and v are not, so they are left out).
Loss otherwise: L = sum over hit rays of w_t t + w_u u + w_v v, w ~ N(0, 1).
"""
import numpy as np
import pytest

from rtbvh_amd import scenes

torch = pytest.importorskip("torch")
import rtbvh_amd as rt  # noqa: E402
from rtbvh_amd.diff import trace_tuv  # noqa: E402

pytestmark = pytest.mark.gpu


def _tree(verts, dev):
    aabbs, centers = scenes.triangle_aabbs_centers(verts)
    bvh = rt.Builder(torch.from_numpy(centers).to(dev),
                     aabbs=torch.from_numpy(aabbs).to(dev)).construct_locally_ordered_clustered()
    return rt.Mbvh.from_bvh(bvh)


def _setup(seed=3, n_rays=20000):
    # rays from a sphere of radius 3 aimed at the middle of the displaced
    # sphere (radius 0.35 +- 0.1): no grazing hits, no silhouette crossings
    verts = scenes.noisy_sphere(60, 40, seed=5).astype(np.float32)
    rays = scenes.rays_sphere_to_box(n_rays, (0.42, 0.42, 0.42), (0.58, 0.58, 0.58), seed=seed)
    dev = torch.device("cuda:0")
    tree = _tree(verts, dev)
    rng = np.random.default_rng(seed)
    w = rng.normal(size=(n_rays, 3))
    return verts, rays, tree, w, dev


def _solve_f64(verts64, o, d, prim):
    """(t, u, v) and M for the given triangles, by a float64 linear solve."""
    tri = verts64[prim].reshape(-1, 3, 3)
    v0, v1, v2 = tri[:, 0], tri[:, 1], tri[:, 2]
    m = np.stack([-d, v1 - v0, v2 - v0], axis=-1)
    x = np.linalg.solve(m, (o - v0)[..., None])[..., 0]
    return x, m


def _grad_f64(verts64, o, d, prim, g):
    x, m = _solve_f64(verts64, o, d, prim)
    lam = np.linalg.solve(np.transpose(m, (0, 2, 1)), g[..., None])[..., 0]
    t, u, v = x[:, 0], x[:, 1], x[:, 2]
    gv = np.zeros((len(verts64), 9))
    np.add.at(gv[:, 0:3], prim, -(1.0 - u - v)[:, None] * lam)
    np.add.at(gv[:, 3:6], prim, -u[:, None] * lam)
    np.add.at(gv[:, 6:9], prim, -v[:, None] * lam)
    return gv, lam, t[:, None] * lam, x


def _loss_f64(verts64, o, d, prim, w):
    x, _ = _solve_f64(verts64, o, d, prim)
    return float(np.sum(w * x))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _forward_backward(verts, rays, tree, w, dev):
    vt = torch.from_numpy(verts).to(dev).requires_grad_(True)
    o = torch.from_numpy(np.ascontiguousarray(rays[:, 0:3])).to(dev).requires_grad_(True)
    d = torch.from_numpy(np.ascontiguousarray(rays[:, 4:7])).to(dev).requires_grad_(True)
    t, u, v, prim = trace_tuv(tree, vt, o, d)
    hit = prim >= 0
    wt = torch.from_numpy(w).to(dev, torch.float32)
    loss = (wt[:, 0] * torch.where(hit, t, 0) + wt[:, 1] * u + wt[:, 2] * v)[hit].sum()
    loss.backward()
    torch.cuda.synchronize()
    return (t.detach().cpu().numpy(), u.detach().cpu().numpy(), v.detach().cpu().numpy(),
            prim.cpu().numpy(), vt.grad.cpu().numpy(), o.grad.cpu().numpy(), d.grad.cpu().numpy())


def test_adjoint_matches_fp64_gradient():
    print("[adjoint_fp64]", flush=True)
    verts, rays, tree, w, dev = _setup()
    t, u, v, prim, g_v, g_o, g_d = _forward_backward(verts, rays, tree, w, dev)
    hit = prim >= 0
    assert hit.mean() > 0.5
    verts64 = verts.astype(np.float64)
    o64 = rays[hit, 0:3].astype(np.float64)
    d64 = rays[hit, 4:7].astype(np.float64)
    gv, go, gd, x = _grad_f64(verts64, o64, d64, prim[hit], w[hit])
    # forward agrees with the float64 solve
    assert _rel(np.stack([t[hit], u[hit], v[hit]], 1), x) < 1e-5
    errs = (_rel(g_v, gv), _rel(g_o[hit], go), _rel(g_d[hit], gd))
    print(f"[adjoint_fp64] relative error vertices {errs[0]:.2e} origin {errs[1]:.2e} "
          f"direction {errs[2]:.2e}", flush=True)
    assert max(errs) <= 1e-4, errs
    assert not g_o[~hit].any() and not g_d[~hit].any()


def test_fp64_gradient_matches_central_differences():
    print("[adjoint_fd64]", flush=True)
    verts, rays, tree, w, dev = _setup(seed=4, n_rays=5000)
    _, _, _, prim, _, _, _ = _forward_backward(verts, rays, tree, w, dev)
    hit = prim >= 0
    verts64 = verts.astype(np.float64)
    o64 = rays[hit, 0:3].astype(np.float64)
    d64 = rays[hit, 4:7].astype(np.float64)
    p, wh = prim[hit], w[hit]
    gv, go, gd, _ = _grad_f64(verts64, o64, d64, p, wh)
    # h = 1e-3: truncation error ~ (h / edge)^2 with edges ~0.03, so ~1e-2;
    # h = 1e-5 shrinks it to ~1e-7 and pins the formula tightly
    for h, tol in ((1e-3, 2e-2), (1e-5, 1e-5)):
        rng = np.random.default_rng(0)
        fd, an = [], []
        for tri in rng.choice(np.unique(p), 12, replace=False):
            for c in rng.choice(9, 3, replace=False):
                vp, vm = verts64.copy(), verts64.copy()
                vp[tri, c] += h
                vm[tri, c] -= h
                fd.append((_loss_f64(vp, o64, d64, p, wh) - _loss_f64(vm, o64, d64, p, wh))
                          / (2 * h))
                an.append(gv[tri, c])
        for r in rng.choice(len(p), 12, replace=False):
            for arr, g in ((o64, go), (d64, gd)):
                c = int(rng.integers(3))
                ap, am = arr.copy(), arr.copy()
                ap[r, c] += h
                am[r, c] -= h
                if arr is o64:
                    lp, lm = _loss_f64(verts64, ap, d64, p, wh), _loss_f64(verts64, am, d64, p, wh)
                else:
                    lp, lm = _loss_f64(verts64, o64, ap, p, wh), _loss_f64(verts64, o64, am, p, wh)
                fd.append((lp - lm) / (2 * h))
                an.append(g[r, c])
        err = _rel(np.asarray(an), np.asarray(fd))
        print(f"[adjoint_fd64] h={h:g}: relative error {err:.2e} over {len(fd)} coordinates",
              flush=True)
        assert err <= tol, (h, err)


def test_gpu_gradient_matches_gpu_central_differences():
    print("[adjoint_fd_gpu]", flush=True)
    verts, rays, tree, w, dev = _setup(seed=5, n_rays=20000)
    # t alone: with shared positions moved together t stays continuous, while
    # u and v jump where a ray changes triangle
    w[:, 1:] = 0.0
    _, _, _, prim, g_v, _, _ = _forward_backward(verts, rays, tree, w, dev)
    corners = verts.reshape(-1, 3)
    keys, inverse = np.unique(corners, axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)
    rays_d = torch.from_numpy(rays).to(dev)
    hit0 = prim >= 0

    def loss(vs):
        # the closest hit under the lowest-(t, id) rule does not depend on the
        # tree, so each perturbed mesh gets a tree of its own
        vs = np.ascontiguousarray(vs)
        vd = torch.from_numpy(vs).to(dev)
        hits = _tree(vs, dev).intersect(vd, rays_d).cpu().numpy()
        t, u, v = (hits[:, k].astype(np.float64) for k in range(3))
        hit = hits[:, 3].view(np.uint32) != 0xFFFFFFFF
        return w[:, 0] * t + w[:, 1] * u + w[:, 2] * v, hit

    rng = np.random.default_rng(1)
    used = np.unique(prim[hit0])
    h = 1e-3
    fd, an = [], []
    for tri in rng.choice(used, 10, replace=False):
        corner = int(tri) * 3 + int(rng.integers(3))
        group = np.nonzero(inverse == inverse[corner])[0]  # every copy of this position
        c = int(rng.integers(3))
        vp, vm = corners.copy(), corners.copy()
        vp[group, c] += np.float32(h)
        vm[group, c] -= np.float32(h)
        step = float(vp[group[0], c]) - float(vm[group[0], c])  # 2h after fp32 rounding
        lp, hp = loss(vp.reshape(-1, 9))
        lm, hm = loss(vm.reshape(-1, 9))
        assert np.array_equal(hp, hit0) and np.array_equal(hm, hit0)
        fd.append((lp[hit0].sum() - lm[hit0].sum()) / step)
        an.append(float(sum(g_v.reshape(-1, 3)[k, c] for k in group)))
    err = _rel(np.asarray(an), np.asarray(fd))
    print(f"[adjoint_fd_gpu] relative error {err:.2e} over {len(fd)} shared vertices", flush=True)
    assert err <= 5e-2


def test_trace_tree_snapshot_guard():
    """A TraceTree traces the snapshot of the vertices it was built from;
    trace_tuv refuses vertices that are another tensor or were modified in
    place since, because its gradients would belong to other geometry."""
    print("[adjoint_snapshot_guard]", flush=True)
    verts_np = scenes.noisy_sphere(60, 40, seed=5).astype(np.float32)
    rays = scenes.rays_sphere_to_box(2000, (0.42, 0.42, 0.42), (0.58, 0.58, 0.58), seed=6)
    dev = torch.device("cuda:0")
    aabbs, centers = scenes.triangle_aabbs_centers(verts_np)
    verts = torch.from_numpy(verts_np).to(dev)
    bvh = rt.Builder(torch.from_numpy(centers).to(dev),
                     aabbs=torch.from_numpy(aabbs).to(dev)).construct_locally_ordered_clustered()
    tree = bvh.trace_tree(verts)
    o = torch.from_numpy(np.ascontiguousarray(rays[:, 0:3])).to(dev)
    d = torch.from_numpy(np.ascontiguousarray(rays[:, 4:7])).to(dev)
    t, _, _, prim = trace_tuv(tree, verts.detach().requires_grad_(True), o, d)
    assert (prim >= 0).float().mean().item() > 0.5
    with pytest.raises(ValueError):
        trace_tuv(tree, verts.clone(), o, d)  # equal values, another tensor
    verts[:, 1] += 0.01                       # moved in place after the build
    with pytest.raises(ValueError):
        trace_tuv(tree, verts, o, d)
    # a tree rebuilt over the moved vertices is accepted
    a2, c2 = scenes.triangle_aabbs_centers(verts.cpu().numpy())
    bvh2 = rt.Builder(torch.from_numpy(c2).to(dev),
                      aabbs=torch.from_numpy(a2).to(dev)).construct_locally_ordered_clustered()
    tree2 = bvh2.trace_tree(verts)
    t2, _, _, _ = trace_tuv(tree2, verts, o, d)
    assert not torch.equal(t, t2)
