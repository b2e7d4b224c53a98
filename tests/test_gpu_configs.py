"""GPU parity at the benchmark workloads, against the oracle.

Each test runs the product path of one benchmark configuration at its real
size and checks its output against the CPU oracle (never against another GPU
path):
  * 1M triangles (noisy sphere + soup, GPU PLOC): the compressed 8-wide
    traversal tree's closest hits of 2.36M cosine-bounce rays (the headline
    kernel, in the launch shape the benchmark's autotuner picks and in the
    default shape) and of the primary rays, bit for bit against the oracle's
    lowest-(t, id) MBVH traversal; any-hit flags likewise; and the MBVH
    traversal's node and triangle counters against the oracle's count of the
    same traversal;
  * 256 scenes of 1k-500k triangles (log-uniform) with 16 rays per triangle,
    built and traced as one recorded graph: sampled scenes' trees node for node
    equal to the oracle's LOCB, their rays' hits equal to the oracle's
    lowest-id traversal;
  * 100k triangles, 2^24 rays, L = sum w t: the adjoint's gradients against
    a float64 reverse-mode gradient on 2^18 rays, and central differences
    (h = 1e-3) of the GPU forward against the adjoint kernel on 2^20 rays,
    hit triangles fixed as the adjoint fixes them;
  * 20M triangles animated, one frame as one recorded graph (PLOC rebuild +
    traversal tree + 2^28 any-hit shadow rays): the frame's build bit-identical
    to the oracle's LOCB, and the shadow flags of a 400k-ray sample equal to
    the oracle's any-hit traversal; then the same frame with the rays ordered
    by the cells of their origins inside the graph: all 2^28 flags equal.
"""
import os

import numpy as np
import pytest

from oracle import refcpu as ref
from rtbvh_amd import scenes
from tests.helpers import flat, hits_equal

torch = pytest.importorskip("torch")
import rtbvh_amd as rt  # noqa: E402

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
# the launch shape the benchmark's autotuner picked on the bounce rays
BENCH_SHAPE = dict(block=128, lds_stack=8, occupancy_pct=1600, kernel=0, refill=16,
                   schedule=1, postpone=0)


def _np_rays(t):
    return t.cpu().numpy().view(ref.RAY).reshape(-1)


def _np_hits(t):
    return t.cpu().numpy().view(ref.HIT).reshape(-1)


@pytest.fixture(scope="module")
def scene_1m():
    print("[c2] building the 1M-triangle scene", flush=True)
    dev = torch.device("cuda:0")
    verts = scenes.scene_1m()
    aabbs, centers = scenes.triangle_aabbs_centers(verts)
    v_d = torch.from_numpy(verts).to(dev)
    bvh = rt.Builder(torch.from_numpy(centers).to(dev),
                     aabbs=torch.from_numpy(aabbs).to(dev)).construct_locally_ordered_clustered()
    nodes, idx = ref.build_locb(aabbs.view(ref.AABB).reshape(-1), centers)
    got = bvh.nodes().cpu().numpy().view(ref.AABB).reshape(-1)
    assert np.array_equal(flat(got).view(np.uint32), flat(nodes).view(np.uint32))
    assert np.array_equal(bvh.indices().cpu().numpy().view(np.uint32), idx)
    m = ref.build_mbvh(nodes)
    tree = bvh.trace_tree(v_d, width=8)
    # rays as in the benchmark, from a 1536^2 image instead of 8192^2
    w = 1536
    primary = scenes.pinhole_rays_torch(w, w, dev)
    hits_p = tree.intersect(primary)
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    bounce = scenes.bounce_rays_torch(v_d, primary, hits_p, gen)
    torch.cuda.synchronize()
    return dict(verts=verts, v_d=v_d, bvh=bvh, nodes=nodes, idx=idx, m=m, tree=tree,
                primary=primary, bounce=bounce, aabbs=aabbs, centers=centers)


@pytest.mark.parametrize("ppl", [1, 4])
def test_c2_binned_sah_matches_oracle(scene_1m, ppl):
    """The binned SAH builder at 1M triangles (the benchmark's scene; its
    levels hold up to 300k tasks, so every tile look-back, bin-slot copy and
    straddling partition tile is exercised): the same topology, leaf sets and
    bounds as the oracle's builder with the reference's two box defects
    fixed, and a valid tree."""
    s = scene_1m
    print(f"[c2] binned SAH at 1M, primitives per leaf {ppl}", flush=True)
    dev = torch.device("cuda:0")
    bvh = rt.Builder(torch.from_numpy(s["centers"]).to(dev),
                     aabbs=torch.from_numpy(s["aabbs"]).to(dev),
                     primitives_per_leaf=ppl).construct_binned_sah()
    got_n = bvh.nodes().cpu().numpy().view(ref.AABB).reshape(-1)
    got_i = bvh.indices().cpu().numpy().view(np.uint32)
    want_n, want_i = ref.build_binned_sah(s["aabbs"].view(ref.AABB).reshape(-1), s["centers"],
                                          prims_per_leaf=ppl, contained=True)
    assert len(got_n) == len(want_n)
    assert ref.compare_bvh(got_n, got_i, want_n, want_i, as_sets=True) == 0
    code, viol, _ = ref.check_bvh(got_n, got_i, len(s["verts"]))
    assert code == 0 and viol == 0


def test_c2_compressed_tree_matches_oracle(scene_1m):
    s = scene_1m
    old = rt.get_trace_config()
    try:
        for name, rays in (("bounce", s["bounce"]), ("primary", s["primary"])):
            want = ref.trace_mbvh(s["m"], s["idx"], s["verts"], _np_rays(rays), threads=THREADS)
            assert (want["prim"] != 0xFFFFFFFF).mean() > 0.3
            for shape in (BENCH_SHAPE, None):
                print(f"[c2] {name} rays={rays.shape[0]} shape={shape}", flush=True)
                if shape is None:
                    rt.set_trace_config()  # the library's default shape
                else:
                    rt.set_trace_config(**shape)
                got = _np_hits(s["tree"].intersect(rays))
                s["tree"].status()
                bad = ~hits_equal(got, want)
                assert not bad.any(), f"{name}: {bad.sum()} rays differ, first {np.nonzero(bad)[0][:5]}"
        # any hit: t_max at random fractions of the closest distance and beyond
        r = s["bounce"].clone()
        g = torch.Generator(device=r.device)
        g.manual_seed(3)
        r[:, 7] = torch.rand(r.shape[0], device=r.device, generator=g) * 0.3
        rt.set_trace_config(**BENCH_SHAPE)
        got = s["tree"].occluded(r).cpu().numpy()
        want = ref.occluded_mbvh(s["m"], s["idx"], s["verts"], _np_rays(r), threads=THREADS)
        assert 0.05 < want.mean() < 0.95
        assert np.array_equal(got, want)
    finally:
        rt.set_trace_config(**old)


def test_c2_work_counters_match_oracle(scene_1m):
    """The per-ray node and triangle counts behind the traversal's bytes: the
    GPU MBVH traversal's counters equal the oracle's count of the same
    traversal (same visit order and pruning), ray for ray summed, on the
    oracle's own MBVH of the same tree."""
    s = scene_1m
    print("[c2] work counters", flush=True)
    m_d = rt.Mbvh.from_arrays(torch.from_numpy(flat(s["m"])).cuda(),
                              torch.from_numpy(s["idx"].view(np.int32)).cuda())
    rays = s["bounce"]
    import ctypes

    from rtbvh_amd import _lib

    counters = torch.zeros(2, dtype=torch.int64, device=rays.device)
    out = torch.empty((rays.shape[0], 4), dtype=torch.float32, device=rays.device)
    _lib.check(_lib.lib().rtbvh_gpu_intersect_mbvh_stats(
        m_d._h, ctypes.c_void_p(s["v_d"].data_ptr()), len(s["verts"]),
        ctypes.c_void_p(rays.data_ptr()), rays.shape[0], ctypes.c_void_p(out.data_ptr()), 0,
        ctypes.c_void_p(counters.data_ptr()), None), "stats")
    torch.cuda.synchronize()
    want, stats = ref.trace_mbvh(s["m"], s["idx"], s["verts"], _np_rays(rays), threads=THREADS,
                                 return_stats=True)
    assert hits_equal(_np_hits(out), want).all()
    got = counters.cpu().numpy().astype(np.uint64)
    print(f"[c2] nodes/ray {got[0] / rays.shape[0]:.3f} tris/ray {got[1] / rays.shape[0]:.3f}",
          flush=True)
    assert got[0] == stats[0] and got[1] == stats[1]


def test_c2_reference_order_traversals_match_oracle(scene_1m):
    """The reference's own benchmark loop at the headline size
    (examples/benchmark.rs:15-41: the MBVH and BVH iterators in their visit
    order, the first accepted candidate at the lowest t) on the GPU over the
    1M tree, against the oracle's restatement of the same loop, on every
    ninth bounce ray (262,144 rays), bit for bit."""
    s = scene_1m
    rays = s["bounce"][::9].contiguous()
    print(f"[c2] reference-order traversals, {rays.shape[0]} rays", flush=True)
    dev = rays.device
    idx_d = torch.from_numpy(s["idx"].view(np.int32)).to(dev)
    mbvh = rt.Mbvh.from_arrays(torch.from_numpy(flat(s["m"])).to(dev), idx_d)
    bvh = rt.Bvh.from_arrays(torch.from_numpy(flat(s["nodes"])).to(dev), idx_d)
    np_rays = _np_rays(rays)
    for name, tree, want in (
            ("mbvh", mbvh, ref.trace_mbvh(s["m"], s["idx"], s["verts"], np_rays,
                                          mode=ref.MODE_VISIT_ORDER, threads=THREADS)),
            ("bvh", bvh, ref.trace_bvh(s["nodes"], s["idx"], s["verts"], np_rays,
                                       mode=ref.MODE_VISIT_ORDER, threads=THREADS))):
        got = _np_hits(tree.intersect(s["v_d"], rays, reference_order=True))
        tree.status()
        bad = ~hits_equal(got, want)
        assert not bad.any(), f"{name}: {bad.sum()} rays differ, first {np.nonzero(bad)[0][:5]}"
        assert (want["prim"] != 0xFFFFFFFF).mean() > 0.3


def _subtree_equal(na, ia, ra, off_a, nb, ib, rb):
    """Node-for-node equality of the subtree of (na, ia) at ra, whose leaves
    hold batch-wide ids offset by off_a, and the tree (nb, ib) at rb."""
    ua = flat(na).view(np.uint32).reshape(-1, 8)
    ub = flat(nb).view(np.uint32).reshape(-1, 8)
    stack = [(ra, rb)]
    count = 0
    while stack:
        a, b = stack.pop()
        count += 1
        if not np.array_equal(ua[a][[0, 1, 2, 4, 5, 6]], ub[b][[0, 1, 2, 4, 5, 6]]):
            return False
        ca, cb = int(na[a]["count"]), int(nb[b]["count"])
        if ca != cb:
            return False
        la, lb = int(na[a]["left_first"]), int(nb[b]["left_first"])
        if ca >= 0:
            if not np.array_equal(ia[la:la + ca].astype(np.int64) - off_a,
                                  ib[lb:lb + cb].astype(np.int64)):
                return False
        else:
            stack.append((la + 1, lb + 1))
            stack.append((la, lb))
    return count


def test_c3_batch_of_256_scenes_matches_oracle():
    print("[c3] generating 256 scenes", flush=True)
    dev = torch.device("cuda:0")
    vl = scenes.batch_scenes(256, seed=7)
    counts = [len(v) for v in vl]
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    verts = np.concatenate(vl)
    aabbs, centers = scenes.triangle_aabbs_centers(verts)
    a8 = aabbs.view(np.float32).reshape(-1, 8)
    box_lo = np.stack([a8[offs[s]:offs[s + 1], 0:3].min(0) for s in range(256)])
    box_hi = np.stack([a8[offs[s]:offs[s + 1], 4:7].max(0) for s in range(256)])
    v_d = torch.from_numpy(verts).to(dev)
    a_d = torch.from_numpy(aabbs).to(dev)
    c_d = torch.from_numpy(centers).to(dev)
    ray_counts = [16 * c for c in counts]
    roffs = np.concatenate([[0], np.cumsum(ray_counts)]).astype(np.int64)
    gen = torch.Generator(device=dev)
    gen.manual_seed(8)
    rays = scenes.batch_rays_torch(box_lo, box_hi, ray_counts, dev, gen)
    roff_d = rt.SceneBatch.ray_offsets(ray_counts, dev)
    print(f"[c3] {len(verts)} triangles, {rays.shape[0]} rays: graph launch", flush=True)
    plan = rt.BatchPlan(c_d, offs.tolist(), v_d, aabbs=a_d)
    hits = torch.empty((rays.shape[0], 4), dtype=torch.float32, device=dev)
    for _ in range(2):  # the first call records the graph, the second replays it
        plan.launch(c_d, v_d, rays, roff_d, aabbs=a_d, out=hits)
    plan.status()
    batch = rt.SceneBatch(c_d, offs.tolist(), aabbs=a_d)
    bnodes = batch.bvh_nodes().cpu().numpy().view(ref.AABB).reshape(-1)
    bidx = batch.indices().cpu().numpy().view(np.uint32)
    rng = np.random.default_rng(9)
    sample = sorted({int(np.argmin(counts)), int(np.argmax(counts))}
                    | set(rng.choice(256, 6, replace=False).tolist()))
    n_hits = 0
    for s in sample:
        p0, p1 = int(offs[s]), int(offs[s + 1])
        print(f"[c3] scene {s}: {p1 - p0} triangles", flush=True)
        nodes, idx = ref.build_locb(aabbs[p0:p1].view(ref.AABB).reshape(-1), centers[p0:p1])
        assert _subtree_equal(bnodes, bidx, s, p0, nodes, idx, 0) == 2 * (p1 - p0) - 1
        m = ref.build_mbvh(nodes)
        r0, r1 = int(roffs[s]), int(roffs[s + 1])
        pick = np.unique(np.linspace(r0, r1 - 1, min(r1 - r0, 150_000)).astype(np.int64))
        pick_d = torch.from_numpy(pick).to(dev)
        want = ref.trace_mbvh(m, idx, vl[s], _np_rays(rays[pick_d]), threads=THREADS)
        got = _np_hits(hits[pick_d])
        bad = ~hits_equal(got, want)
        assert not bad.any(), f"scene {s}: {bad.sum()} of {len(pick)} rays differ"
        n_hits += int((want["prim"] != 0xFFFFFFFF).sum())
    # soups of 0.01-edge triangles are hit by few of the rays, spheres by most
    assert n_hits > 100_000


def test_c4_adjoint_at_100k_triangles():
    from rtbvh_amd.diff import trace_tuv

    print("[c4] 100k-triangle sphere, 2^24 rays", flush=True)
    dev = torch.device("cuda:0")
    verts_np = scenes.noisy_sphere(250, 200, seed=7)
    assert len(verts_np) == 100_000
    aabbs, centers = scenes.triangle_aabbs_centers(verts_np)
    verts = torch.from_numpy(verts_np).to(dev)

    def tree_of(v_np, v_dev=None):
        a, c = scenes.triangle_aabbs_centers(v_np)
        b = rt.Builder(torch.from_numpy(c).to(dev),
                       aabbs=torch.from_numpy(a).to(dev)).construct_locally_ordered_clustered()
        if v_dev is None:
            v_dev = torch.from_numpy(np.ascontiguousarray(v_np)).to(dev)
        return b.trace_tree(v_dev)

    tree = tree_of(verts_np, verts)
    n = 1 << 24
    gen = torch.Generator(device=dev)
    gen.manual_seed(10)
    o = torch.randn((n, 3), device=dev, generator=gen)
    o = (3.0 * o / o.norm(dim=1, keepdim=True)).contiguous()
    tgt = 0.25 + 0.5 * torch.rand((n, 3), device=dev, generator=gen)
    d = tgt - o
    d = (d / d.norm(dim=1, keepdim=True)).contiguous()
    w = torch.randn(n, device=dev, generator=gen)

    def run(k):
        vt = verts.detach().requires_grad_(True)
        oo = o[:k].detach().requires_grad_(True)
        dd = d[:k].detach().requires_grad_(True)
        t, u, v, prim = trace_tuv(tree, vt, oo, dd)
        (w[:k] * torch.where(prim >= 0, t, torch.zeros_like(t))).sum().backward()
        torch.cuda.synchronize()
        return t.detach(), prim, vt.grad, oo.grad, dd.grad

    # the whole workload runs; ray gradients of a sample and the vertex
    # gradient of the first 2^18 rays are checked against float64
    t, prim, g_v, g_o, g_d = run(n)
    assert (prim >= 0).float().mean().item() > 0.9
    k = 1 << 18
    sel = torch.arange(0, n, n // k, device=dev)[:k]
    p_ = prim[sel].cpu().numpy()
    hit = p_ >= 0
    v64 = verts_np.astype(np.float64)[p_[hit]].reshape(-1, 3, 3)
    o64 = o[sel].cpu().numpy().astype(np.float64)[hit]
    d64 = d[sel].cpu().numpy().astype(np.float64)[hit]
    m = np.stack([-d64, v64[:, 1] - v64[:, 0], v64[:, 2] - v64[:, 0]], axis=-1)
    x = np.linalg.solve(m, (o64 - v64[:, 0])[..., None])[..., 0]
    g = np.zeros_like(x)
    g[:, 0] = w[sel].cpu().numpy().astype(np.float64)[hit]
    lam = np.linalg.solve(np.transpose(m, (0, 2, 1)), g[..., None])[..., 0]

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))

    err_o = rel(g_o[sel].cpu().numpy()[hit], lam)
    err_d = rel(g_d[sel].cpu().numpy()[hit], x[:, 0:1] * lam)
    err_t = rel(t[sel].cpu().numpy()[hit], x[:, 0])
    _, prim_k, g_vk, _, _ = run(k)
    pk = prim_k.cpu().numpy()
    hk = pk >= 0
    vk = verts_np.astype(np.float64)[pk[hk]].reshape(-1, 3, 3)
    ok_ = o[:k].cpu().numpy().astype(np.float64)[hk]
    dk = d[:k].cpu().numpy().astype(np.float64)[hk]
    mk = np.stack([-dk, vk[:, 1] - vk[:, 0], vk[:, 2] - vk[:, 0]], axis=-1)
    xk = np.linalg.solve(mk, (ok_ - vk[:, 0])[..., None])[..., 0]
    gk = np.zeros_like(xk)
    gk[:, 0] = w[:k].cpu().numpy().astype(np.float64)[hk]
    lk = np.linalg.solve(np.transpose(mk, (0, 2, 1)), gk[..., None])[..., 0]
    gv = np.zeros((len(verts_np), 9))
    np.add.at(gv[:, 0:3], pk[hk], -(1.0 - xk[:, 1] - xk[:, 2])[:, None] * lk)
    np.add.at(gv[:, 3:6], pk[hk], -xk[:, 1:2] * lk)
    np.add.at(gv[:, 6:9], pk[hk], -xk[:, 2:3] * lk)
    err_v = rel(g_vk.cpu().numpy(), gv)
    print(f"[c4] float64 relative errors: t {err_t:.2e} vertices {err_v:.2e} origins {err_o:.2e} "
          f"directions {err_d:.2e}", flush=True)
    assert err_t <= 1e-5 and err_o <= 1e-4 and err_d <= 1e-4 and err_v <= 2e-4

    # central differences of the GPU forward (fresh tree per mesh) on 2^20
    # rays against the adjoint kernel, with the topology fixed as the adjoint
    # fixes it: every copy of a vertex position moves together (the mesh stays
    # watertight), and only rays that hit the same triangle at -h, 0 and +h
    # carry weight on both sides (at h = 1e-3, 11-18 % of an edge, many rays
    # near the moved vertex cross to a neighbour, where t has a kink).
    # Grazing hits (|cos| < 0.3 to the hit triangle's normal) also get weight
    # 0: their t is the most curved in the vertex position.
    from rtbvh_amd.diff import adjoint

    kf = 1 << 20
    rays_f = rt.rays_tensor(o[:kf], d[:kf])
    h0_d = tree.intersect(rays_f)
    h0 = h0_d.cpu().numpy()
    p0 = h0[:, 3].view(np.int32)
    hit0 = p0 >= 0
    tri = verts_np.astype(np.float64)[np.where(hit0, p0, 0)].reshape(-1, 3, 3)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    cos = np.abs((nrm * d[:kf].cpu().numpy().astype(np.float64)).sum(1))
    base_w = np.where(hit0 & (cos >= 0.3), w[:kf].cpu().numpy().astype(np.float64), 0.0)
    corners = verts_np.reshape(-1, 3)
    _, inverse = np.unique(corners, axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)

    def trace(vs):
        hh = tree_of(vs.reshape(-1, 9)).intersect(rays_f).cpu().numpy()
        return hh[:, 0].astype(np.float64), hh[:, 3].view(np.int32)

    rng = np.random.default_rng(11)
    used = np.unique(p0[base_w != 0])
    h = 1e-3
    fd, an, kept, moved = [], [], [], []
    for tri_id in rng.choice(used, 12, replace=False):
        corner = int(tri_id) * 3 + int(rng.integers(3))
        group = np.nonzero(inverse == inverse[corner])[0]
        c = int(rng.integers(3))
        vp, vm = corners.copy(), corners.copy()
        vp[group, c] += np.float32(h)
        vm[group, c] -= np.float32(h)
        step = float(vp[group[0], c]) - float(vm[group[0], c])
        tp, pp = trace(vp)
        tm, pm = trace(vm)
        keep = (base_w != 0) & (pp == p0) & (pm == p0)
        wk = np.where(keep, base_w, 0.0)
        fd.append(float(np.sum(wk * (tp - tm))) / step)
        g_tuv = torch.zeros((kf, 3), dtype=torch.float32, device=dev)
        g_tuv[:, 0] = torch.from_numpy(wk.astype(np.float32)).to(dev)
        gv, _, _ = adjoint(verts, rays_f, h0_d, g_tuv, grad_rays=False)
        gv = gv.reshape(-1, 3).cpu().numpy().astype(np.float64)
        an.append(float(gv[group, c].sum()))
        touching = np.isin(p0, group // 3) & (base_w != 0)
        kept.append(int((keep & touching).sum()))
        moved.append(int(touching.sum()))
    err_fd = rel(np.asarray(an), np.asarray(fd))
    print(f"[c4] central differences (h = 1e-3) vs GPU adjoint: relative error {err_fd:.2e}; "
          f"rays on the moved triangles kept {sum(kept)} of {sum(moved)}", flush=True)
    assert sum(kept) >= 0.5 * sum(moved) and sum(kept) > 200
    assert err_fd <= 5e-2


def test_c5_animated_20m_frame_matches_oracle():
    print("[c5] generating the 20M-triangle sphere grid", flush=True)
    dev = torch.device("cuda:0")
    base = torch.from_numpy(scenes.sphere_grid()).to(dev)
    n = base.shape[0]
    verts = torch.empty_like(base)
    aabbs = torch.empty((n, 8), dtype=torch.float32, device=dev)
    centers = torch.empty((n, 3), dtype=torch.float32, device=dev)
    scenes.animate_y_torch(base, 0, verts)
    scenes.triangle_aabbs_centers_torch(verts, aabbs, centers)
    gen = torch.Generator(device=dev)
    gen.manual_seed(12)
    n_rays = 1 << 28
    light = (32.0, 20.0, 32.0)
    rays = scenes.shadow_rays_torch(verts, n_rays, light, dev, gen)
    occ = torch.empty((n_rays,), dtype=torch.uint8, device=dev)
    plan = rt.BatchPlan(centers, [0, n], verts, aabbs=aabbs)
    pick = torch.from_numpy(np.linspace(0, n_rays - 1, 400_000).astype(np.int64)).to(dev)
    for frame in (1, 2):
        print(f"[c5] frame {frame}: animate, graph launch", flush=True)
        scenes.animate_y_torch(base, frame, verts)
        scenes.triangle_aabbs_centers_torch(verts, aabbs, centers)
        plan.launch(centers, verts, rays, None, aabbs=aabbs, any_hit=True, out=occ)
        plan.status()
        # the oracle over the same inputs: LOCB (bit-identical to the GPU's
        # eager build of this frame), its MBVH, any hit of the sampled rays
        a_np = aabbs.cpu().numpy().view(ref.AABB).reshape(-1)
        c_np = centers.cpu().numpy()
        v_np = verts.cpu().numpy()
        print(f"[c5] frame {frame}: oracle build", flush=True)
        nodes, idx = ref.build_locb(a_np, c_np)
        if frame == 1:
            bvh = rt.Builder(centers, aabbs=aabbs).construct_locally_ordered_clustered()
            got = bvh.nodes().cpu().numpy().view(ref.AABB).reshape(-1)
            assert np.array_equal(flat(got).view(np.uint32), flat(nodes).view(np.uint32))
            del bvh, got
        m = ref.build_mbvh(nodes)
        want = ref.occluded_mbvh(m, idx, v_np, _np_rays(rays[pick]), threads=THREADS)
        got = occ[pick].cpu().numpy()
        print(f"[c5] frame {frame}: occluded fraction {want.mean():.3f}", flush=True)
        assert 0.2 < want.mean() < 0.95
        assert np.array_equal(got, want), f"frame {frame}: {(got != want).sum()} flags differ"
        del nodes, idx, m
    # frame 2 again, the rays ordered by origin cell inside the graph
    print("[c5] frame 2, ordered rays", flush=True)
    plan.order_rays(n_rays)
    occ_ordered = torch.empty_like(occ)
    plan.launch(centers, verts, rays, None, aabbs=aabbs, any_hit=True, out=occ_ordered)
    plan.status()
    torch.cuda.synchronize()
    assert torch.equal(occ_ordered, occ)
