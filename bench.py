#!/usr/bin/env python3
"""Benchmark of the GPU hot path on one MI355X.

Headline workload (1M-triangle scene):
  1,000,000 FP32 triangles — a 750x600-quad noisy sphere (900k) plus 100k
  random triangles (edge 0.01) in the unit cube; GPU PLOC build (search
  radius 14), the reference-layout MBVH collapse and the compressed 8-wide
  traversal tree; 8192^2 = 67,108,864 pinhole primary rays and the same
  number of cosine-weighted bounce rays from the primary hits.

A step is one closest-hit launch over the 67,108,864 bounce rays (the
incoherent set), inputs resident in HBM, lowest-(t, id) rule (hit records
identical to brute force; a sample is checked against the CPU oracle).  The
headline value is bounce-ray Mrays/s of the traversal tree; the JSON line
also carries the build times (PLOC, collapse, traversal tree, binned SAH),
the primary-ray rate, the 4-wide traversal tree's and the reference-layout
MBVH's rates on the same rays (same hits), the spatial-split SAH tree's build
time and rates (same hits), the roofline of the trace kernel, the CPU
oracle's rate on a sample of the same rays (host threads stated), the
batched-vs-plain comparison over 256 scenes of uneven size, and the animated
20M-triangle any-hit frame (rays in the caller's order, then grouped by the
cells of their origins inside the frame).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s closest-hit (1M tris, incoherent); PLOC build ms; batched vs plain"
HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "cw8_bounce_1m_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_batched(args, dev, rt, scenes, torch):
    """Many scenes of uneven size: per-scene launches (plain) against one
    batched build and one batched launch.  Whole job = PLOC build + compressed
    traversal tree + closest hit of every scene's rays, the same tree and
    kernel either way."""
    log(f"generating {args.batch_scenes} batch scenes")
    vl = scenes.batch_scenes(args.batch_scenes, seed=args.seed)
    counts = [len(v) for v in vl]
    offs = [0]
    for c in counts:
        offs.append(offs[-1] + c)
    verts = np.concatenate(vl)
    del vl
    aabbs, centers = scenes.triangle_aabbs_centers(verts)
    lo = aabbs.view(np.float32).reshape(-1, 8)
    v_d = torch.from_numpy(verts).to(dev)
    a_d = torch.from_numpy(aabbs).to(dev)
    c_d = torch.from_numpy(centers).to(dev)
    box_lo = np.stack([lo[offs[s]:offs[s + 1], 0:3].min(0) for s in range(len(counts))])
    box_hi = np.stack([lo[offs[s]:offs[s + 1], 4:7].max(0) for s in range(len(counts))])
    del verts, aabbs, centers, lo
    ray_counts = [16 * c for c in counts]
    roffs = [0]
    for c in ray_counts:
        roffs.append(roffs[-1] + c)
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 1)
    rays = scenes.batch_rays_torch(box_lo, box_hi, ray_counts, dev, gen)
    roff_d = rt.SceneBatch.ray_offsets(ray_counts, dev)
    n_rays = rays.shape[0]
    hits_plain = torch.empty((n_rays, 4), dtype=torch.float32, device=dev)
    hits_batch = torch.empty_like(hits_plain)
    log(f"batch: {len(counts)} scenes, {offs[-1]} tris, {n_rays} rays")

    def plain(first, last, keep):
        trees = []
        for s in range(first, last):
            p0, p1, r0, r1 = offs[s], offs[s + 1], roffs[s], roffs[s + 1]
            bvh = rt.Builder(c_d[p0:p1], aabbs=a_d[p0:p1]).construct_locally_ordered_clustered()
            tree = bvh.trace_tree(v_d[p0:p1])
            tree.intersect(rays[r0:r1], out=hits_plain[r0:r1])
            if keep:
                trees.append((bvh, tree))
        torch.cuda.synchronize()
        return trees

    def batched():
        b = rt.SceneBatch(c_d, offs, aabbs=a_d)
        b.attach_triangles(v_d)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        b.intersect(rays, roff_d, out=hits_batch)
        torch.cuda.synchronize()
        return b, t1

    # warm-up: kernels loaded, allocator primed; the launch shape is chosen
    # by the autotuner on the batch's own rays (untimed set-up, like the
    # builds) and serves the per-scene launches too
    plain(0, 4, False)
    b, _ = batched()
    outer_shape = rt.get_trace_config()
    shape = rt.autotune(b, rays, roff_d)
    log(f"batch launch shape {shape}")
    del b
    plain_ms, batch_ms, batch_build_ms, batch_trace_ms = [], [], [], []
    trees = None
    for _ in range(2):
        trees = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        trees = plain(0, len(counts), True)
        plain_ms.append((time.perf_counter() - t0) * 1e3)
    # per-scene trace launches alone, trees already built
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, (_, tree) in enumerate(trees):
        tree.intersect(rays[roffs[s]:roffs[s + 1]], out=hits_plain[roffs[s]:roffs[s + 1]])
    torch.cuda.synchronize()
    plain_trace_ms = (time.perf_counter() - t0) * 1e3
    trees[-1][1].status()
    del trees
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b, t1 = batched()
        t2 = time.perf_counter()
        batch_ms.append((t2 - t0) * 1e3)
        batch_build_ms.append((t1 - t0) * 1e3)
        batch_trace_ms.append((t2 - t1) * 1e3)
        del b
    same = bool(torch.equal(hits_plain.view(torch.int32), hits_batch.view(torch.int32)))
    # the same frame as one hipGraph (build + traversal tree + trace)
    plan = rt.BatchPlan(c_d, offs, v_d, aabbs=a_d)
    hits_graph = torch.empty_like(hits_plain)
    graph_ms = []
    for k in range(4):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.launch(c_d, v_d, rays, roff_d, aabbs=a_d, out=hits_graph)
        torch.cuda.synchronize()
        if k > 0:  # the first call records the graph
            graph_ms.append((time.perf_counter() - t0) * 1e3)
    plan.status()
    same_graph = bool(torch.equal(hits_graph.view(torch.int32), hits_batch.view(torch.int32)))
    del hits_graph, plan
    rt.set_trace_config(**outer_shape)
    p_ms, b_ms = statistics.median(plain_ms), statistics.median(batch_ms)
    g_ms = statistics.median(graph_ms)
    bt_ms = statistics.median(batch_trace_ms)
    res = {
        "scenes": len(counts),
        "tris": offs[-1],
        "rays": n_rays,
        "plain_ms": round(p_ms, 2),
        "batched_ms": round(b_ms, 2),
        "plain_trace_ms": round(plain_trace_ms, 2),
        "launch_shape": shape,
        "batched_build_ms": round(statistics.median(batch_build_ms), 2),
        "batched_trace_ms": round(bt_ms, 2),
        "batched_mrays_s": round(n_rays / (bt_ms * 1e-3) / 1e6, 1),
        "plain_mrays_s": round(n_rays / (plain_trace_ms * 1e-3) / 1e6, 1),
        "same_hits": same,
        "graph_ms": round(g_ms, 2),
        "graph_same_hits": same_graph,
        "graph_vs_plain": round(p_ms / g_ms, 3),
        "workload": "scene sizes round(exp(U(ln 1e3, ln 5e5))), alternating noisy sphere / "
                    "soup; 16 rays per triangle, origin uniform in the scene box, uniform "
                    "direction; whole job = PLOC + compressed traversal tree + closest hit "
                    "(plain: per-scene calls; batched: one build and one launch)",
    }
    log(f"batched {b_ms:.1f} ms vs plain {p_ms:.1f} ms (x{p_ms / b_ms:.2f}); trace "
        f"{bt_ms:.1f} vs {plain_trace_ms:.1f} ms; same hits {same}; graph {g_ms:.1f} ms "
        f"(same {same_graph})")
    del rays, hits_plain, hits_batch, v_d, a_d, c_d
    torch.cuda.empty_cache()
    # the batched path proper is the recorded frame (one graph: build + trace)
    return res, p_ms / g_ms


def run_adjoint(args, dev, rt, scenes, torch):
    """100k-triangle noisy sphere, 16M rays, L = sum w_i t_i: forward closest
    hit and the adjoint kernel's backward to vertices, origins and
    directions, timed; gradients of a ray sample checked against a float64
    solve on the host."""
    from rtbvh_amd.diff import trace_tuv

    log("differentiable: 100k-triangle sphere, 16M rays")
    verts_np = scenes.noisy_sphere(250, 200, seed=args.seed)
    aabbs, centers = scenes.triangle_aabbs_centers(verts_np)
    verts = torch.from_numpy(verts_np).to(dev)
    bvh = rt.Builder(torch.from_numpy(centers).to(dev),
                     aabbs=torch.from_numpy(aabbs).to(dev)).construct_locally_ordered_clustered()
    tree = bvh.trace_tree(verts)
    n = 1 << 24
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 3)
    o = torch.randn((n, 3), device=dev, generator=gen)
    o = 3.0 * o / o.norm(dim=1, keepdim=True)
    tgt = 0.25 + 0.5 * torch.rand((n, 3), device=dev, generator=gen)
    d = tgt - o
    d = (d / d.norm(dim=1, keepdim=True)).contiguous()
    o = o.contiguous()
    w = torch.randn(n, device=dev, generator=gen)

    def step():
        vt = verts.detach().requires_grad_(True)
        oo = o.detach().requires_grad_(True)
        dd = d.detach().requires_grad_(True)
        t, u, v, prim = trace_tuv(tree, vt, oo, dd)
        loss = (w * torch.where(prim >= 0, t, torch.zeros_like(t))).sum()
        return loss, vt, oo, dd

    step()[0].backward()
    torch.cuda.synchronize()
    fwd, tot = [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss, vt, oo, dd = step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append((t1 - t0) * 1e3)
        tot.append((t2 - t0) * 1e3)
    # float64 check on the first 200k rays: per-ray gradients by linear solves
    k = 200_000
    vt = verts.detach().requires_grad_(True)
    oo = o[:k].detach().requires_grad_(True)
    dd = d[:k].detach().requires_grad_(True)
    t, u, v, prim = trace_tuv(tree, vt, oo, dd)
    (w[:k] * torch.where(prim >= 0, t, torch.zeros_like(t))).sum().backward()
    p_ = prim.cpu().numpy()
    hit = p_ >= 0
    v64 = verts_np.astype(np.float64)[p_[hit]].reshape(-1, 3, 3)
    o64 = o[:k].cpu().numpy().astype(np.float64)[hit]
    d64 = d[:k].cpu().numpy().astype(np.float64)[hit]
    m = np.stack([-d64, v64[:, 1] - v64[:, 0], v64[:, 2] - v64[:, 0]], axis=-1)
    x = np.linalg.solve(m, (o64 - v64[:, 0])[..., None])[..., 0]
    g = np.zeros_like(x)
    g[:, 0] = w[:k].cpu().numpy().astype(np.float64)[hit]
    lam = np.linalg.solve(np.transpose(m, (0, 2, 1)), g[..., None])[..., 0]
    gv = np.zeros((len(verts_np), 9))
    uu, vv = x[:, 1], x[:, 2]
    np.add.at(gv[:, 0:3], p_[hit], -(1.0 - uu - vv)[:, None] * lam)
    np.add.at(gv[:, 3:6], p_[hit], -uu[:, None] * lam)
    np.add.at(gv[:, 6:9], p_[hit], -vv[:, None] * lam)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))

    errs = {"vertices": rel(vt.grad.cpu().numpy(), gv),
            "origins": rel(oo.grad.cpu().numpy()[hit], lam),
            "directions": rel(dd.grad.cpu().numpy()[hit], x[:, 0:1] * lam)}
    fm, tm = statistics.median(fwd), statistics.median(tot)
    res = {"tris": len(verts_np), "rays": n, "forward_ms": round(fm, 2),
           "forward_backward_ms": round(tm, 2),
           "forward_backward_mrays_s": round(n / (tm * 1e-3) / 1e6, 1),
           "hit_fraction": round(float(hit.mean()), 4),
           "fp64_rel_error_200k_rays": {a: float(f"{b:.3e}") for a, b in errs.items()},
           "workload": "noisy sphere 250x200 quads (100k tris), 2^24 rays from radius 3 at the "
                       "cube [0.25, 0.75]^3, L = sum w t, w ~ N(0,1); forward = trace tree closest "
                       "hit, backward = adjoint kernel (vertex grads by float atomics)"}
    log(f"differentiable: forward {fm:.1f} ms, forward+backward {tm:.1f} ms, errors {errs}")
    del tree, bvh, verts, o, d, w
    torch.cuda.empty_cache()
    return res


def copy_calibration(args, torch, dev):
    """One plain device copy of known size, so that rocprofv3's FETCH_SIZE /
    WRITE_SIZE can be scaled to bytes."""
    if args.copy_calibration <= 0:
        return
    nfl = int(args.copy_calibration * (1 << 30)) // 4
    src = torch.ones(nfl, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy calibration: {nfl * 4} bytes read and written")


def run_large(args, dev, rt, scenes, torch):
    """64 x 64 displaced spheres (19,988,480 triangles) animated per frame:
    each frame moves the vertices, recomputes boxes and centres, rebuilds
    (PLOC + compressed traversal tree) and traces 2^28 any-hit shadow rays to
    a point light — the whole frame as one recorded graph.  For comparison,
    the same rays are also traced eagerly against the reference-layout MBVH."""
    log("generating the 20M-triangle sphere grid")
    base = torch.from_numpy(scenes.sphere_grid()).to(dev)
    n = base.shape[0]
    verts = torch.empty_like(base)
    aabbs = torch.empty((n, 8), dtype=torch.float32, device=dev)
    centers = torch.empty((n, 3), dtype=torch.float32, device=dev)
    scenes.animate_y_torch(base, 0, verts)
    scenes.triangle_aabbs_centers_torch(verts, aabbs, centers)
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 2)
    n_rays = 1 << args.large_rays_log2
    rays = scenes.shadow_rays_torch(verts, n_rays, (32.0, 20.0, 32.0), dev, gen)
    occ = torch.empty((n_rays,), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bvh = rt.Builder(centers, aabbs=aabbs).construct_locally_ordered_clustered()
    torch.cuda.synchronize()
    eager_build_ms = (time.perf_counter() - t0) * 1e3
    del bvh
    plan = rt.BatchPlan(centers, [0, n], verts, aabbs=aabbs)

    def frames():
        ms = []
        for f in range(1, 5):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            scenes.animate_y_torch(base, f, verts)
            scenes.triangle_aabbs_centers_torch(verts, aabbs, centers)
            plan.launch(centers, verts, rays, None, aabbs=aabbs, any_hit=True, out=occ)
            torch.cuda.synchronize()
            if f > 1:  # the first launch records the graph
                ms.append((time.perf_counter() - t0) * 1e3)
        plan.status()
        return ms

    # the rays in the caller's (random) order, then grouped by the cells of
    # their origins (sorted inside each frame's graph; same flags per ray)
    caller_order_ms = frames()
    occ_caller = occ.clone()
    plan.order_rays(n_rays)
    frame_ms = frames()
    same_order = bool(torch.equal(occ, occ_caller))
    del occ_caller
    occluded = float(occ.float().mean().item())
    # the last frame against an eager build + any-hit of the same vertices
    bvh = rt.Builder(centers, aabbs=aabbs).construct_locally_ordered_clustered()
    m = rt.Mbvh.from_bvh(bvh)
    m.attach_triangles(verts)
    want = m.occluded_prepared(rays)
    same = bool(torch.equal(want, occ))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.occluded_prepared(rays, out=want)
    torch.cuda.synchronize()
    trace_ms = (time.perf_counter() - t0) * 1e3
    fm = statistics.median(frame_ms)
    fc = statistics.median(caller_order_ms)
    res = {
        "tris": n,
        "rays": n_rays,
        "frame_ms": round(fm, 2),
        "frame_mrays_s": round(n_rays / (fm * 1e-3) / 1e6, 1),
        "frame_ms_caller_order": round(fc, 2),
        "ordered_same_as_caller_order": same_order,
        "eager_ploc_build_ms": round(eager_build_ms, 2),
        "eager_mbvh_any_hit_ms": round(trace_ms, 2),
        "eager_mbvh_any_hit_mrays_s": round(n_rays / (trace_ms * 1e-3) / 1e6, 1),
        "occluded_fraction": round(occluded, 4),
        "plan_iterations": plan.iterations,
        "same_as_eager": same,
        "workload": "64x64 seeded displaced spheres (61x40 quads each); per frame y += "
                    "0.02 sin(2 pi (x + f/60)), boxes/centres, PLOC rebuild, compressed "
                    "traversal tree and 2^28 any-hit shadow rays to a point light (random "
                    "surface points, in random order), one graph; frame_ms with the rays "
                    "grouped by the cells of their origins inside the graph "
                    "(rtbvh_gpu_batch_plan_order_rays), frame_ms_caller_order without",
    }
    log(f"large: frame {fm:.1f} ms ({res['frame_mrays_s']} Mrays/s incl. rebuild; caller's "
        f"order {fc:.1f} ms, same {same_order}), eager MBVH any-hit {trace_ms:.1f} ms, eager "
        f"PLOC {eager_build_ms:.1f} ms, same {same}")
    del plan, m, bvh, rays, occ, want, base, verts, aabbs, centers
    torch.cuda.empty_cache()
    return res


def cpu_info():
    """Host CPU model, logical CPUs, and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return model, os.cpu_count() or 1, allowed


def cpu_threads():
    """Threads for the CPU oracle: the CPUs this process may use, at most 16
    (one GPU's share of the host on the benchmark machines)."""
    return max(1, min(16, cpu_info()[2]))


def oracle_rate(fn, rays, threads, target_s, reps=5, probe=20_000):
    """Median-of-`reps` rays/s of the oracle traversal fn(rays, threads) on a
    sample sized for about target_s seconds per run."""
    probe = min(probe, len(rays))
    t0 = time.perf_counter()
    fn(rays[:probe], threads)
    rate = probe / max(time.perf_counter() - t0, 1e-9)
    n = int(min(len(rays), max(probe, rate * target_s)))
    sample = rays[:n]
    runs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn(sample, threads)
        runs.append(time.perf_counter() - t0)
    return n / statistics.median(runs), n


def gpu_launch_ms(torch, stream, fn, reps=20):
    """Median per-launch time (HIP events on the launch stream) of fn()."""
    fn()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        ms.append(a.elapsed_time(b))
    return statistics.median(ms)


def run_c1(args, dev, rt, scenes, torch, stream):
    """The reference's own benchmark shape, small meshes with the binned SAH
    builder (examples/benchmark.rs): the teapot (6,320 triangles, the
    reference's fixture) with the benchmark's 1000 x 1000 camera, and the
    80 x 40 torus (6,400 triangles) with 1M rays from a radius-3 sphere aimed
    at points in its box.  GPU: binned SAH build, MBVH collapse, the
    reference-order traversals (BVH2 and MBVH, the reference's visit order and
    tie rule) and the lowest-(t, id) compressed tree, each checked against the
    CPU oracle's traversal of the same tree; CPU oracle: the same build and
    reference-order traversals at 1 thread and at the host share."""
    from oracle import refcpu as ref

    log("C1: teapot and torus, binned SAH")
    golden = os.path.join(ROOT, "tests", "golden", "teapot.obj")
    threads = cpu_threads()
    out = {}
    for name in ("teapot", "torus"):
        if name == "teapot":
            verts = scenes.load_obj(golden)
            rays_np = scenes.reference_camera_rays(1000, 1000)
        else:
            verts = scenes.torus(80, 40)
            v3 = verts.reshape(-1, 3)
            rays_np = scenes.rays_sphere_to_box(1_000_000, v3.min(0), v3.max(0), radius=3.0,
                                                seed=args.seed)
        aabbs, centers = scenes.triangle_aabbs_centers(verts)
        v_d = torch.from_numpy(verts).to(dev)
        a_d = torch.from_numpy(aabbs).to(dev)
        c_d = torch.from_numpy(centers).to(dev)
        r_d = torch.from_numpy(rays_np).to(dev)
        build_ms, coll_ms = [], []
        bvh = m = None
        for i in range(6):
            bvh = m = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bvh = rt.Builder(c_d, aabbs=a_d).construct_binned_sah()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            m = rt.Mbvh.from_bvh(bvh)
            torch.cuda.synchronize()
            if i > 0:
                build_ms.append((t1 - t0) * 1e3)
                coll_ms.append((time.perf_counter() - t1) * 1e3)
        tree = bvh.trace_tree(v_d)
        m.attach_triangles(v_d)
        n = r_d.shape[0]
        hits = torch.empty((n, 4), dtype=torch.float32, device=dev)
        gpu = {}
        for key, fn in (("bvh2_reference_order", lambda: bvh.intersect(v_d, r_d, True, out=hits)),
                        ("mbvh_reference_order",
                         lambda: m.intersect_prepared(r_d, True, out=hits)),
                        ("compressed_tree_lowest_id", lambda: tree.intersect(r_d, out=hits))):
            ms = gpu_launch_ms(torch, stream, fn)
            gpu[key] = round(n / (ms * 1e-3) / 1e6, 1)
        # parity on the GPU-built tree: reference order against the oracle's
        # reference-order traversal, lowest id against the oracle's lowest id
        nodes = bvh.nodes().cpu().numpy().view(ref.AABB).reshape(-1)
        idx = bvh.indices().cpu().numpy().view(np.uint32)
        mn = m.quad_nodes().cpu().numpy().view(ref.MBVH).reshape(-1)
        rays = rays_np.view(ref.RAY).reshape(-1)
        same = {}
        for key, got, want in (
                ("bvh2_reference_order", bvh.intersect(v_d, r_d, True),
                 ref.trace_bvh(nodes, idx, verts, rays, mode=ref.MODE_VISIT_ORDER,
                               threads=threads)),
                ("mbvh_reference_order", m.intersect_prepared(r_d, True),
                 ref.trace_mbvh(mn, idx, verts, rays, mode=ref.MODE_VISIT_ORDER,
                                threads=threads)),
                ("compressed_tree_lowest_id", tree.intersect(r_d),
                 ref.trace_mbvh(mn, idx, verts, rays, threads=threads))):
            g = got.cpu().numpy().view(np.uint32).reshape(-1, 4)
            same[key] = bool(np.array_equal(g, want.view(np.uint32).reshape(-1, 4)))
        # the CPU oracle: the same binned build (1 thread) and traversals
        a_rec = aabbs.view(ref.AABB).reshape(-1)
        cpu_build = []
        for _ in range(5):
            t0 = time.perf_counter()
            ref.build_binned_sah(a_rec, centers)
            cpu_build.append((time.perf_counter() - t0) * 1e3)
        t0 = time.perf_counter()
        ref.build_mbvh(nodes)
        cpu_collapse_ms = (time.perf_counter() - t0) * 1e3
        cpu = {}
        for key, fn in (("bvh2", lambda r, t: ref.trace_bvh(nodes, idx, verts, r,
                                                           mode=ref.MODE_VISIT_ORDER, threads=t)),
                        ("mbvh", lambda r, t: ref.trace_mbvh(mn, idx, verts, r,
                                                            mode=ref.MODE_VISIT_ORDER, threads=t))):
            r1, _ = oracle_rate(fn, rays, 1, 0.4)
            rn, _ = oracle_rate(fn, rays, threads, 0.4)
            cpu[f"{key}_1t_mrays_s"] = round(r1 / 1e6, 3)
            cpu[f"{key}_{threads}t_mrays_s"] = round(rn / 1e6, 3)
        out[name] = {
            "tris": len(verts),
            "rays": n,
            "rays_kind": ("examples/benchmark.rs camera, 1000 x 1000" if name == "teapot" else
                          "radius-3 sphere aimed at uniform points in the mesh box"),
            "gpu_binned_sah_build_ms": round(statistics.median(build_ms), 3),
            "gpu_mbvh_collapse_ms": round(statistics.median(coll_ms), 3),
            "gpu_mrays_s": gpu,
            "gpu_same_as_oracle": same,
            "hit_fraction": round(float((hits[:, 3].view(torch.int32) >= 0).float().mean()), 4),
            "cpu_oracle": dict(cpu, binned_sah_build_ms_1t=round(statistics.median(cpu_build), 3),
                               mbvh_collapse_ms_1t=round(cpu_collapse_ms, 3), threads=threads),
        }
        log(f"C1 {name}: GPU build {out[name]['gpu_binned_sah_build_ms']} ms, {gpu}, same {same}; "
            f"CPU {cpu}")
        del tree, m, bvh, v_d, a_d, c_d, r_d, hits
    out["reference_published"] = {
        "note": "README.md:31-41, teapot, AMD 5950X, 32 threads, coherent camera rays",
        "binned_sah_build_ms": 3.81, "mbvh_collapse_ms": 1.156,
        "bvh2_1t_mrays_s": 4.704701, "bvh2_32t_mrays_s": 77.38225,
        "mbvh_1t_mrays_s": 5.970805, "mbvh_32t_mrays_s": 125.44958}
    return out


def run_build_only(kind, rt, torch, a_d, c_d, v_d, reps):
    """Timed binned-SAH, spatial-split or compressed traversal-tree builds of
    the 1M scene (the last over one PLOC tree), each from a synchronised
    device after the previous tree was released."""
    log(f"{kind} builds")
    ms, tree = [], None
    if kind == "tree":
        bvh = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered()
        for i in range(reps):
            tree = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tree = bvh.trace_tree(v_d)
            torch.cuda.synchronize()
            if i > 0:
                ms.append((time.perf_counter() - t0) * 1e3)
        log(f"build only: tree {statistics.median(ms):.3f} ms (median of {len(ms)}), "
            f"{tree.node_count} nodes")
        return
    for i in range(reps):
        tree = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = rt.Builder(c_d, aabbs=a_d)
        tree = b.construct_binned_sah() if kind == "binned" else b.construct_spatial_sah(v_d)
        torch.cuda.synchronize()
        if i > 0:
            ms.append((time.perf_counter() - t0) * 1e3)
    log(f"build only: {kind} {statistics.median(ms):.3f} ms (median of {len(ms)}), "
        f"{tree.index_count} references, {tree.node_count} nodes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--image", type=int, default=8192, help="primary image side")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget for the CPU oracle baseline (0 disables it)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--quick", action="store_true",
                    help="one build of each kind and no binned SAH (for profiled runs)")
    ap.add_argument("--max-leaf", type=int, default=1,
                    help="primitives per leaf of the 4-wide traversal tree")
    ap.add_argument("--width", type=int, default=8, choices=(4, 8),
                    help="traversal tree of the headline: 8 compressed (default) or 4-wide")
    ap.add_argument("--batch-scenes", type=int, default=256,
                    help="scenes in the batched-vs-plain comparison (0 skips it)")
    ap.add_argument("--large", type=int, default=1,
                    help="1: also run the animated 20M-triangle any-hit frame (0 skips it)")
    ap.add_argument("--large-rays-log2", type=int, default=28)
    ap.add_argument("--spatial", type=int, default=1,
                    help="1: also build the spatial-split SAH tree and trace the same rays")
    ap.add_argument("--c1", type=int, default=1,
                    help="1: also run the reference's own benchmark shape (teapot and torus, "
                         "binned SAH, 1M rays) on the GPU and the CPU oracle")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="1: also time forward + adjoint backward on 100k triangles (0 skips it)")
    ap.add_argument("--build-only", nargs="?", const="ploc", default=None,
                    choices=("ploc", "binned", "spatial", "tree"),
                    help="only the timed builds of the 1M scene with one builder (default PLOC), "
                         "for profiler runs; no JSON line")
    ap.add_argument("--trace-only", action="store_true",
                    help="build, then only the timed bounce-ray steps of the headline trace "
                         "(and the copy calibration) for profiler runs; no JSON line")
    ap.add_argument("--large-only", action="store_true",
                    help="only the animated 20M-triangle frames (caller's order, then ordered "
                         "rays) and their eager comparison, for profiler runs; no JSON line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of autotuning")
    ap.add_argument("--shape", default="",
                    help="launch shape block,lds_stack,occupancy_pct,kernel,refill,schedule,"
                         "postpone instead of autotuning (to profile the shape a run picked)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="GiB copied by one plain device copy at the end (scales rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs "
              "on a single GPU (rays and scenes would shard across GPUs with no exchange)",
              flush=True)
        sys.exit(2)

    import torch

    import rtbvh_amd as rt
    from rtbvh_amd import _lib, scenes

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)

    if args.large_only:
        run_large(args, dev, rt, scenes, torch)
        return

    log("generating the 1M-triangle scene")
    verts = scenes.scene_1m()
    aabbs, centers = scenes.triangle_aabbs_centers(verts)
    v_d = torch.from_numpy(verts).to(dev)
    a_d = torch.from_numpy(aabbs).to(dev)
    c_d = torch.from_numpy(centers).to(dev)
    n_tris = len(verts)

    if args.build_only in ("binned", "spatial", "tree"):
        run_build_only(args.build_only, rt, torch, a_d, c_d, v_d, 2 if args.quick else 6)
        return

    # ---- build: PLOC + collapse ------------------------------------------
    log("PLOC builds")
    build_ms = []
    bvh = None
    reps = 2 if args.quick else 6
    # (each timed region starts after the previous result is released: a
    # release frees device memory, which waits for the device)
    for i in range(reps):
        bvh = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bvh = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered()
        torch.cuda.synchronize()
        if i > 0:
            build_ms.append((time.perf_counter() - t0) * 1e3)
    if args.build_only:
        log(f"build only: PLOC {statistics.median(build_ms):.3f} ms (median of {len(build_ms)})")
        return
    collapse_ms = []
    mbvh = None
    for i in range(2 if args.quick else 4):
        mbvh = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mbvh = rt.Mbvh.from_bvh(bvh)
        torch.cuda.synchronize()
        if i > 0:
            collapse_ms.append((time.perf_counter() - t0) * 1e3)
    binned_ms = [float("nan")]
    for i in range(0 if args.quick else 3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sah = rt.Builder(c_d, aabbs=a_d).construct_binned_sah()
        torch.cuda.synchronize()
        if i > 0:
            binned_ms.append((time.perf_counter() - t0) * 1e3)
        del sah
    if len(binned_ms) > 1:
        binned_ms = binned_ms[1:]
    log(f"PLOC {statistics.median(build_ms):.2f} ms, collapse {statistics.median(collapse_ms):.2f} ms,"
        f" binned SAH {statistics.median(binned_ms):.2f} ms; MBVH nodes {mbvh.node_count}")

    # ---- rays ---------------------------------------------------------------
    W = args.image
    n_rays = W * W
    log(f"generating {n_rays} primary rays")
    primary = scenes.pinhole_rays_torch(W, W, dev)
    hits_p = torch.empty((n_rays, 4), dtype=torch.float32, device=dev)
    mbvh.intersect(v_d, primary, out=hits_p)
    torch.cuda.synchronize()
    mbvh.status()
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    bounce = torch.empty((n_rays, 8), dtype=torch.float32, device=dev)
    chunk = 1 << 23
    log("generating bounce rays")
    for s in range(0, n_rays, chunk):
        e = min(n_rays, s + chunk)
        bounce[s:e] = scenes.bounce_rays_torch(v_d, primary[s:e], hits_p[s:e], gen)
    torch.cuda.synchronize()

    attach_ms = []
    for i in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mbvh.attach_triangles(v_d)
        torch.cuda.synchronize()
        if i > 0:
            attach_ms.append((time.perf_counter() - t0) * 1e3)

    tree_ms = []
    ttree = None
    for i in range(3):
        ttree = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ttree = bvh.trace_tree(v_d, max_leaf=args.max_leaf if args.width == 4 else 1,
                               width=args.width)
        torch.cuda.synchronize()
        if i > 0:
            tree_ms.append((time.perf_counter() - t0) * 1e3)
    log(f"traversal tree (width {args.width}): {ttree.node_count} nodes, "
        f"{statistics.median(tree_ms):.2f} ms")
    # the other traversal tree, for comparison on the same rays
    other_width = 4 if args.width == 8 else 8
    otree = bvh.trace_tree(v_d, width=other_width)

    def trace_vertex(rays, out):
        mbvh.intersect(v_d, rays, out=out)

    def trace_prepared(rays, out):
        mbvh.intersect_prepared(rays, out=out)

    def trace_tree(rays, out):
        ttree.intersect(rays, out=out)

    def trace_other(rays, out):
        otree.intersect(rays, out=out)

    def timed(rays, steps, warmup, fn=trace_tree):
        out = torch.empty((rays.shape[0], 4), dtype=torch.float32, device=dev)
        for _ in range(warmup):
            fn(rays, out)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(stream)
            fn(rays, out)
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        mbvh.status()
        per = [a.elapsed_time(b) for a, b in ev]
        return out, per, wall

    # launch shape chosen by the autotuner on the bounce rays themselves
    # (untimed set-up, like the builds; a smaller sample favours smaller
    # persistent grids than the full launch wants)
    if args.shape:
        keys = ("block", "lds_stack", "occupancy_pct", "kernel", "refill", "schedule", "postpone")
        vals = [int(x) for x in args.shape.split(",")]
        if len(vals) != len(keys):
            raise SystemExit(f"--shape takes {len(keys)} integers: {','.join(keys)}")
        rt.set_trace_config(**dict(zip(keys, vals)))
        tuned = rt.get_trace_config()
    elif args.no_autotune:
        tuned = rt.get_trace_config()
    else:
        tuned = rt.autotune(ttree, bounce)
    log(f"launch shape {tuned}")

    if args.trace_only:
        del hits_p
        _, per, _ = timed(bounce, args.steps, args.warmup)
        log(f"trace only: {n_rays / (sum(per) / len(per) * 1e-3) / 1e6:.1f} Mrays/s")
        copy_calibration(args, torch, dev)
        return

    log("timing primary rays")
    _, per_p, _ = timed(primary, 3, 1)
    primary_mrays = n_rays / (statistics.median(per_p) * 1e-3) / 1e6
    del hits_p

    log(f"timing {args.steps} steps over the bounce rays")
    hits, per, wall = timed(bounce, args.steps, args.warmup)
    ms_per_step = sum(per) / len(per)
    value = n_rays / (ms_per_step * 1e-3) / 1e6
    log(f"bounce {value:.1f} Mrays/s ({ms_per_step:.2f} ms/step), primary {primary_mrays:.1f} Mrays/s")
    # the reference-layout MBVH (the collapse of src/bvh.rs:381-404) traced over
    # leaf-ordered triangles and over the caller's vertex array: same hits
    hits_m, per_m, _ = timed(bounce, 3, 1, fn=trace_prepared)
    mbvh_mrays = n_rays / (statistics.median(per_m) * 1e-3) / 1e6
    same_mbvh_hits = bool(torch.equal(hits_m.view(torch.int32), hits.view(torch.int32)))
    del hits_m
    hits_v, per_v, _ = timed(bounce, 3, 1, fn=trace_vertex)
    vertex_mrays = n_rays / (statistics.median(per_v) * 1e-3) / 1e6
    same_vertex_hits = bool(torch.equal(hits_v.view(torch.int32), hits.view(torch.int32)))
    del hits_v
    log(f"MBVH + leaf-ordered triangles {mbvh_mrays:.1f} Mrays/s (same {same_mbvh_hits}), "
        f"MBVH + vertex array {vertex_mrays:.1f} Mrays/s (same {same_vertex_hits})")
    hits_o, per_o, _ = timed(bounce, 3, 1, fn=trace_other)
    other_mrays = n_rays / (statistics.median(per_o) * 1e-3) / 1e6
    same_other_hits = bool(torch.equal(hits_o.view(torch.int32), hits.view(torch.int32)))
    del hits_o
    log(f"width-{other_width} traversal tree {other_mrays:.1f} Mrays/s (same {same_other_hits})")
    del otree

    # ---- spatial-split SAH tree (src/builders/spatial_sah.rs), same rays ----
    spatial = None
    if args.spatial:
        sp_ms = []
        sp = None
        for i in range(3):
            sp = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sp = rt.Builder(c_d, aabbs=a_d).construct_spatial_sah(v_d)
            torch.cuda.synchronize()
            if i > 0:
                sp_ms.append((time.perf_counter() - t0) * 1e3)
        sp_tree = sp.trace_tree(v_d, width=args.width)

        def trace_spatial(rays, out):
            sp_tree.intersect(rays, out=out)

        hits_s, per_s, _ = timed(bounce, 3, 1, fn=trace_spatial)
        sp_bounce = n_rays / (statistics.median(per_s) * 1e-3) / 1e6
        same_sp = bool(torch.equal(hits_s.view(torch.int32), hits.view(torch.int32)))
        del hits_s
        hits_s, per_s, _ = timed(primary, 3, 1, fn=trace_spatial)
        sp_primary = n_rays / (statistics.median(per_s) * 1e-3) / 1e6
        del hits_s
        spatial = {
            "build_ms": round(statistics.median(sp_ms), 2),
            "references": sp.index_count,
            "trace_tree_nodes": sp_tree.node_count,
            "bounce_mrays_s": round(sp_bounce, 1),
            "primary_mrays_s": round(sp_primary, 1),
            "same_hits": same_sp,
        }
        log(f"spatial SAH: build {spatial['build_ms']} ms, {sp.index_count} references, "
            f"bounce {sp_bounce:.1f} Mrays/s (same {same_sp}), primary {sp_primary:.1f} Mrays/s")
        del sp_tree, sp

    # ---- kernel work counters (nodes and triangles per ray) ----------------
    import ctypes

    def work_counts(call):
        counters = torch.zeros(2, dtype=torch.int64, device=dev)
        stats_out = torch.empty_like(hits)
        _lib.check(call(stats_out, counters), "stats")
        torch.cuda.synchronize()
        same = bool(torch.equal(stats_out.view(torch.int32), hits.view(torch.int32)))
        return counters[0].item() / n_rays, counters[1].item() / n_rays, same

    nodes_per_ray, tris_per_ray, same_stats_hits = work_counts(
        lambda out, cnt: _lib.lib().rtbvh_gpu_trace_tree_intersect_stats(
            ttree._h, ctypes.c_void_p(bounce.data_ptr()), n_rays, ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(cnt.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))
    m_nodes_per_ray, m_tris_per_ray, _ = work_counts(
        lambda out, cnt: _lib.lib().rtbvh_gpu_intersect_mbvh_stats(
            mbvh._h, ctypes.c_void_p(v_d.data_ptr()), n_tris, ctypes.c_void_p(bounce.data_ptr()),
            n_rays, ctypes.c_void_p(out.data_ptr()), 0, ctypes.c_void_p(cnt.data_ptr()),
            ctypes.c_void_p(stream.cuda_stream)))
    log(f"per ray: traversal tree {nodes_per_ray:.2f} nodes / {tris_per_ray:.2f} triangles, "
        f"MBVH {m_nodes_per_ray:.2f} / {m_tris_per_ray:.2f}")
    # ray 32 B + hit 16 B + one node record (80 B compressed, 128 B 4-wide)
    # per visit + 48 B per leaf-ordered triangle
    node_bytes = 80 if args.width == 8 else 128
    b_ray = 32 + 16 + node_bytes * nodes_per_ray + 48 * tris_per_ray
    achieved = value * 1e6 * b_ray / 1e9
    # HBM bytes per launch, measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # of this kernel (scaled by a plain copy), kept beside the kernel trace
    # (the entry of the launch shape this run picked when there is one)
    traffic, traffic_shape = None, None
    if os.path.exists(PMC_SUMMARY):
        with open(PMC_SUMMARY) as fh:
            pmc = json.load(fh)
        traffic, traffic_shape = pmc.get("hbm_bytes_per_launch"), pmc.get("launch_shape")
        for entry in pmc.get("by_shape", []):
            if entry.get("launch_shape") == dict(tuned):
                traffic, traffic_shape = entry["hbm_bytes_per_launch"], entry["launch_shape"]

    # ---- CPU oracle on a sample of the same rays ----------------------------
    # the reference's algorithm restated in C++ (oracle/refcpu.cpp): its
    # BVH2 and MBVH iterators in their visit order with the first-accepted
    # closest hit (examples/benchmark.rs:15-41), over the same PLOC tree (the
    # oracle's LOCB, which the GPU build equals bit for bit) and its MBVH, at
    # 1 thread and at the host's share; median of 5
    cpu = None
    parity = None
    if args.cpu_seconds > 0:
        from oracle import refcpu as ref

        threads = cpu_threads()
        model, nproc, allowed = cpu_info()
        m_nodes = mbvh.quad_nodes().cpu().numpy().view(ref.MBVH).reshape(-1)
        m_idx = mbvh.indices().cpu().numpy().view(np.uint32)
        b_nodes = bvh.nodes().cpu().numpy().view(ref.AABB).reshape(-1)
        b_idx = bvh.indices().cpu().numpy().view(np.uint32)
        budget = args.cpu_seconds / 20.0  # per timed run: 2 trees x 2 thread counts x 5
        pool = bounce[:4_000_000].cpu().numpy().view(ref.RAY).reshape(-1)
        rates = {}
        for key, fn in (("mbvh", lambda r, t: ref.trace_mbvh(m_nodes, m_idx, verts, r,
                                                            mode=ref.MODE_VISIT_ORDER, threads=t)),
                        ("bvh2", lambda r, t: ref.trace_bvh(b_nodes, b_idx, verts, r,
                                                           mode=ref.MODE_VISIT_ORDER, threads=t))):
            for t in (threads, 1):
                rates[(key, t)] = oracle_rate(fn, pool, t, budget)
        cpu_mrays = rates[("mbvh", threads)][0] / 1e6
        # parity of the headline's lowest-(t, id) hits on a 1M-ray sample
        n_par = min(n_rays, 1_000_000)
        want = ref.trace_mbvh(m_nodes, m_idx, verts, pool[:n_par], threads=threads)
        got = hits[:n_par].cpu().numpy().view(np.uint32).reshape(-1, 4)
        mism = int(np.count_nonzero(np.any(got != want.view(np.uint32).reshape(-1, 4), axis=1)))
        parity = {"rays_checked": n_par, "mismatches": mism}
        t0 = time.perf_counter()
        ref.build_locb(aabbs.view(ref.AABB).reshape(-1), centers)
        cpu_build_ms = (time.perf_counter() - t0) * 1e3
        cpu = {"value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"first {rates[('mbvh', threads)][1]} of the {n_rays} bounce rays "
                         f"({rates[('mbvh', 1)][1]} at 1 thread), the oracle's MBVH iterator in "
                         f"the reference's visit order with the first-accepted closest hit "
                         f"(examples/benchmark.rs:15-41) over the same PLOC tree, median of 5, "
                         f"{threads} threads (this process may use {allowed} of {nproc} host "
                         f"CPUs; one GPU's share is 16)",
               "cpu_model": model,
               "mbvh_1t_mrays_s": round(rates[("mbvh", 1)][0] / 1e6, 3),
               "bvh2_mrays_s": round(rates[("bvh2", threads)][0] / 1e6, 3),
               "bvh2_1t_mrays_s": round(rates[("bvh2", 1)][0] / 1e6, 3),
               "locb_build_ms_1t": round(cpu_build_ms, 1)}
        log(f"CPU oracle ({model}, {threads} threads): MBVH {cpu_mrays:.2f} Mrays/s "
            f"(1T {cpu['mbvh_1t_mrays_s']}), BVH2 {cpu['bvh2_mrays_s']} "
            f"(1T {cpu['bvh2_1t_mrays_s']}); parity {parity}")

    c1_info = run_c1(args, dev, rt, scenes, torch, stream) if args.c1 else None

    batch_info, batch_ratio = None, None
    if args.batch_scenes > 0:
        del bounce
        torch.cuda.empty_cache()
        batch_info, batch_ratio = run_batched(args, dev, rt, scenes, torch)

    large_info = run_large(args, dev, rt, scenes, torch) if args.large else None
    adjoint_info = run_adjoint(args, dev, rt, scenes, torch) if args.adjoint else None

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"1M-tri noisy sphere + soup; GPU PLOC (r=14) + "
                        f"{'compressed 8-wide' if args.width == 8 else '4-wide'} traversal tree; "
                        f"closest hit of 8192^2 cosine-bounce rays",
            "tris": n_tris,
            "rays_per_step": n_rays,
            "tie_rule": "lowest (t, prim id)",
        },
        "ploc_build_ms": round(statistics.median(build_ms), 3),
        "mbvh_collapse_ms": round(statistics.median(collapse_ms), 3),
        "attach_triangles_ms": round(statistics.median(attach_ms), 3),
        "trace_tree_build_ms": round(statistics.median(tree_ms), 3),
        "trace_tree_width": args.width,
        "trace_tree_max_leaf": args.max_leaf if args.width == 4 else 1,
        "trace_tree_nodes": ttree.node_count,
        f"width{other_width}_tree_mrays_s": round(other_mrays, 2),
        f"width{other_width}_tree_same_hits": same_other_hits,
        "spatial_sah": spatial,
        "trace_launch_shape": tuned,
        "binned_sah_build_ms": round(statistics.median(binned_ms), 3),
        "primary_mrays_s": round(primary_mrays, 2),
        "mbvh_prepared_mrays_s": round(mbvh_mrays, 2),
        "mbvh_prepared_same_hits": same_mbvh_hits,
        "mbvh_vertex_mrays_s": round(vertex_mrays, 2),
        "mbvh_vertex_same_hits": same_vertex_hits,
        "mbvh_nodes_per_ray": round(m_nodes_per_ray, 3),
        "mbvh_tris_per_ray": round(m_tris_per_ray, 3),
        "wall_s_timed": round(wall, 4),
        "nodes_per_ray": round(nodes_per_ray, 3),
        "tris_per_ray": round(tris_per_ray, 3),
        "stats_kernel_same_hits": same_stats_hits,
        "batched_vs_plain": None if batch_ratio is None else round(batch_ratio, 3),
        "batched": batch_info,
        "large_animated": large_info,
        "differentiable": adjoint_info,
        "plain_small_meshes": c1_info,
        "parity_vs_cpu_oracle": parity,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_ray": round(b_ray, 1),
            # the measured HBM bytes over this run's launch time: the share of
            # HBM bandwidth the kernel really draws (most node fetches hit L2)
            "hbm_frac": (None if traffic is None else
                         round(traffic / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)),
            "traffic_launch_shape": traffic_shape,
            "bytes_per_ray_compulsory": 48,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    copy_calibration(args, torch, dev)


if __name__ == "__main__":
    main()
