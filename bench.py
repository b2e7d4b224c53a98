#!/usr/bin/env python3
"""Benchmark of the GPU hot path on one MI355X.

Workload (1M-triangle scene, the benchmark configuration's second case):
  1,000,000 FP32 triangles — a 750x600-quad noisy sphere (900k) plus 100k
  random triangles (edge 0.01) in the unit cube; GPU PLOC build (search
  radius 14) + MBVH collapse; 8192^2 = 67,108,864 pinhole primary rays and
  the same number of cosine-weighted bounce rays from the primary hits.

A step is one closest-hit pass over the 67,108,864 bounce rays (the
incoherent set), inputs resident in HBM.  The headline value is
bounce-ray Mrays/s; the JSON line also carries the PLOC build time, the MBVH
collapse time, the primary-ray rate, the roofline of the trace kernel and
the CPU oracle's rate on a sample of the same rays (host threads stated).

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
PMC_SUMMARY = os.path.join(ROOT, "profiles", "trace_mbvh_bounce_1m_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--image", type=int, default=8192, help="primary image side")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="budget for the CPU oracle baseline (0 disables it)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--quick", action="store_true",
                    help="one build of each kind and no binned SAH (for profiled runs)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="GiB copied by one plain device copy at the end (scales rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (rays and scenes "
              "shard with no exchange, see DESIGN.md)", flush=True)
        sys.exit(2)

    import torch

    import rtbvh_amd as rt
    from rtbvh_amd import _lib, scenes

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)

    log("generating the 1M-triangle scene")
    verts = scenes.scene_1m()
    aabbs, centers = scenes.triangle_aabbs_centers(verts)
    v_d = torch.from_numpy(verts).to(dev)
    a_d = torch.from_numpy(aabbs).to(dev)
    c_d = torch.from_numpy(centers).to(dev)
    n_tris = len(verts)

    # ---- build: PLOC + collapse ------------------------------------------
    log("PLOC builds")
    build_ms = []
    bvh = None
    reps = 2 if args.quick else 6
    for i in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bvh = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered()
        torch.cuda.synchronize()
        if i > 0:
            build_ms.append((time.perf_counter() - t0) * 1e3)
    collapse_ms = []
    mbvh = None
    for i in range(2 if args.quick else 4):
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
        rt.Builder(c_d, aabbs=a_d).construct_binned_sah()
        torch.cuda.synchronize()
        if i > 0:
            binned_ms.append((time.perf_counter() - t0) * 1e3)
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

    def timed(rays, steps, warmup):
        out = torch.empty((rays.shape[0], 4), dtype=torch.float32, device=dev)
        for _ in range(warmup):
            mbvh.intersect(v_d, rays, out=out)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(stream)
            mbvh.intersect(v_d, rays, out=out)
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        mbvh.status()
        per = [a.elapsed_time(b) for a, b in ev]
        return out, per, wall

    log("timing primary rays")
    _, per_p, _ = timed(primary, 3, 1)
    primary_mrays = n_rays / (statistics.median(per_p) * 1e-3) / 1e6
    del hits_p

    log(f"timing {args.steps} steps over the bounce rays")
    hits, per, wall = timed(bounce, args.steps, args.warmup)
    ms_per_step = sum(per) / len(per)
    value = n_rays / (ms_per_step * 1e-3) / 1e6
    log(f"bounce {value:.1f} Mrays/s ({ms_per_step:.2f} ms/step), primary {primary_mrays:.1f} Mrays/s")

    # ---- kernel work counters (MBVH nodes and triangles per ray) -----------
    counters = torch.zeros(2, dtype=torch.int64, device=dev)
    stats_out = torch.empty_like(hits)
    import ctypes

    _lib.check(_lib.lib().rtbvh_gpu_intersect_mbvh_stats(
        mbvh._h, ctypes.c_void_p(v_d.data_ptr()), n_tris, ctypes.c_void_p(bounce.data_ptr()),
        n_rays, ctypes.c_void_p(stats_out.data_ptr()), 0, ctypes.c_void_p(counters.data_ptr()),
        ctypes.c_void_p(stream.cuda_stream)), "rtbvh_gpu_intersect_mbvh_stats")
    torch.cuda.synchronize()
    nodes_per_ray = counters[0].item() / n_rays
    tris_per_ray = counters[1].item() / n_rays
    same_stats_hits = bool(torch.equal(stats_out.view(torch.int32), hits.view(torch.int32)))
    del stats_out
    b_ray = 32 + 16 + 128 * nodes_per_ray + 40 * tris_per_ray
    achieved = value * 1e6 * b_ray / 1e9
    traffic = None
    if os.path.exists(PMC_SUMMARY):
        with open(PMC_SUMMARY) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    # ---- CPU oracle on a sample of the same rays ----------------------------
    cpu = None
    parity = None
    if args.cpu_seconds > 0:
        from oracle import refcpu as ref

        threads = min(16, os.cpu_count() or 1)
        m_nodes = mbvh.quad_nodes().cpu().numpy().view(ref.MBVH).reshape(-1)
        m_idx = mbvh.indices().cpu().numpy().view(np.uint32)
        probe = 100_000
        sample = bounce[:probe].cpu().numpy().view(ref.RAY).reshape(-1)
        t0 = time.perf_counter()
        ref.trace_mbvh(m_nodes, m_idx, verts, sample, threads=threads)
        rate = probe / (time.perf_counter() - t0)
        n_sample = int(min(n_rays, max(probe, rate * args.cpu_seconds / 3)))
        n_sample -= n_sample % 1000
        sample = bounce[:n_sample].cpu().numpy().view(ref.RAY).reshape(-1)
        runs = []
        want = None
        for _ in range(3):
            t0 = time.perf_counter()
            want = ref.trace_mbvh(m_nodes, m_idx, verts, sample, threads=threads)
            runs.append(time.perf_counter() - t0)
        cpu_mrays = n_sample / statistics.median(runs) / 1e6
        got = hits[:n_sample].cpu().numpy().view(np.uint32).reshape(-1, 4)
        mism = int(np.count_nonzero(np.any(got != want.view(np.uint32).reshape(-1, 4), axis=1)))
        parity = {"rays_checked": n_sample, "mismatches": mism}
        t0 = time.perf_counter()
        ref.build_locb(aabbs.view(ref.AABB).reshape(-1), centers)
        cpu_build_ms = (time.perf_counter() - t0) * 1e3
        cpu = {"value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"first {n_sample} of the {n_rays} bounce rays, same MBVH and triangles, "
                         f"oracle lowest-id closest hit, median of 3; oracle LOCB build of the "
                         f"1M scene {cpu_build_ms:.0f} ms on 1 thread"}
        log(f"CPU oracle {cpu_mrays:.2f} Mrays/s on {threads} threads; parity {parity}")

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
            "workload": "1M-tri noisy sphere + soup; GPU PLOC (r=14) + MBVH; closest hit of "
                        "8192^2 cosine-bounce rays",
            "tris": n_tris,
            "rays_per_step": n_rays,
            "tie_rule": "lowest (t, prim id)",
        },
        "ploc_build_ms": round(statistics.median(build_ms), 3),
        "mbvh_collapse_ms": round(statistics.median(collapse_ms), 3),
        "binned_sah_build_ms": round(statistics.median(binned_ms), 3),
        "primary_mrays_s": round(primary_mrays, 2),
        "wall_s_timed": round(wall, 4),
        "nodes_per_ray": round(nodes_per_ray, 3),
        "tris_per_ray": round(tris_per_ray, 3),
        "stats_kernel_same_hits": same_stats_hits,
        "batched_vs_plain": None,
        "parity_vs_cpu_oracle": parity,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_ray": round(b_ray, 1),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    if args.copy_calibration > 0:
        nfl = int(args.copy_calibration * (1 << 30)) // 4
        src = torch.ones(nfl, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nfl * 4} bytes read and written")


if __name__ == "__main__":
    main()
