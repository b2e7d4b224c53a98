"""Capturable build + trace (rtbvh_gpu_batch_plan_*): the same frame as the
eager batched calls, enqueued without allocation or host synchronisation and
replayed as a hipGraph.

  * enqueue and every graph replay return, bit for bit, the hits of the eager
    SceneBatch (which equal per-scene traces and brute force, see
    test_gpu_batch.py);
  * replays are checked one by one with eager work and synchronisation in
    between, and after the vertices move in place (an animated frame) the
    replay equals an eager build of the moved scenes;
  * a one-scene plan's any-hit flags equal Mbvh.occluded_prepared;
  * the first launch (warm-up and capture) is ordered after the caller's
    stream.
"""
import numpy as np
import pytest

from oracle import refcpu as ref
from rtbvh_amd import scenes
from tests.helpers import as_rays, flat, hits_equal, prep, ray_set, teapot

torch = pytest.importorskip("torch")
import rtbvh_amd as rt  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).view(dtype).cuda()


def _batch():
    vl = [teapot(), scenes.noisy_sphere(70, 50, seed=3), scenes.random_soup(9000, 0.02, seed=4),
          scenes.random_soup(5, 0.2, seed=5), scenes.torus(60, 30)]
    counts = [len(v) for v in vl]
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64).tolist()
    rays_l = [ray_set(v, 4000, seed=30 + s, kind="box") for s, v in enumerate(vl)]
    return vl, offs, rays_l


def _eager(verts, offs, rays, roff):
    aabbs, centers = prep(verts)
    b = rt.SceneBatch(_dev(centers), offs, aabbs=_dev(flat(aabbs)))
    b.attach_triangles(_dev(verts))
    out = b.intersect(_dev(rays), roff).cpu().numpy()
    torch.cuda.synchronize()
    return out


def test_plan_enqueue_and_graph_replays():
    print("[plan_graph]", flush=True)
    vl, offs, rays_l = _batch()
    verts = np.concatenate(vl)
    rays = np.concatenate(rays_l)
    aabbs, centers = prep(verts)
    a_d, c_d, v_d, r_d = _dev(flat(aabbs)), _dev(centers), _dev(verts), _dev(rays)
    roff = rt.SceneBatch.ray_offsets([len(r) for r in rays_l], "cuda")
    want = _eager(verts, offs, rays, roff)
    plan = rt.BatchPlan(c_d, offs, v_d, aabbs=a_d)
    assert plan.iterations > 0
    got = plan.enqueue(c_d, v_d, r_d, roff, aabbs=a_d).cpu().numpy()
    plan.status()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    out = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    for k in range(4):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        plan.launch(c_d, v_d, r_d, roff, aabbs=a_d, out=out)
        res = out.cpu().numpy()
        plan.status()
        assert np.array_equal(res.view(np.uint32), want.view(np.uint32)), f"replay {k}"
        # unrelated eager work between replays
        torch.cuda.synchronize()
        _ = (v_d * 2.0).sum().item()
    # animated frame: move the vertices in place, refresh boxes and centres in place
    moved = verts.copy()
    moved[:, 1::3] += np.float32(0.01) * np.sin(7.0 * moved[:, 0::3]).astype(np.float32)
    m_aabbs, m_centers = prep(moved)
    v_d.copy_(_dev(moved))
    a_d.copy_(_dev(flat(m_aabbs)))
    c_d.copy_(_dev(m_centers))
    want2 = _eager(moved, offs, rays, roff)
    plan.launch(c_d, v_d, r_d, roff, aabbs=a_d, out=out)
    res2 = out.cpu().numpy()
    plan.status()
    assert np.array_equal(res2.view(np.uint32), want2.view(np.uint32))
    assert not np.array_equal(res2.view(np.uint32), want.view(np.uint32))
    # and every scene against brute force
    start = 0
    for s, v in enumerate(vl):
        n = len(rays_l[s])
        mv = moved[offs[s]:offs[s + 1]]
        brute = ref.trace_brute(mv, as_rays(rays_l[s]), threads=8)
        assert hits_equal(res2[start:start + n].view(ref.HIT).reshape(-1), brute).all(), s
        start += n


def test_single_scene_plan_any_hit():
    print("[plan_any_hit]", flush=True)
    verts = scenes.noisy_sphere(80, 60, seed=9)
    aabbs, centers = prep(verts)
    a_d, c_d, v_d = _dev(flat(aabbs)), _dev(centers), _dev(verts)
    rays = ray_set(verts, 30000, seed=2, kind="box")
    rays[:, 7] = np.random.default_rng(3).uniform(0.0, 1.0, len(rays)).astype(np.float32)
    r_d = _dev(rays)
    plan = rt.BatchPlan(c_d, [0, len(verts)], v_d, aabbs=a_d)
    occ = plan.launch(c_d, v_d, r_d, None, aabbs=a_d, any_hit=True).cpu().numpy()
    plan.status()
    bvh = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered()
    m = rt.Mbvh.from_bvh(bvh)
    m.attach_triangles(v_d)
    want = m.occluded_prepared(r_d).cpu().numpy()
    assert np.array_equal(occ, want)
    assert 0 < occ.mean() < 1


def test_first_launch_waits_for_the_callers_stream():
    """The first launch of a plan (eager warm-up, then capture) runs after the
    work already queued on the caller's stream: a copy of `out` queued behind
    slow work still sees the old contents, and the result equals the eager
    path, with no synchronisation between producing the inputs and the
    launch."""
    print("[plan_stream_order]", flush=True)
    vl, offs, rays_l = _batch()
    verts = np.concatenate(vl)
    rays = np.concatenate(rays_l)
    aabbs, centers = prep(verts)
    a_d, c_d, v_d, r_src = _dev(flat(aabbs)), _dev(centers), _dev(verts), _dev(rays)
    roff = rt.SceneBatch.ray_offsets([len(r) for r in rays_l], "cuda")
    want = _eager(verts, offs, rays, roff)
    plan = rt.BatchPlan(c_d, offs, v_d, aabbs=a_d)
    out = torch.full((len(rays), 4), -1.0, dtype=torch.float32, device="cuda")
    big = torch.randn((6144, 6144), device="cuda")
    torch.cuda.synchronize()
    for _ in range(6):  # tens of milliseconds of work ahead of the copy
        big = big @ big
        big = big / big.norm()
    seen = out.clone()            # queued behind the slow work
    r_d = r_src * 1.0             # the rays, produced asynchronously
    plan.launch(c_d, v_d, r_d, roff, aabbs=a_d, out=out)
    plan.status()
    torch.cuda.synchronize()
    assert bool((seen == -1.0).all())
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32))
