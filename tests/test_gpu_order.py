"""Rays traced in the order of their origins (RTBVH_TRACE_ORDER_RAYS,
rtbvh_gpu_batch_plan_order_rays; no counterpart in the reference).

A ray's result does not depend on the order in which the rays are traced,
so the ordered traversal must give, bit for bit, the results of the
unordered one (itself pinned to the oracle in test_gpu_trace.py), written at
each ray's own index:

  * closest hit and any hit on compressed and 4-wide traversal trees, for
    random rays, shadow rays from random surface points, and rays whose
    origins are NaN, infinite, all equal, or far outside the others (border
    cells of the grid);
  * ray counts of 1, 255, 257 and one past 2^20 (the origin sample's
    stride), and two calls in a row, the second smaller (the cached block
    reused); compressed trees read the rays through the order, 4-wide trees
    trace a sorted copy and gather the results back; one call of 2^24 + 2^22
    rays on a 4-wide tree, whose 64 bytes per ray exceed the cache's 1 GiB
    (allocated and freed per call);
  * closest hits of a small case against the oracle's brute force;
  * a plan with ordering on: eager enqueue and graph replays (moved vertices
    in between) equal the plan without it; more rays than it was set up for
    and multi-scene plans are refused; order_rays(0) returns to the caller's
    order;
  * the flag is refused where it does not apply (batches, MBVH calls);
  * eager ordered calls are stream-ordered: several in a row on one stream
    and on a second stream, with no synchronisation between them, reuse the
    cached block one after the other and each returns its own results.
"""
import numpy as np
import pytest

from oracle import refcpu as ref
from rtbvh_amd import scenes
from tests.helpers import as_rays, prep, ray_set, teapot

torch = pytest.importorskip("torch")
import rtbvh_amd as rt  # noqa: E402
from rtbvh_amd import _lib as L  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).view(dtype).cuda()


def _scene(verts):
    aabbs, centers = prep(verts)
    a_d = _dev(np.ascontiguousarray(aabbs).view(np.float32).reshape(len(verts), 8))
    c_d, v_d = _dev(centers), _dev(verts)
    return a_d, c_d, v_d


def _odd_rays(verts, n, seed):
    """Random rays with a few NaN, infinite, repeated and far-away origins."""
    r = ray_set(verts, n, seed=seed, kind="box")
    k = max(1, n // 50)
    rng = np.random.default_rng(seed)
    r[rng.integers(0, n, k), 0] = np.nan
    r[rng.integers(0, n, k), 1] = np.inf
    r[rng.integers(0, n, k), 2] = -np.inf
    r[rng.integers(0, n, k), 0:3] = np.float32(1e30)
    r[rng.integers(0, n, k), 0:3] = r[0, 0:3]
    return r


def _same_bits(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


@pytest.mark.parametrize("width", [8, 4])
def test_ordered_equals_unordered(width):
    print(f"[order_tree width={width}]", flush=True)
    verts = scenes.noisy_sphere(90, 70, seed=12)
    a_d, c_d, v_d = _scene(verts)
    bvh = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered()
    tree = bvh.trace_tree(v_d, width=width)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    for n in (1, 255, 257, (1 << 20) + 1, 70000, 3000):
        rays = _dev(_odd_rays(verts, n, seed=n))
        want = tree.intersect(rays)
        got = tree.intersect(rays, order_rays=True)
        assert _same_bits(got, want), n
        occ_want = tree.occluded(rays)
        occ = tree.occluded(rays, order_rays=True)
        assert torch.equal(occ, occ_want), n
    shadow = scenes.shadow_rays_torch(v_d, 200000, (3.0, 2.0, 3.0), "cuda", gen)
    want = tree.occluded(shadow)
    got = tree.occluded(shadow, order_rays=True)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    assert 0 < float(want.float().mean()) < 1
    assert _same_bits(tree.intersect(shadow, order_rays=True), tree.intersect(shadow))


def test_ordered_calls_in_stream_order():
    print("[order_streams]", flush=True)
    verts = scenes.noisy_sphere(90, 70, seed=14)
    a_d, c_d, v_d = _scene(verts)
    tree = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered().trace_tree(v_d)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(21)
    sets = [scenes.shadow_rays_torch(v_d, m, (2.0 + k, 3.0, 1.0 - k), "cuda", gen)
            for k, m in enumerate((300000, 120000, 300000, 50000))]
    want_occ = [tree.occluded(r) for r in sets]
    want_hit = [tree.intersect(r) for r in sets]
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    got = []
    for k, r in enumerate(sets):  # no synchronisation between the calls
        if k == 2:
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                got.append((tree.occluded(r, order_rays=True), tree.intersect(r, order_rays=True)))
        else:
            got.append((tree.occluded(r, order_rays=True), tree.intersect(r, order_rays=True)))
    torch.cuda.synchronize()
    for k in range(len(sets)):
        assert torch.equal(got[k][0], want_occ[k]), k
        assert _same_bits(got[k][1], want_hit[k]), k


def test_ordered_against_brute_force():
    print("[order_brute]", flush=True)
    verts = teapot()
    a_d, c_d, v_d = _scene(verts)
    tree = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered().trace_tree(v_d)
    rays = ray_set(verts, 3000, seed=8, kind="sphere")
    got = tree.intersect(_dev(rays), order_rays=True).cpu().numpy()
    want = ref.trace_brute(verts, as_rays(rays), threads=8)
    want = np.ascontiguousarray(want).view(np.uint32).reshape(len(rays), -1)
    assert np.array_equal(got.view(np.uint32), want)
    occ = tree.occluded(_dev(rays), order_rays=True).cpu().numpy()
    assert np.array_equal(occ.astype(bool), ref.occluded_brute(verts, as_rays(rays), threads=8))


def test_ordered_beyond_the_cached_block():
    """2^24 + 2^22 rays on a 4-wide tree (a sorted copy: 64 bytes per ray)
    need more than the 1 GiB the per-device cache holds for this kind: the
    call allocates and frees its own block."""
    print("[order_uncached]", flush=True)
    verts = scenes.random_soup(20000, 0.02, seed=6)
    a_d, c_d, v_d = _scene(verts)
    bvh = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered()
    tree = bvh.trace_tree(v_d, width=4)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    n = (1 << 24) + (1 << 22)
    rays = scenes.shadow_rays_torch(v_d, n, (0.5, 2.0, 0.5), "cuda", gen)
    want = tree.occluded(rays)
    got = tree.occluded(rays, order_rays=True)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    del rays, want, got
    torch.cuda.empty_cache()


def test_plan_with_ordered_rays():
    print("[order_plan]", flush=True)
    base = scenes.noisy_sphere(80, 60, seed=9)
    a_d, c_d, v_d = _scene(base)
    base_d = v_d.clone()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    n = 150000
    rays = scenes.shadow_rays_torch(v_d, n, (2.0, 3.0, 1.0), "cuda", gen)
    plain = rt.BatchPlan(c_d, [0, len(base)], v_d, aabbs=a_d)
    ordered = rt.BatchPlan(c_d, [0, len(base)], v_d, aabbs=a_d)
    ordered.order_rays(n)
    occ_p = torch.empty((n,), dtype=torch.uint8, device="cuda")
    occ_o = torch.empty((n,), dtype=torch.uint8, device="cuda")
    hit_p = plain.enqueue(c_d, v_d, rays, None, aabbs=a_d)
    hit_o = ordered.enqueue(c_d, v_d, rays, None, aabbs=a_d)
    torch.cuda.synchronize()
    assert _same_bits(hit_o, hit_p)
    for f in range(4):  # graph recorded on the first launch, replayed after
        v = base_d.view(-1, 3).clone()
        v[:, 1] += 0.01 * torch.sin(v[:, 0] * 3.0 + f)
        v_d.copy_(v.view_as(v_d))
        aabbs, centers = prep(v_d.cpu().numpy())
        a_d.copy_(_dev(np.ascontiguousarray(aabbs).view(np.float32).reshape(len(base), 8)))
        c_d.copy_(_dev(centers))
        plain.launch(c_d, v_d, rays, None, aabbs=a_d, any_hit=True, out=occ_p)
        ordered.launch(c_d, v_d, rays, None, aabbs=a_d, any_hit=True, out=occ_o)
        plain.status()
        ordered.status()
        torch.cuda.synchronize()
        assert torch.equal(occ_o, occ_p), f
    assert 0 < float(occ_p.float().mean()) < 1
    # more rays than the buffers hold
    more = torch.cat([rays, rays[:10]])
    with pytest.raises(L.RTBvhError) as ei:
        ordered.launch(c_d, v_d, more, None, aabbs=a_d, any_hit=True)
    assert ei.value.code == L.INVALID_ARGUMENT
    # back to the caller's order: the same flags again
    ordered.order_rays(0)
    ordered.launch(c_d, v_d, more, None, aabbs=a_d, any_hit=True)
    ordered.launch(c_d, v_d, rays, None, aabbs=a_d, any_hit=True, out=occ_o)
    ordered.status()
    torch.cuda.synchronize()
    assert torch.equal(occ_o, occ_p)


def test_ordering_refused_where_it_does_not_apply():
    print("[order_refused]", flush=True)
    vl = [teapot(), scenes.torus(40, 20)]
    verts = np.concatenate(vl)
    a_d, c_d, v_d = _scene(verts)
    offs = [0, len(vl[0]), len(verts)]
    plan = rt.BatchPlan(c_d, offs, v_d, aabbs=a_d)
    with pytest.raises(L.RTBvhError) as ei:
        plan.order_rays(1000)
    assert ei.value.code == L.INVALID_ARGUMENT
    batch = rt.SceneBatch(c_d, offs, aabbs=a_d)
    batch.attach_triangles(v_d)
    rays = _dev(ray_set(verts, 100, seed=1, kind="box"))
    roff = rt.SceneBatch.ray_offsets([50, 50], "cuda")
    out = torch.empty((100, 4), dtype=torch.float32, device="cuda")
    rc = L.lib().rtbvh_gpu_intersect_batch(batch._h, rays.data_ptr(), roff.data_ptr(), 100,
                                           out.data_ptr(), L.TRACE_ORDER_RAYS, None)
    assert rc == L.INVALID_ARGUMENT
    torch.cuda.synchronize()
    bvh = rt.Builder(c_d, aabbs=a_d).construct_locally_ordered_clustered()
    m = rt.Mbvh.from_bvh(bvh)
    m.attach_triangles(v_d)
    rc = L.lib().rtbvh_gpu_intersect_mbvh_prepared(m._h, rays.data_ptr(), 100, out.data_ptr(),
                                                   L.TRACE_ORDER_RAYS, None)
    assert rc == L.INVALID_ARGUMENT
