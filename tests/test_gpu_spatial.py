"""Spatial-split SAH builds (rtbvh_gpu_create_spatial_bvh).

The reference's SpatialSahBuilder (src/builders/spatial_sah.rs:247-1034)
cannot be run here and has defects that lose hits (its clipped boxes come
from a * t * (b - a), :91-94), so its node arrays are no oracle.  The bar is
the structure of a valid spatial-split tree and the hits:
  * every primitive is referenced, references stay within the split budget
    (0.75 n extra), the leaves partition the index array, child boxes lie in
    their parents;
  * on sliver-heavy scenes spatial splits do happen (more references than
    primitives) and lower the SAH cost against the binned-SAH tree;
  * traced through the reference-layout MBVH, the BVH itself and both
    traversal trees, the hits equal the oracle's brute force bit for bit.
"""
import numpy as np
import pytest

from oracle import refcpu as ref
from rtbvh_amd import scenes
from tests.helpers import (as_rays, flat, flat_grid, flat_grid_rays, geometric_chain, hits_equal,
                           prep, ray_set, tree_depth)

torch = pytest.importorskip("torch")
import rtbvh_amd as rt  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).view(dtype).cuda()


def slivers(n, length=0.5, width=0.004, seed=0):
    """Long thin triangles in the unit cube: the case spatial splits are for."""
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3))
    d = rng.normal(size=(n, 2, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    v1 = v0 + length * d[:, 0]
    v2 = v1 + width * d[:, 1]
    return np.ascontiguousarray(np.concatenate([v0, v1, v2], axis=1), dtype=np.float32)


def _build(verts, kind, ppl=None):
    aabbs, centers = prep(verts)
    b = rt.Builder(_dev(centers), aabbs=_dev(flat(aabbs)), primitives_per_leaf=ppl)
    if kind == "spatial":
        return b.construct_spatial_sah(_dev(verts))
    return b.construct_binned_sah()


def _check_structure(bvh, n):
    nodes = bvh.nodes().cpu().numpy().view(ref.AABB).reshape(-1)
    idx = bvh.indices().cpu().numpy().view(np.uint32)
    m = bvh.index_count
    assert n <= m <= n + int(np.float32(n) * np.float32(0.75))
    assert np.array_equal(np.unique(idx), np.arange(n, dtype=np.uint32))
    covered = np.zeros(m, np.int64)
    seen = np.zeros(len(nodes), bool)
    stack = [0]
    lo = flat(nodes)[:, 0:3].astype(np.float64)
    hi = flat(nodes)[:, 4:7].astype(np.float64)
    while stack:
        i = stack.pop()
        assert 0 <= i < len(nodes) and not seen[i]
        seen[i] = True
        c, lf = int(nodes[i]["count"]), int(nodes[i]["left_first"])
        if c >= 0:
            assert c >= 1 and 0 <= lf and lf + c <= m
            covered[lf:lf + c] += 1
        else:
            for ch in (lf, lf + 1):
                assert np.all(lo[ch] >= lo[i]) and np.all(hi[ch] <= hi[i]), (i, ch)
                stack.append(ch)
    assert np.all(covered == 1)
    return nodes, idx


def _sah(nodes):
    """Surface-area cost (traversal 1, intersection 1) relative to the root."""
    f = flat(nodes).astype(np.float64)
    d = f[:, 4:7] - f[:, 0:3]
    area = (d[:, 0] + d[:, 1]) * d[:, 2] + d[:, 0] * d[:, 1]
    cost, stack = 0.0, [0]
    while stack:
        i = stack.pop()
        c, lf = int(nodes[i]["count"]), int(nodes[i]["left_first"])
        if c >= 0:
            cost += area[i] * c
        else:
            cost += area[i]
            stack += [lf, lf + 1]
    return cost / area[0]


def _hits_all_paths(bvh, verts, rays8):
    dv, dr = _dev(verts), _dev(rays8)
    want = ref.trace_brute(verts, as_rays(rays8), threads=8)
    got = {
        "bvh": bvh.intersect(dv, dr),
        "mbvh": rt.Mbvh.from_bvh(bvh).intersect(dv, dr),
        "tree8": bvh.trace_tree(dv, width=8).intersect(dr),
        "tree4": bvh.trace_tree(dv, width=4).intersect(dr),
    }
    for name, g in got.items():
        g = g.cpu().numpy().view(ref.HIT).reshape(-1)
        same = hits_equal(g, want)
        for i in np.flatnonzero(~same)[:3]:
            print(f"  {name} ray {i}: got {g[i]} want {want[i]}", flush=True)
        assert same.all(), name
    occ = rt.Mbvh.from_bvh(bvh).occluded(dv, dr).cpu().numpy().astype(bool)
    assert np.array_equal(occ, want["prim"] != 0xFFFFFFFF)
    return want


@pytest.mark.parametrize("scene", ["slivers", "mixed", "sphere"])
def test_spatial_structure_quality_and_hits(scene):
    print(f"[spatial] scene={scene}", flush=True)
    if scene == "slivers":
        verts = slivers(6000, seed=3)
    elif scene == "mixed":
        verts = np.concatenate([scenes.noisy_sphere(60, 40, seed=4), slivers(2000, seed=5)])
    else:
        verts = scenes.noisy_sphere(80, 60, seed=6)
    n = len(verts)
    bvh = _build(verts, "spatial")
    assert bvh.primitive_count == n
    nodes, _ = _check_structure(bvh, n)
    sah_s = _sah(nodes)
    binned = _build(verts, "binned")
    sah_b = _sah(binned.nodes().cpu().numpy().view(ref.AABB).reshape(-1))
    print(f"  references {bvh.index_count} for {n} triangles; SAH spatial {sah_s:.2f} vs "
          f"binned {sah_b:.2f}", flush=True)
    if scene != "sphere":
        assert bvh.index_count > n  # spatial splits happened
        assert sah_s < sah_b
    else:
        assert sah_s <= sah_b * 1.05
    rays8 = ray_set(verts, 20000, seed=7, kind="box")
    rays8[1::3, 7] = np.float32(0.2)
    want = _hits_all_paths(bvh, verts, rays8)
    assert (want["prim"] != 0xFFFFFFFF).sum() > len(rays8) // 20


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 200])
def test_spatial_small(n):
    print(f"[spatial_small] n={n}", flush=True)
    verts = slivers(n, length=0.3, seed=100 + n)
    bvh = _build(verts, "spatial")
    _check_structure(bvh, n)
    rays8 = ray_set(verts, 3000, seed=n, kind="box")
    _hits_all_paths(bvh, verts, rays8)


def test_spatial_leaf_size_and_refit():
    """primitives_per_leaf caps the leaves the SAH may keep; a refit with the
    primitives' own boxes (looser than the clipped ones) keeps the hits."""
    print("[spatial_leaf_refit]", flush=True)
    verts = np.concatenate([slivers(1500, seed=11), scenes.random_soup(1500, 0.03, seed=12)])
    bvh = _build(verts, "spatial", ppl=4)
    nodes, _ = _check_structure(bvh, len(verts))
    assert int(nodes["count"][nodes["count"] >= 0].max()) >= 1
    rays8 = ray_set(verts, 8000, seed=13, kind="box")
    _hits_all_paths(bvh, verts, rays8)
    aabbs, _ = prep(verts)
    bvh.refit(_dev(flat(aabbs)))
    _check_structure(bvh, len(verts))
    dv, dr = _dev(verts), _dev(rays8)
    want = ref.trace_brute(verts, as_rays(rays8), threads=8)
    got = bvh.intersect(dv, dr).cpu().numpy().view(ref.HIT).reshape(-1)
    assert hits_equal(got, want).all()


def test_spatial_duplicates_degenerates_and_flat():
    """Exact duplicates, zero-area triangles, triangles collapsed to a point,
    and a flat scene (one axis of zero extent) hit by axis-parallel rays:
    the spatial build stays valid and its hits equal brute force."""
    print("[spatial_degenerate]", flush=True)
    rng = np.random.default_rng(81)
    soup = slivers(400, length=0.4, seed=82)
    dup = np.repeat(soup[:60], 4, axis=0)
    line = soup[60:100].copy()
    line[:, 6:9] = line[:, 3:6]
    point = np.tile(np.float32([0.5, 0.5, 0.5]), (30, 3))
    verts = np.concatenate([dup, soup[100:], line, point])
    verts = verts[rng.permutation(len(verts))].astype(np.float32)
    bvh = _build(verts, "spatial")
    _check_structure(bvh, len(verts))
    _hits_all_paths(bvh, verts, ray_set(verts, 8000, seed=83, kind="box"))
    flatv = flat_grid(g=30, offset=0.25)
    bvh = _build(flatv, "spatial")
    _check_structure(bvh, len(flatv))
    want = _hits_all_paths(bvh, flatv, flat_grid_rays(flatv, 4000))
    assert (want["prim"] != 0xFFFFFFFF).sum() > 400


@pytest.mark.parametrize("ppl", [None, 4])
def test_spatial_deep_chain(ppl):
    """A tree deeper than the first batch of levels the builder enqueues
    (2 lg n + 2) before it reads a level record: the later batches take
    their grids from the record read last."""
    print(f"[spatial_chain] ppl={ppl}", flush=True)
    verts = geometric_chain()
    bvh = _build(verts, "spatial", ppl=ppl)
    nodes, _ = _check_structure(bvh, len(verts))
    assert tree_depth(nodes) > 2 * int(np.ceil(np.log2(len(verts)))) + 2
    _hits_all_paths(bvh, verts, ray_set(verts, 4000, seed=91, kind="box"))


def test_spatial_errors():
    verts = slivers(10, seed=1)
    aabbs, centers = prep(verts)
    b = rt.Builder(_dev(centers), aabbs=_dev(flat(aabbs)))
    with pytest.raises(ValueError):
        b.construct_spatial_sah(_dev(verts[:5]))
    with pytest.raises(rt.BuildError):
        rt.Builder(_dev(centers[:0])).construct_spatial_sah(_dev(verts[:0]))
