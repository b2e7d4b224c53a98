"""The reference's own C ABI, end to end through the library: create_bvh ->
create_mbvh -> intersect / intersect_packet / intersect_mbvh /
intersect_mbvh_packet -> refit -> free_*, with host arrays in and host
mirrors out, exactly as an existing rtbvh_ffi caller uses it.

Ports of the reference's FFI tests (rtbvh_ffi/src/lib.rs:868-943
`create_delete`, :945-1019 `intersect` with its callback :1027-1060), then the
same calls on larger scenes, checked against the oracle:
  * the mirrored node arrays: PLOC bit-identical to the oracle's LOCB, binned
    SAH equal in topology and bounds to the oracle's builder with its two box
    defects fixed, for center strides 12 and 16, with and without aabbs;
  * create_mbvh: structurally identical to the oracle's Mbvh::construct of
    the mirrored nodes;
  * the callback traversals: the same final t as the oracle's
    reference-order traversal of the mirrored arrays, and the first candidate
    the library offered at that t is the oracle's hit (the first accepted
    candidate wins, examples/benchmark.rs:33-37); the packet calls give every
    lane its brute-force closest t;
  * refit: the mirror is updated in place (same pointers) and equals the
    oracle's Bvh::refit bit for bit.
"""
import ctypes

import numpy as np
import pytest

from oracle import refcpu as ref
from rtbvh_amd import _lib as L
from rtbvh_amd import scenes
from tests.helpers import as_rays, flat, prep, ray_set, teapot

pytestmark = pytest.mark.gpu

RAY_T_MIN = np.float32(1e-4)


def _aabb_macro(tris):
    """aabb!(v0, v1, v2) (src/aabb.rs:472-482): grow from empty, then pad
    by 1e-4."""
    v = tris.reshape(-1, 3, 3)
    a = np.zeros(len(v), ref.AABB)
    a["min"] = v.min(1) - np.float32(1e-4)
    a["max"] = v.max(1) + np.float32(1e-4)
    return a


def _centers_of(aabbs):
    """Aabb::center (src/aabb.rs:326-329): (min + max) * 0.5."""
    return ((aabbs["min"] + aabbs["max"]) * np.float32(0.5)).astype(np.float32)


def _strided(centers, stride):
    if stride == 12:
        return np.ascontiguousarray(centers, dtype=np.float32)
    return np.ascontiguousarray(np.concatenate(
        [centers, np.zeros((len(centers), 1), np.float32)], 1), dtype=np.float32)


def _create(aabbs, centers, stride, ppl, kind):
    c = _strided(centers, stride)
    out = L.RTBvh()
    rc = L.lib().create_bvh(None if aabbs is None else aabbs.ctypes.data, len(c), c.ctypes.data,
                            stride, ppl, kind, ctypes.byref(out))
    return rc, out


def _mirror(handle, dtype, size):
    """The library-owned host arrays a handle points at."""
    nodes = np.frombuffer(ctypes.string_at(handle.nodes, handle.node_count * size),
                          dtype=dtype).copy()
    idx = np.frombuffer(ctypes.string_at(handle.indices, handle.index_count * 4),
                        dtype=np.uint32).copy()
    return nodes, idx


def _mt_callback(verts, o, d, t_min, calls=None):
    """intersect_test (rtbvh_ffi/src/lib.rs:1027-1060), generalised to the
    ray's t_min, using the oracle's triangle test; records the candidates."""

    def cb(prim, t_ptr, _user):
        if calls is not None:
            calls.append(int(prim))
        ok, t, _, _ = ref.mt(verts[prim], o, d, t_min, t_ptr[0])
        if ok:
            t_ptr[0] = t
        return False

    return L.CALLBACK(cb)


# ------------------------------------------------------- reference tests --
def test_create_delete():
    """rtbvh_ffi/src/lib.rs:868-943: 27 triangles from a 10 x 10 point grid.
    (The reference passes 12-byte Vec3 centres with stride 16, which reads
    past its array; here the centres are laid out with the stride used.)"""
    print("[ffi_create_delete]", flush=True)
    grid = np.array([[x, y, 0] for x in range(10) for y in range(10)], np.float32)
    tris = grid[:81].reshape(27, 9)
    aabbs = _aabb_macro(tris)
    centers = _centers_of(aabbs)
    lib = L.lib()
    out = L.RTBvh()
    assert lib.create_bvh(None, 27, None, 16, 1, L.BINNED_SAH, ctypes.byref(out)) == L.ERROR
    rc, bvh = _create(None, centers, 16, 1, L.BINNED_SAH)
    assert rc == L.OK and bvh.index_count == 27 and 27 <= bvh.node_count <= 53
    lib.free_bvh(bvh)
    rc, bvh = _create(aabbs, centers, 16, 1, L.BINNED_SAH)
    assert rc == L.OK
    m = L.RTMbvh()
    assert lib.create_mbvh(bvh, ctypes.byref(m)) == L.OK
    assert m.index_count == 27 and m.node_count >= 1
    lib.free_bvh(bvh)
    lib.free_mbvh(m)


def test_intersect_kat():
    """rtbvh_ffi/src/lib.rs:945-1019: a quad of two triangles at z = 1,
    stride 12 (size_of::<Vec3>()), binned SAH; a ray from the origin along +z
    with t = 1e26 ends with t == 1 through intersect and intersect_mbvh."""
    print("[ffi_intersect_kat]", flush=True)
    v = np.array([[-1, -1, 1, 1, -1, 1, 1, 1, 1], [-1, -1, 1, 1, 1, 1, -1, 1, 1]], np.float32)
    aabbs = _aabb_macro(v)
    rc, bvh = _create(aabbs, _centers_of(aabbs), 12, 1, L.BINNED_SAH)
    assert rc == L.OK
    lib = L.lib()
    m = L.RTMbvh()
    assert lib.create_mbvh(bvh, ctypes.byref(m)) == L.OK
    o = np.zeros(3, np.float32)
    d = np.array([0, 0, 1], np.float32)
    for fn, h in ((lib.intersect, bvh), (lib.intersect_mbvh, m)):
        t = np.array([1e26], np.float32)
        cb = _mt_callback(v, o, d, 1e-5)
        assert fn(h, o.ctypes.data, d.ctypes.data, t.ctypes.data, None, cb) == L.OK
        assert abs(float(t[0]) - 1.0) < np.finfo(np.float32).eps
    lib.free_bvh(bvh)
    lib.free_mbvh(m)


# ----------------------------------------------- the same calls, at scale --
SCENES = {
    "teapot": teapot,
    "soup": lambda: scenes.random_soup(2500, 0.05, seed=31),
}


def _oracle_tree(aabbs, centers, kind, ppl):
    if kind == L.LOCALLY_ORDERED_CLUSTERED:
        return ref.build_locb(aabbs, centers)
    return ref.build_binned_sah(aabbs, centers, prims_per_leaf=ppl, contained=True)


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("kind", [L.LOCALLY_ORDERED_CLUSTERED, L.BINNED_SAH])
@pytest.mark.parametrize("stride", [12, 16])
@pytest.mark.parametrize("with_aabbs", [True, False])
def test_create_bvh_mirrors_match_oracle(scene, kind, stride, with_aabbs):
    print(f"[ffi_create] scene={scene} kind={kind} stride={stride} aabbs={with_aabbs}",
          flush=True)
    verts = SCENES[scene]()
    aabbs, centers = prep(verts)
    if not with_aabbs:  # null aabbs: point boxes at the centres (in order)
        aabbs = np.zeros(len(centers), ref.AABB)
        aabbs["min"] = centers
        aabbs["max"] = centers
    lib = L.lib()
    for ppl in ((1,) if kind == L.LOCALLY_ORDERED_CLUSTERED else (1, 4)):
        rc, bvh = _create(aabbs if with_aabbs else None, centers, stride, ppl, kind)
        assert rc == L.OK, L.lib().rtbvh_last_error()
        nodes, idx = _mirror(bvh, ref.AABB, 32)
        want_n, want_i = _oracle_tree(aabbs, centers, kind, ppl)
        assert bvh.index_count == len(verts) and bvh.node_count == len(want_n)
        if kind == L.LOCALLY_ORDERED_CLUSTERED:
            assert np.array_equal(flat(nodes).view(np.uint32), flat(want_n).view(np.uint32))
            assert np.array_equal(idx, want_i)
        else:
            assert ref.compare_bvh(nodes, idx, want_n, want_i, as_sets=True) == 0
        m = L.RTMbvh()
        assert lib.create_mbvh(bvh, ctypes.byref(m)) == L.OK
        mn, mi = _mirror(m, ref.MBVH, 128)
        want_m = ref.build_mbvh(nodes)
        assert m.node_count == len(want_m) and np.array_equal(mi, idx)
        assert ref.compare_mbvh(mn, mi, want_m, idx) == 0
        lib.free_mbvh(m)
        lib.free_bvh(bvh)


@pytest.mark.parametrize("kind", [L.LOCALLY_ORDERED_CLUSTERED, L.BINNED_SAH])
def test_callback_traversal_refit_and_free(kind):
    """intersect / intersect_mbvh over the library-built mirrors equal the
    oracle's reference-order traversal of the same arrays (t and the
    accepted primitive); the packet calls give each lane the brute-force
    closest t; refit updates the mirror in place to the oracle's refit; the
    traversals of the refitted tree follow the moved triangles."""
    print(f"[ffi_traverse] kind={kind}", flush=True)
    verts = teapot()
    aabbs, centers = prep(verts)
    lib = L.lib()
    rc, bvh = _create(aabbs, centers, 16, 1, kind)
    assert rc == L.OK
    m = L.RTMbvh()
    assert lib.create_mbvh(bvh, ctypes.byref(m)) == L.OK
    nodes_ptr, idx_ptr = bvh.nodes, bvh.indices

    def check(vv, seed):
        nodes, idx = _mirror(bvh, ref.AABB, 32)
        mn, mi = _mirror(m, ref.MBVH, 128)
        rays8 = ray_set(vv, 120, seed=seed, kind="box")
        rays = as_rays(rays8)
        want_b = ref.trace_bvh(nodes, idx, vv, rays, mode=ref.MODE_VISIT_ORDER)
        want_m = ref.trace_mbvh(mn, mi, vv, rays, mode=ref.MODE_VISIT_ORDER)
        brute = ref.trace_brute(vv, rays)
        for i, r in enumerate(rays8):
            o, d = r[0:3].copy(), r[4:7].copy()
            for fn, h, want in ((lib.intersect, bvh, want_b), (lib.intersect_mbvh, m, want_m)):
                t = np.array([r[7]], np.float32)
                calls = []
                cb = _mt_callback(vv, o, d, float(r[3]), calls)
                assert fn(h, o.ctypes.data, d.ctypes.data, t.ctypes.data, None, cb) == L.OK
                assert t[0] == want["t"][i]
                if want["prim"][i] != L.MISS:
                    # the accepted candidate is the first one at the final t
                    first = next(p for p in calls
                                 if ref.mt(vv[p], o, d, float(r[3]), np.inf)[1] == t[0]
                                 and ref.mt(vv[p], o, d, float(r[3]), np.inf)[0])
                    assert first == want["prim"][i]
        # packets: each lane's closest t is the brute-force one
        for fn, h in ((lib.intersect_packet, bvh), (lib.intersect_mbvh_packet, m)):
            for p in range(0, 40, 4):
                r = rays8[p:p + 4]
                cols = [np.ascontiguousarray(r[:, k]) for k in (0, 1, 2, 4, 5, 6)]
                t = np.ascontiguousarray(r[:, 7])

                def cb(prim, t_ptr, _u, r=r):
                    for lane in range(4):
                        ok, tt, _, _ = ref.mt(vv[prim], r[lane, 0:3], r[lane, 4:7], RAY_T_MIN,
                                              t_ptr[lane])
                        if ok:
                            t_ptr[lane] = tt
                    return False

                c = L.CALLBACK(cb)
                assert fn(h, *[x.ctypes.data for x in cols], t.ctypes.data, None, c) == L.OK
                assert np.array_equal(t, brute["t"][p:p + 4])

    check(verts, 41)
    # refit with moved triangles: in place, the oracle's Bvh::refit bit for bit
    before, idx = _mirror(bvh, ref.AABB, 32)
    moved = verts + np.random.default_rng(7).normal(scale=0.01, size=verts.shape).astype(
        np.float32)
    new_aabbs, _ = prep(moved)
    assert lib.refit(new_aabbs.ctypes.data, bvh) == L.OK
    assert bvh.nodes == nodes_ptr and bvh.indices == idx_ptr
    after, idx2 = _mirror(bvh, ref.AABB, 32)
    want = ref.refit(before, idx, new_aabbs)
    assert np.array_equal(flat(after).view(np.uint32), flat(want).view(np.uint32))
    assert np.array_equal(idx2, idx)
    # the MBVH made before the refit is rebuilt from the refitted tree
    lib.free_mbvh(m)
    m = L.RTMbvh()
    assert lib.create_mbvh(bvh, ctypes.byref(m)) == L.OK
    check(moved, 43)
    lib.free_mbvh(m)
    lib.free_bvh(bvh)
    # a second free is reported on stderr, not fatal (rtbvh_ffi/src/lib.rs:838-849)
    lib.free_bvh(bvh)
