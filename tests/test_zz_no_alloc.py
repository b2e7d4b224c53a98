"""rtbvh_gpu_create_bvh_noalloc: the PLOC build into caller arrays with the
library's static per-device workspace.  Its nodes and indices equal, bit for
bit, those of rtbvh_gpu_create_bvh (and of the oracle's LOCB) on two calls in
a row, the second smaller (and with point boxes, stride 16), so nothing left
in the shared workspace by one call leaks into the next; sizes beyond the
stated maximum are refused."""
import ctypes

import numpy as np
import pytest

from oracle import refcpu as ref
from rtbvh_amd import _lib as L
from rtbvh_amd import scenes
from tests.helpers import flat, prep

torch = pytest.importorskip("torch")
import rtbvh_amd as rt  # noqa: E402

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _noalloc(aabbs, centers, n):
    nodes = torch.full((max(1, 2 * n - 1), 8), -7.0, dtype=torch.float32, device="cuda")
    idx = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    rc = L.lib().rtbvh_gpu_create_bvh_noalloc(_p(aabbs), n, _p(centers), 4 * centers.shape[1],
                                               _p(nodes), _p(idx), None)
    assert rc == L.OK, L.lib().rtbvh_last_error()
    torch.cuda.synchronize()
    return (nodes.cpu().numpy().view(ref.AABB).reshape(-1),
            idx.cpu().numpy().view(np.uint32))


def _create(aabbs, centers):
    bvh = rt.Builder(centers, aabbs=aabbs).construct_locally_ordered_clustered()
    return (bvh.nodes().cpu().numpy().view(ref.AABB).reshape(-1),
            bvh.indices().cpu().numpy().view(np.uint32))


def test_noalloc_build_equals_create_bvh():
    print("[noalloc] two calls, the second smaller", flush=True)
    big = np.concatenate([scenes.noisy_sphere(300, 240, seed=21),
                          scenes.random_soup(100_000, 0.02, seed=22)])  # 244,000 triangles
    small = scenes.random_soup(3000, 0.05, seed=23)
    assert len(big) <= L.NOALLOC_MAX_PRIMS
    a_big, c_big = prep(big)
    a_b = torch.from_numpy(flat(a_big)).cuda()
    c_b = torch.from_numpy(c_big).cuda()
    got = _noalloc(a_b, c_b, len(big))
    want = _create(a_b, c_b)
    assert np.array_equal(flat(got[0]).view(np.uint32), flat(want[0]).view(np.uint32))
    assert np.array_equal(got[1], want[1])
    # second call: fewer primitives, null aabbs (point boxes), stride 16
    _, c_small = prep(small)
    c16 = np.concatenate([c_small, np.zeros((len(c_small), 1), np.float32)], 1)
    c_s = torch.from_numpy(np.ascontiguousarray(c16)).cuda()
    got = _noalloc(None, c_s, len(small))
    want = _create(None, c_s)
    assert np.array_equal(flat(got[0]).view(np.uint32), flat(want[0]).view(np.uint32))
    assert np.array_equal(got[1], want[1])
    pts = np.zeros(len(c_small), ref.AABB)
    pts["min"] = c_small
    pts["max"] = c_small
    o_nodes, o_idx = ref.build_locb(pts, c_small)
    assert np.array_equal(flat(got[0]).view(np.uint32), flat(o_nodes).view(np.uint32))
    # two primitives: the tree is one root leaf (locb.rs:258-269) in node 0
    got = _noalloc(a_b[:2].contiguous(), c_b[:2].contiguous(), 2)
    want = _create(a_b[:2].contiguous(), c_b[:2].contiguous())
    assert len(want[0]) == 1
    assert np.array_equal(flat(got[0][:1]).view(np.uint32), flat(want[0]).view(np.uint32))
    assert np.array_equal(got[1], want[1])


def test_noalloc_refuses_more_than_the_maximum():
    n = L.NOALLOC_MAX_PRIMS + 1
    c = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    nodes = torch.empty((2 * n - 1, 8), dtype=torch.float32, device="cuda")
    idx = torch.empty((n,), dtype=torch.int32, device="cuda")
    lib = L.lib()
    assert lib.rtbvh_gpu_create_bvh_noalloc(None, n, _p(c), 12, _p(nodes), _p(idx),
                                            None) == L.INVALID_ARGUMENT
    assert b"at most" in lib.rtbvh_last_error()
    assert lib.rtbvh_gpu_create_bvh_noalloc(None, 0, _p(c), 12, _p(nodes), _p(idx),
                                            None) == L.NO_PRIMITIVES
    assert lib.rtbvh_gpu_create_bvh_noalloc(None, 10, _p(c), 20, _p(nodes), _p(idx),
                                            None) == L.INVALID_ARGUMENT
