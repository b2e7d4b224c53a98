"""Eager builds take their scratch from a per-device workspace cache (one
block per builder kind, kept between builds).  Builds in a row that grow and
shrink the request, and builds from several host threads at once, each on
its own stream (one of them finds the block busy and allocates its own),
return the same trees as builds one at a time: PLOC and binned SAH node
arrays and indices, and the compressed traversal tree's hits, bit for bit.
"""
import threading

import numpy as np
import pytest

from rtbvh_amd import scenes
from tests.helpers import flat, prep, ray_set, teapot

torch = pytest.importorskip("torch")
import rtbvh_amd as rt  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


SCENES = [teapot(), scenes.random_soup(20000, 0.02, seed=11), scenes.torus(40, 20),
          scenes.noisy_sphere(120, 90, seed=12), scenes.random_soup(300, 0.1, seed=13)]


def _build(verts, kind):
    aabbs, centers = prep(verts)
    b = rt.Builder(_dev(centers), aabbs=_dev(flat(aabbs)))
    bvh = b.construct_locally_ordered_clustered() if kind == "ploc" else b.construct_binned_sah()
    v_d = _dev(verts)
    tree = bvh.trace_tree(v_d)
    rays = _dev(ray_set(verts, 3000, seed=5, kind="box")).view(torch.float32)
    hits = tree.intersect(rays)
    torch.cuda.current_stream().synchronize()
    return (bvh.nodes().cpu().numpy().view(np.uint32).copy(),
            bvh.indices().cpu().numpy().view(np.uint32).copy(),
            hits.cpu().numpy().view(np.uint32).copy())


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_builds_in_a_row_and_from_threads_equal_one_at_a_time():
    print("[workspace] sequential", flush=True)
    jobs = [(i, kind) for i in range(len(SCENES)) for kind in ("ploc", "binned")]
    want = {job: _build(SCENES[job[0]], job[1]) for job in jobs}
    # in a row, largest and smallest alternating (the cached block grows once)
    for job in sorted(jobs, key=lambda j: (len(SCENES[j[0]]) % 7, j[1])):
        assert _same(_build(SCENES[job[0]], job[1]), want[job]), job
    print("[workspace] threads", flush=True)
    errors = []

    def worker(seed):
        try:
            stream = torch.cuda.Stream()
            order = list(jobs)
            np.random.default_rng(seed).shuffle(order)
            with torch.cuda.stream(stream):
                for _ in range(2):
                    for job in order:
                        if not _same(_build(SCENES[job[0]], job[1]), want[job]):
                            errors.append((seed, job))
        except Exception as e:  # reported below, on the test's thread
            errors.append((seed, repr(e)))

    threads = [threading.Thread(target=worker, args=(s,)) for s in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors[:5]
