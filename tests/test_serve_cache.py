"""Host logic of the serving loop: byte-budget LRU with pinning, first come
first served admission, the request stream (no GPU)."""

import numpy as np
import pytest

from del_msh_dlpack.Serve.cache import ByteLRU, Request, admit


def test_lru_evicts_oldest_unpinned_first():
    c = ByteLRU(100)
    for k in range(4):
        c.insert(k, 25, None)
    c.touch(0)           # order now 1, 2, 3, 0
    c.pin(1)
    assert c.reserve(30)  # needs 30: evicts 2 (oldest unpinned) then 3
    assert 1 in c and 0 in c and 2 not in c and 3 not in c
    assert c.used == 50 and c.evicted == 2


def test_lru_refuses_when_pinned_bytes_block():
    c = ByteLRU(100)
    c.insert(0, 60, None, pins=1)
    c.insert(1, 30, None)
    assert not c.reserve(50)          # 30 evictable + 10 free < 50
    assert 1 in c and c.used == 90    # nothing evicted on refusal
    c.unpin(0)
    assert c.reserve(50) and 0 not in c


def _req(k, mid, side=10):
    return Request(k, 0.0, mid, side, side, np.eye(4))


def test_admit_fifo_with_budget_and_step_limits():
    c = ByteLRU(100)
    c.insert(7, 40, "cached")
    size = {7: 40, 1: 50, 2: 50, 3: 5}
    queue = [_req(0, 7), _req(1, 1), _req(2, 2), _req(3, 1), _req(4, 3)]
    batch, new_ids, planned = admit(queue, c, size.__getitem__, max_requests=10,
                                    max_pixels=10_000)
    # 7 is cached (pinned); 1 fits (50 + 40 = 90); 2 does not (nothing evictable);
    # the second request for 1 rides along; 3 fits in the last 10 bytes
    assert [r.rid for r in batch] == [0, 1, 3, 4]
    assert new_ids == [1, 3] and planned == {1: 2, 3: 1}
    assert [r.rid for r in queue] == [2]
    assert c.entries[7].pins == 1
    assert c.used == 40  # new trees are added by insert() after their build
    c.insert(1, 50, "t1", pins=planned[1])
    c.insert(3, 5, "t3", pins=planned[3])
    assert c.used == 95


def test_admit_respects_step_pixels():
    c = ByteLRU(1000)
    queue = [_req(k, k, side=100) for k in range(5)]  # 10_000 px each
    batch, _, _ = admit(queue, c, lambda m: 10, max_requests=10, max_pixels=25_000)
    assert [r.rid for r in batch] == [0, 1]
    assert len(queue) == 3


def test_request_stream_is_seeded():
    import del_msh_dlpack.Serve.torch as Serve

    a = Serve.request_stream(50, rate=1000.0, num_ids=100, seed=3)
    b = Serve.request_stream(50, rate=1000.0, num_ids=100, seed=3)
    assert [r.mesh_id for r in a] == [r.mesh_id for r in b]
    assert all(np.array_equal(x.ndc2world, y.ndc2world) for x, y in zip(a, b))
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) >= 0) and 0.02 < t[-1] < 0.2
    sides = np.array([r.width for r in a])
    assert sides.min() >= 256 and sides.max() <= 1024
    # Zipf: the most popular id shows up more than once in 50 draws
    assert np.bincount([r.mesh_id for r in a]).max() > 1


def test_catalog_meshes_are_fixed_by_id():
    import del_msh_dlpack.Serve.torch as Serve

    cat = Serve.MeshCatalog(tri_range=(100, 2000), seed=1)
    t0, v0 = cat.get(5)
    cat2 = Serve.MeshCatalog(tri_range=(100, 2000), seed=1)
    t1, v1 = cat2.get(5)
    assert np.array_equal(t0, t1) and np.array_equal(v0, v1)
    assert 50 <= t0.shape[0] <= 4000
