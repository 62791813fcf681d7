"""Pins of the restated reference bookkeeping (oracle/residency.py) by
scenarios worked out by hand from src/TreeletMemoryManager.cpp and
src/RenderPassAnalyse.cpp, independent of the product manager:

  * freeIncorePosition (:560-653): a treelet is evictable when its slot and
    its parent's slot both have visibility < 0; the evicted treelet's
    resident descendants are freed first (pushed on the free-slot stack in
    breadth-first order), then the treelet itself, so the next insert takes
    the treelet's own slot; the root slot is never aged, so children of the
    root are never evicted;
  * the probe cursor persists between calls and one call tests at most 100
    slots;
  * performAnalysePass's host half (:248-356): the visible set propagates to
    the root with the largest error, the root's error is 99, requests are
    ordered by descending error (ties: ascending gid) and capped at
    12 MiB / treelet bytes.
"""
import numpy as np

from oracle import residency as R


def _headers(parents):
    h = np.zeros((len(parents), 8), np.uint32)
    h[:, 4] = parents
    return h


def test_eviction_frees_descendants_then_the_treelet():
    # 0 <- 1, 2, 3 ; 1 <- 4 ; 4 <- 5
    m = R.Residency(_headers([0, 0, 0, 0, 1, 4]), 4)
    assert m.slot_gid == [0, -1, -1, -1] and m.free == [3, 2, 1]
    for g in (1, 4, 5):
        assert m.insert(g, R.MAX_VISIBILITY, 1.0)
    assert m.slot_gid == [0, 1, 4, 5] and not m.free
    for _ in range(R.MAX_VISIBILITY + 1):
        assert m.update_client_side([], []) == 0
    assert m.vis == [R.MAX_VISIBILITY, -1, -1, -1]
    # slot 1 (treelet 1) is skipped: its parent is the never-aged root;
    # slot 2 (treelet 4, parent 1 at -2) is evicted with its child 5
    assert m.update_client_side([(0, 99.0)], [(2, 0.5)]) == 1
    assert m.evictions == 2
    assert m.slot_gid == [0, 1, 2, -1]
    assert m.free == [3] and m.search_pos == 2
    assert m.treelet_slot == {0: 0, 1: 1, 2: 2}
    assert m.vis[2] == R.MAX_VISIBILITY and m.err[2] == 0.5
    # no evictable slot left (only the root's children are resident)
    assert m.insert(3, R.MAX_VISIBILITY, 0.1) and m.slot_gid == [0, 1, 2, 3]
    for _ in range(20):
        m.update_client_side([], [])
    assert not m.insert(4, R.MAX_VISIBILITY, 0.1)


def test_probe_limit_and_persistent_cursor():
    # 0 <- 1 ; 1 <- 2..150 ; a 150-slot pool holds 0..149
    parents = [0, 0] + [1] * 149
    m = R.Residency(_headers(parents), 150)
    for g in range(1, 150):
        assert m.insert(g, R.MAX_VISIBILITY, 0.0)
    seen = [(g, 0.0) for g in range(2, 150) if g != 120]
    for _ in range(R.MAX_VISIBILITY + 1):
        m.update_client_side(seen, [])
    assert m.vis[1] == -1 and m.vis[120] == -1 and m.vis[2] == R.MAX_VISIBILITY
    # the only evictable slot is 120; the first call stops after 100 probes
    assert m.update_client_side(seen, [(150, 1.0)]) == 0
    assert m.search_pos == 101 and m.evictions == 0
    # the next call goes on from slot 101 and finds it
    assert m.update_client_side(seen, [(150, 1.0)]) == 1
    assert m.slot_gid[120] == 150 and m.evictions == 1 and m.search_pos == 120


def test_remove_pushes_the_slot_back():
    m = R.Residency(_headers([0, 0, 0]), 3)
    assert m.insert(1, 8, 0.0) and m.insert(2, 8, 0.0)
    m.remove(1)
    assert m.free == [1] and m.slot_gid == [0, -1, 2]
    assert m.insert(1, 8, 0.0) and m.slot_gid == [0, 1, 2]


def test_aggregate_records_by_hand():
    # slots 0..3 hold treelets 0, 1, 4, 2 ; 0 <- 1, 2 ; 1 <- 4 ; 4 <- 5, 6, 7
    parent = [0, 0, 0, 0, 1, 4, 4, 4]
    slot_gid = [0, 1, 4, 2]
    N = 16
    rec = {
        # kind: 0 miss, 1 leaf (sub = asked treelet or 0), 2 inner LOD stop
        "rec_kind": np.array([0, 2, 1, 1, 1, 1, 2, 1], np.uint8),
        "rec_err": np.array([5.0, 1.0, 0.25, 0.75, 0.25, -3.0, 0.0, 0.5], np.float32),
        "rec_sub": np.array([0, 0, 5, 6, 7, 0, 0, 5], np.uint32),
        "node": np.array([0, 2 * N + 3, 2 * N + 9, 2 * N + 10, 2 * N + 11, 3 * N + 1, 1 * N, 2 * N + 9], np.uint32),
    }
    old, new = R.aggregate_records(rec, slot_gid, parent, N, 4 << 20)
    # treelet 4's records (max 1.0) reach 1 and the root; treelet 2 has -3.0
    assert old == [(0, 99.0), (1, 1.0), (2, -3.0), (4, 1.0)]
    # requests: 6 (0.75), then 5 and ... 5 keeps its max 0.5, 7 has 0.25
    assert new == [(6, 0.75), (5, 0.5), (7, 0.25)]
    # the cap: 12 MiB of 8 MiB treelets is one request
    _, new = R.aggregate_records(rec, slot_gid, parent, N, 8 << 20)
    assert new == [(6, 0.75)]
    # ties keep ascending gid
    rec["rec_err"][:] = 0.5
    _, new = R.aggregate_records(rec, slot_gid, parent, N, 4 << 20)
    assert new == [(5, 0.5), (6, 0.5), (7, 0.5)]
