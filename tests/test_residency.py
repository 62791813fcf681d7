"""Residency manager on the host (device = -1): slot assignment, stitching,
removal, eviction errors and the merged-upload rule, checked against the
oracle's independent stitching and a restatement of the merge loop."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import _capi, workloads
from tests import helpers as H


@pytest.mark.parametrize("fixture", ["terrain8", "terrain8_bfs", "menger8"])
def test_resident_all_matches_oracle_stitching(request, fixture):
    import oracle
    sc, base = request.getfixturevalue(fixture)
    ib = workloads.incore_bytes_for(sc, spare_slots=2)
    mm = svo.TreeletMemoryManagerHip(sc, ib, device=-1, resident_all=True)
    o = oracle.OracleScene(base, ib).incore()
    nodes, attr = mm.getIncoreBuffer()
    assert np.array_equal(nodes, o["nodes"])
    assert np.array_equal(attr[:, 1], o["colors"]) and np.array_equal(attr[:, 0], o["normals"])
    slots = mm.getSlots()
    assert np.array_equal(np.array(slots, np.int32), o["slot_gid"])
    assert mm.info["resident_treelets"] == sc.info["num_treelets"]
    assert mm.info["free_slots"] == 2


def test_stitch_and_unstitch(terrain8_bfs):
    sc, _ = terrain8_bfs
    N = sc.info["nodes_per_treelet"]
    mm = svo.TreeletMemoryManagerHip(sc, 64 * sc.info["treelet_bytes"], device=-1)
    assert mm.getSlots()[0] == 0 and mm.info["resident_treelets"] == 1
    hdr = sc.headers()
    children = [int(g) for g in np.nonzero(hdr[:, 4] == 0)[0] if g > 0][:3]
    for k, g in enumerate(children):
        mm.insertTreeletIntoIncoreBuffer(g, 1.0)
        assert mm.getSlotGid(g) == k + 1  # the free-slot stack yields 1, 2, 3 ...
    nodes, _ = mm.getIncoreBuffer(0, 4 * N)
    root_nodes = sc.treelet(0)[1]
    for k, g in enumerate(children):
        h, tn, _ = sc.treelet(g)
        leaf, lp = h["parent_leaf_pos"], h["parent_leaf_parent_pos"]
        slot_pos = (k + 1) * N
        assert nodes[leaf, 0] == ((slot_pos + 1 - leaf) & 0xFFFFFFFF)
        assert nodes[leaf, 1] == tn[0, 1]
        assert not (nodes[lp, 1] >> h["parent_leaf_idx"]) & 1
        assert np.array_equal(nodes[slot_pos:slot_pos + h["num_nodes"]], tn)
        assert not nodes[slot_pos + h["num_nodes"]:slot_pos + N].any()
    # removal restores the pristine leaf and its leaf bit, frees the slot
    g = children[1]
    h = sc.treelet(g)[0]
    mm.removeTreeletFromIncoreBuffer(g)
    nodes, _ = mm.getIncoreBuffer(0, N)
    assert np.array_equal(nodes[h["parent_leaf_pos"]], root_nodes[h["parent_leaf_pos"]])
    assert (nodes[h["parent_leaf_parent_pos"], 1] >> h["parent_leaf_idx"]) & 1
    assert mm.getSlotGid(g) == -1 and mm.getSlots()[2] == -1
    # the freed slot is reused first
    mm.insertTreeletIntoIncoreBuffer(g)
    assert mm.getSlotGid(g) == 2


def test_remove_frees_resident_descendants(terrain8_bfs):
    sc, _ = terrain8_bfs
    mm = svo.TreeletMemoryManagerHip(sc, workloads.incore_bytes_for(sc), device=-1, resident_all=True)
    hdr = sc.headers()
    parents = hdr[1:, 4]
    g = int(np.bincount(parents).argmax())  # the treelet with the most sub-treelets
    if g == 0:
        g = int(parents[parents > 0][0])
    kids = [int(u) for u in np.nonzero(hdr[:, 4] == g)[0] if u > 0]
    assert kids
    before = mm.info["resident_treelets"]
    mm.removeTreeletFromIncoreBuffer(g)
    assert mm.getSlotGid(g) == -1 and all(mm.getSlotGid(u) == -1 for u in kids)
    assert mm.info["resident_treelets"] <= before - 1 - len(kids)


def test_residency_errors(terrain8_bfs):
    sc, _ = terrain8_bfs
    tb = sc.info["treelet_bytes"]
    mm = svo.TreeletMemoryManagerHip(sc, 3 * tb, device=-1)
    hdr = sc.headers()
    child = next(int(g) for g in range(1, len(hdr)) if hdr[g, 4] == 0 and np.any(hdr[:, 4] == g))
    grandchild = next(int(g) for g in range(1, len(hdr)) if hdr[g, 4] == child)
    with pytest.raises(svo.SvoError) as e:
        mm.insertTreeletIntoIncoreBuffer(grandchild)
    assert e.value.status == _capi.SVO_ERR_STATE  # parent not resident
    mm.insertTreeletIntoIncoreBuffer(child)
    with pytest.raises(svo.SvoError) as e:
        mm.insertTreeletIntoIncoreBuffer(child)
    assert e.value.status == _capi.SVO_ERR_STATE  # already resident
    mm.insertTreeletIntoIncoreBuffer(grandchild)
    other = next(int(g) for g in range(1, len(hdr)) if hdr[g, 4] == 0 and g != child)
    with pytest.raises(svo.SvoError) as e:
        mm.insertTreeletIntoIncoreBuffer(other)
    assert e.value.status == _capi.SVO_ERR_CAPACITY  # 3 slots, all visible
    with pytest.raises(svo.SvoError) as e:
        mm.insertTreeletIntoIncoreBuffer(10 ** 7)
    assert e.value.status == _capi.SVO_ERR_ARG
    with pytest.raises(svo.SvoError) as e:
        mm.removeTreeletFromIncoreBuffer(0)
    assert e.value.status == _capi.SVO_ERR_ARG
    mm.resetIncoreBuffer()
    assert mm.info["resident_treelets"] == 1 and mm.getSlotGid(child) == -1


def merged_ranges(dirty):
    """Range merging of TreeletMemoryManagerCl::updateDeviceMemory
    (src/TreeletMemoryManagerCl.cpp:175-213), FP32 ratio test."""
    it = sorted(dirty)
    limit = np.float32(0.6) * np.float32(2.0)
    i, out = 0, []
    while i < len(it):
        start, ns, nt = it[i], 1, 1
        while i < len(it):
            dist = it[i] - start
            if np.float32(nt + 1) / np.float32(dist + 1) >= limit:
                ns, nt, i = dist + 1, nt + 1, i + 1
            else:
                break
        out.append((start, ns))
    return out


def test_merge_rule_restatement_examples():
    assert merged_ranges([0]) == [(0, 1)]
    assert merged_ranges([0, 1, 2]) == [(0, 3)]
    assert merged_ranges([0, 2]) == [(0, 1), (2, 1)]  # (2+1)/(2+1) = 1.0 < 1.2
    assert merged_ranges([0, 1, 3]) == [(0, 2), (3, 1)]
    assert merged_ranges([4, 5, 6, 8]) == [(4, 3), (8, 1)]  # 5/5 = 1.0 < 1.2
    assert merged_ranges([0, 5]) == [(0, 1), (5, 1)]


def test_update_device_merges_like_the_reference(terrain8_bfs):
    sc, _ = terrain8_bfs
    tb, N = sc.info["treelet_bytes"], sc.info["nodes_per_treelet"]
    rng = np.random.default_rng(3)
    mm = svo.TreeletMemoryManagerHip(sc, 4000 * tb, device=-1)
    stats = mm.updateDeviceMemory()
    assert stats["slots_uploaded"] == 0 and stats["copy_calls"] == 0
    resident = {0}
    order = []
    hdr = sc.headers()
    for g in range(1, sc.info["num_treelets"]):
        if int(hdr[g, 4]) in resident and rng.random() < 0.3:
            resident.add(g)
            order.append(g)
        if len(order) == 60:
            break
    dirty = set()
    for g in order:
        mm.insertTreeletIntoIncoreBuffer(g)
        dirty |= {mm.getSlotGid(g), mm.getSlotGid(int(hdr[g, 4]))}
    assert mm.info["dirty_slots"] == len(dirty)
    ranges = merged_ranges(dirty)
    stats = mm.updateDeviceMemory()
    assert stats["slots_uploaded"] == len(dirty)
    assert stats["copy_calls"] == len(ranges)
    assert stats["bytes"] == sum(n for _, n in ranges) * N * 16
    assert mm.info["dirty_slots"] == 0
