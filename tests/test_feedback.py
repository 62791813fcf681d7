"""Analysis feedback on the host (device = -1): the product's update from
aggregated records (svo_mm_apply_records) against the restated reference
bookkeeping (oracle/residency.py: RenderPassAnalyse's host half and
updateClientSideIncoreBuffer over per-pixel records), cycle after cycle with
a pool small enough to force evictions; and the incore buffer against the
stitch rule."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads


def _records_for_cycle(rng, model, headers, N, n_records):
    """Synthetic per-pixel analysis records over the model's resident
    treelets: inner-node stops and leaf hits inside resident treelets, and
    leaf hits on the anchors of non-resident children (asking for them)."""
    parent = headers[:, 4]
    # a few treelets in view per cycle, so that the others age and get evicted
    resident = sorted(model.treelet_slot)
    resident = [resident[int(i)] for i in rng.choice(len(resident), size=min(3, len(resident)), replace=False)]
    kids = [g for g in range(1, headers.shape[0]) if g not in model.treelet_slot and int(parent[g]) in resident]
    kind = np.zeros(n_records, np.uint8)
    err = np.zeros(n_records, np.float32)
    sub = np.zeros(n_records, np.uint32)
    node = np.zeros(n_records, np.uint32)
    for i in range(n_records):
        r = rng.random()
        e = np.float32(rng.normal())
        if r < 0.1:
            continue  # a miss
        if r < 0.55 and kids:
            g = kids[int(rng.integers(len(kids)))]
            p = int(parent[g])
            node[i] = model.treelet_slot[p] * N + int(headers[g, 5])  # the anchor leaf
            kind[i], err[i], sub[i] = 1, e, g
        else:
            g = resident[int(rng.integers(len(resident)))]
            node[i] = model.treelet_slot[g] * N + int(rng.integers(1, N))
            kind[i], err[i] = (2, np.float32(rng.integers(2))) if r < 0.8 else (1, e)
    return {"rec_kind": kind, "rec_err": err, "rec_sub": sub, "node": node}


def _device_aggregate(rec, N):
    """What the device tables hold after the records: per slot and per asked
    gid the largest error."""
    slots, gids = {}, {}
    for k, e, s, n in zip(rec["rec_kind"], rec["rec_err"], rec["rec_sub"], rec["node"]):
        if not k:
            continue
        sl = int(n) // N
        slots[sl] = max(slots.get(sl, -np.inf), float(e))
        if k == 1 and s:
            gids[int(s)] = max(gids.get(int(s), -np.inf), float(e))
    return slots, gids


@pytest.mark.parametrize("seed", [1, 2])
def test_apply_records_matches_reference_bookkeeping(terrain8, seed):
    from oracle import residency as R
    sc, _ = terrain8
    headers = sc.headers()
    N = sc.info["nodes_per_treelet"]
    num_slots = 24
    mm = svo.TreeletMemoryManagerHip(sc, num_slots * sc.info["treelet_bytes"], device=-1)
    model = R.Residency(headers, num_slots)
    rng = np.random.default_rng(seed)
    total_inserted = total_evicted = 0
    for cycle in range(120):
        rec = _records_for_cycle(rng, model, headers, N, 30)
        old, new = R.aggregate_records(rec, model.slot_gid, model.parent, N, sc.info["treelet_bytes"])
        ev0 = model.evictions
        inserted = model.update_client_side(old, new)
        slots, gids = _device_aggregate(rec, N)
        st = mm.applyRecords(list(slots), list(slots.values()), list(gids), list(gids.values()))
        assert st["visible_treelets"] == len(old), cycle
        assert st["requests"] == len({g for g in gids}), cycle
        assert st["requests_kept"] == len(new), cycle
        assert st["inserted"] == inserted, cycle
        assert st["evicted"] == model.evictions - ev0, cycle
        assert mm.getSlots() == model.slot_gid, f"slot tables differ after cycle {cycle}"
        total_inserted += inserted
        total_evicted += model.evictions - ev0
    assert total_inserted > num_slots and total_evicted > 0  # the pool filled up and evicted
    # the incore buffer of every resident slot follows the stitch rule
    nodes, _ = mm.getIncoreBuffer()
    exp = R.expected_incore(sc, mm.getSlots(), N)
    for s, g in enumerate(mm.getSlots()):
        if g >= 0:
            assert np.array_equal(nodes[s * N:(s + 1) * N], exp[s * N:(s + 1) * N]), f"slot {s} (treelet {g})"


def test_request_order_and_pool_limit(terrain8):
    sc, _ = terrain8
    headers = sc.headers()
    tb = sc.info["treelet_bytes"]
    cap = int((12.0 * 1048576.0) / tb)  # 3072 treelets of 4 KiB
    # every child of the root asked for at once: the pool, not the cap, limits
    mm = svo.TreeletMemoryManagerHip(sc, 64 * tb, device=-1)
    kids = [g for g in range(1, headers.shape[0]) if headers[g, 4] == 0]
    errs = np.linspace(-1, 1, len(kids)).astype(np.float32)[::-1]
    st = mm.applyRecords([0], [0.5], kids, errs)
    assert st["requests"] == len(kids) and st["requests_kept"] == min(cap, len(kids))
    assert st["inserted"] == min(cap, len(kids), 63)
    # ties are inserted in ascending gid order
    mm2 = svo.TreeletMemoryManagerHip(sc, 4 * tb, device=-1)
    st = mm2.applyRecords([0], [0.5], kids[:6][::-1], [1.0] * 6)
    assert st["inserted"] == 3 and mm2.getSlots()[1:] == sorted(kids[:6])[:3]


def test_apply_records_rejects_bad_input(terrain8):
    sc, _ = terrain8
    mm = svo.TreeletMemoryManagerHip(sc, 8 * sc.info["treelet_bytes"], device=-1)
    with pytest.raises(svo.SvoError):
        mm.applyRecords([], [], [0], [1.0])  # the root is never asked for
    with pytest.raises(svo.SvoError):
        mm.applyRecords([], [], [5, 5], [1.0, 2.0])
    with pytest.raises(svo.SvoError):
        mm.applyRecords([], [], [10 ** 6], [1.0])


def test_analysis_camera_is_a_coarser_frame_of_the_same_frustum():
    cam = workloads.terrain_camera(3, 1920, 1080)
    a = svo.analysis_camera(cam, 8)
    assert (a.width, a.height) == (240, 135)
    for k in range(3):
        assert a.pixel_h[k] == np.float32(cam.pixel_h[k] * np.float32(8.0))
        assert a.far_lower_left[k] == cam.far_lower_left[k]
    # the far plane's upper-right corner is the same (up to rounding)
    full = [cam.far_lower_left[k] + cam.pixel_h[k] * 1920 + cam.pixel_v[k] * 1080 for k in range(3)]
    coarse = [a.far_lower_left[k] + a.pixel_h[k] * 240 + a.pixel_v[k] * 135 for k in range(3)]
    assert np.allclose(full, coarse, atol=1e-4)
    j = svo.analysis_camera(cam, 8, (1.0, -1.0))
    for k in range(3):
        want = np.float32(np.float32(cam.far_lower_left[k] + np.float32(np.float32(8.0) * cam.pixel_h[k]) * np.float32(0.5))
                          + np.float32(np.float32(-8.0) * cam.pixel_v[k]) * np.float32(0.5))
        assert j.far_lower_left[k] == want
    with pytest.raises(svo.SvoError):
        svo.analysis_camera(cam, 0)
    with pytest.raises(svo.SvoError):
        svo.analysis_camera(cam, 8, (1.5, 0.0))
