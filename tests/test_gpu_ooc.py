"""GPU: out-of-core residency driven by device-side analysis feedback.

Each frame: analyse pass on the GPU (records aggregated in device tables),
svo_mm_apply_feedback (visibility, page-ins, evictions), upload, render.  The
same frame is replayed on the CPU: the oracle traverses the analysis frame
over the manager's own incore buffer, the restated reference bookkeeping
(oracle/residency.py) aggregates its per-pixel records and updates its slot
table.  Slot tables, page-in counts and rendered pixels must agree exactly."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _same(a, b, what):
    for k in ("node", "treelet", "rgba"):
        bad = np.count_nonzero(a[k] != b[k])
        assert bad == 0, f"{what}: {k} differs on {bad} pixels"
    bad = np.count_nonzero(a["t"].view(np.uint32) != b["t"].view(np.uint32))
    assert bad == 0, f"{what}: t bits differ on {bad} pixels"


def _oracle(mm, cam, policy="treelet", records=False):
    import oracle
    n = mm.info["pool_nodes"]
    nodes, attr = mm.getIncoreBuffer(0, n)
    return oracle.render_buffers(nodes, attr[:, 1], np.asarray(mm.getSlots(), np.int32), mm.getNumNodesPerTreelet(),
                                 mm.scene.info["max_depth"], cam, policy, records=records)


def _path(n, w, h, lod):
    return [workloads.terrain_camera(500 + i, w, h, lod=lod, terrain_seed=workloads.TERRAIN_SEED) for i in range(n)]


def _asked(rec):
    k, s = rec["rec_kind"].ravel(), rec["rec_sub"].ravel()
    return len(set(s[(k == 1) & (s != 0)].tolist()))


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_analyse_apply_render_matches_reference_bookkeeping(terrain8, policy):
    from oracle import residency as R
    sc, _ = terrain8
    headers = sc.headers()
    N, tb = sc.info["nodes_per_treelet"], sc.info["treelet_bytes"]
    num_slots = 48  # of ~1700 treelets: the views below overflow it
    mm = svo.TreeletMemoryManagerHip(sc, num_slots * tb, device=0)
    model = R.Residency(headers, num_slots)
    rng = np.random.default_rng(17)
    s = svo.Stream()
    pages = evicted = 0
    for f, cam in enumerate(_path(24, 320, 200, lod=0.25)):
        jit = tuple(float(v) for v in rng.uniform(-1, 1, 2))
        acam = svo.analysis_camera(cam, 4, jit)
        rec = _oracle(mm, acam, policy, records=True)
        old, new = R.aggregate_records(rec, model.slot_gid, model.parent, N, tb)
        ev0 = model.evictions
        inserted = model.update_client_side(old, new)
        _say(f"frame {f} policy={policy}: oracle asks {_asked(rec)}, inserts {inserted}")
        mm.performAnalysePass(cam, 4, jit, policy, s)
        if f % 2:  # a readback taken ahead (fetchFeedback) is the one later calls use
            mm.fetchFeedback(s)
        vs = mm.readFeedback(s)  # RenderPassAnalyse's visible sets, then applied below
        assert vs["old"] == [(g, float(np.float32(e))) for g, e in old], f"visible old set, frame {f}"
        assert vs["new"] == [(g, float(np.float32(e))) for g, e in new], f"visible new set, frame {f}"
        st = mm.updateClientSideIncoreBuffer(s)
        assert st["visible_treelets"] == len(old), f
        assert st["requests"] == _asked(rec), f
        assert st["inserted"] == inserted and st["evicted"] == model.evictions - ev0, f
        assert mm.getSlots() == model.slot_gid, f"slot tables differ after frame {f}"
        mm.updateDeviceMemory(s)
        if f % 4 == 3:  # the upload kernel left HBM equal to the host incore buffer
            n = mm.info["pool_nodes"]
            hn, ha = mm.getIncoreBuffer(0, n)
            dn, da = mm.getDeviceIncoreBuffer(0, n)
            assert np.array_equal(hn, dn) and np.array_equal(ha, da), f"device incore differs after frame {f}"
        gpu = svo.render(mm, cam, policy)
        _same(gpu, _oracle(mm, cam, policy), f"frame {f}")
        pages += inserted
        evicted += model.evictions - ev0
    _say(f"page-ins {pages}, evictions {evicted}, resident {mm.info['resident_treelets']} of {sc.info['num_treelets']}")
    assert pages > num_slots and evicted > 0


def test_feedback_recorded_during_a_render(terrain8):
    """A view with feedback_stride k records exactly the analysis records of
    its pixels (x % k, y % k) == (0, 0)."""
    from oracle import residency as R
    sc, _ = terrain8
    headers = sc.headers()
    N, tb = sc.info["nodes_per_treelet"], sc.info["treelet_bytes"]
    num_slots = max(8, sc.info["num_treelets"] // 3)
    mm = svo.TreeletMemoryManagerHip(sc, num_slots * tb, device=0)
    model = R.Residency(headers, num_slots)
    r = svo.Renderer(0)
    s = svo.Stream()
    for f, cam in enumerate(_path(6, 256, 160, lod=0.25)):
        frame = svo.Frame.alloc(cam.width, cam.height)
        rec = _oracle(mm, cam, "treelet", records=True)
        sub = {k: v[::3, ::3] for k, v in rec.items()}
        old, new = R.aggregate_records(sub, model.slot_gid, model.parent, N, tb)
        inserted = model.update_client_side(old, new)
        r.render([svo.View(mm, cam, frame, feedback_stride=3)], "treelet", s)
        st = mm.updateClientSideIncoreBuffer(s)
        _same(frame.fetch(), rec, f"render with feedback, frame {f}")
        assert st["requests"] == _asked(sub) and st["inserted"] == inserted, f
        assert mm.getSlots() == model.slot_gid, f"slot tables differ after frame {f}"
        mm.updateDeviceMemory(s)
        _say(f"fused feedback frame {f}: inserted {inserted}")


def test_feedback_without_records_changes_nothing_but_visibility(terrain8):
    sc, _ = terrain8
    mm = svo.TreeletMemoryManagerHip(sc, 16 * sc.info["treelet_bytes"], device=0)
    before = mm.getSlots()
    st = mm.updateClientSideIncoreBuffer()
    assert st["requests"] == 0 and st["inserted"] == 0 and mm.getSlots() == before
    # a camera looking away from the cube records nothing
    cam = svo.look_at((0, 0, 3), (0, 0, 6), (0, 1, 0), 30, 64, 64)
    mm.performAnalysePass(cam, 4)
    st = mm.updateClientSideIncoreBuffer()
    assert st["visible_slots"] == 0 and st["requests"] == 0


def test_fetched_readback_states(terrain8):
    """svo_mm_fetch_feedback enqueues the readback only: a second fetch of the
    cycle does nothing, an update waiting on another stream gets the same
    records as a plain update, and a reset drops a pending readback."""
    sc, _ = terrain8
    tb = sc.info["treelet_bytes"]
    cams = _path(3, 320, 200, lod=0.25)
    s1, s2 = svo.Stream(), svo.Stream()
    ref = svo.TreeletMemoryManagerHip(sc, 48 * tb, device=0)
    mm = svo.TreeletMemoryManagerHip(sc, 48 * tb, device=0)
    for f, cam in enumerate(cams):
        _say(f"frame {f}: plain update against fetch, fetch, update on another stream")
        ref.performAnalysePass(cam, 4, (0.0, 0.0), "treelet", s1)
        want = ref.updateClientSideIncoreBuffer(s1)
        ref.updateDeviceMemory(s1)
        mm.performAnalysePass(cam, 4, (0.0, 0.0), "treelet", s1)
        mm.fetchFeedback(s1)
        mm.fetchFeedback(s1)  # the cycle is read already: nothing more is enqueued
        got = mm.updateClientSideIncoreBuffer(s2)
        mm.updateDeviceMemory(s1)
        assert got == want, f"frame {f}"
        assert mm.getSlots() == ref.getSlots(), f"frame {f}"
    # a reset with a readback pending: the next update reads a new cycle
    mm.performAnalysePass(cams[0], 4, (0.0, 0.0), "treelet", s1)
    mm.fetchFeedback(s1)
    mm.resetIncoreBuffer()
    mm.updateDeviceMemory(s1)
    st = mm.updateClientSideIncoreBuffer(s1)
    assert st["requests"] == 0 and st["inserted"] == 0
    assert mm.info["resident_treelets"] == 1


def test_normals_follow_updates_into_shading(terrain8):
    """Uploads leave the normals of refilled slots for the first launch that
    shades with them: phong and normal views after out-of-core updates match
    the oracle over the manager's own buffers, and the device buffer equals
    the host's once read back."""
    sc, _ = terrain8
    tb = sc.info["treelet_bytes"]
    mm = svo.TreeletMemoryManagerHip(sc, 48 * tb, device=0)
    s = svo.Stream()
    cams = _path(6, 200, 150, lod=0.25)
    for f, cam in enumerate(cams):
        mm.performAnalysePass(cam, 4, (0.0, 0.0), "treelet", s)
        mm.updateClientSideIncoreBuffer(s)
        mm.updateDeviceMemory(s)
        mode = "phong" if f % 2 else "normal"
        _say(f"frame {f}: {mode} after {mm.info['resident_treelets']} resident")
        vcam = svo.with_view_mode(cam, mode)
        g = svo.render(mm, vcam, "treelet", shade=True)
        n = mm.info["pool_nodes"]
        nodes, attr = mm.getIncoreBuffer(0, n)
        r = _oracle_shade(mm, vcam, nodes, attr)
        assert np.array_equal(g["node"], r["node"]), f"frame {f}"
        if mode == "normal":
            same = (g["shade"].view(np.uint32) == r["shade"].view(np.uint32)) | (np.isnan(g["shade"]) & np.isnan(r["shade"]))
            assert same.all(), f"frame {f}: {int((~same).sum())} values differ"
        else:
            assert np.allclose(g["shade"], r["shade"], rtol=0, atol=2e-6), f"frame {f}"
    hn, ha = mm.getIncoreBuffer(0, mm.info["pool_nodes"])
    dn, da = mm.getDeviceIncoreBuffer(0, mm.info["pool_nodes"])
    assert np.array_equal(hn, dn) and np.array_equal(ha, da)


def _oracle_shade(mm, cam, nodes, attr):
    import oracle
    return oracle.render_buffers(nodes, attr[:, 1], np.asarray(mm.getSlots(), np.int32), mm.getNumNodesPerTreelet(),
                                 mm.scene.info["max_depth"], cam, "treelet", normals=attr[:, 0])
