"""GPU parity of the out-of-core workload at its benchmarked size: the
depth-13 terrain (8192^3 leaf grid) in 16384-node treelets, larger than the
device pool, along the seeded 300-frame Catmull-Rom fly-through at 1920x1080
with LOD 1 pixel and the analyse pass at 1/8 per axis.

Every frame runs analyse pass -> feedback update -> upload -> render on the
GPU.  The same frame is replayed on the CPU: the oracle traverses the
analysis frame over the manager's own host incore buffer, the restated
reference bookkeeping (oracle/residency.py, src/RenderPassAnalyse.cpp:248-356
and src/TreeletMemoryManager.cpp:392-700) aggregates its per-pixel records
and updates its slot table.  After each frame the slot tables, visibilities
and page-in / eviction counts must be equal, and the rendered frame must be
bit-identical to the oracle's render over the manager's buffer.

* the 8 GiB pool (32,768 slots): the first 12 frames of the flight;
* a 640-slot pool with every 12th camera of the flight (24 frames): the pool
  fills by frame 12 and freeIncorePosition evicts from then on.
"""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads

pytestmark = pytest.mark.gpu

DIVIDE = 8


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


@pytest.fixture(scope="module")
def terrain13():
    sc = workloads.terrain_scene(depth=13, treelet_bytes=131072)
    _say(f"terrain depth 13: {sc.info}")
    yield sc


def _same(a, b, what):
    for k in ("node", "treelet", "rgba"):
        bad = np.count_nonzero(a[k] != b[k])
        assert bad == 0, f"{what}: {k} differs on {bad} pixels"
    bad = np.count_nonzero(a["t"].view(np.uint32) != b["t"].view(np.uint32))
    assert bad == 0, f"{what}: t bits differ on {bad} pixels"


def _oracle(mm, cam, records=False):
    import oracle
    nodes = mm.getIncoreBufferView()
    _, colors = mm.getAttributeIncoreBufferView()
    return oracle.render_buffers(nodes, colors, np.asarray(mm.getSlots(), np.int32), mm.getNumNodesPerTreelet(),
                                 mm.scene.info["max_depth"], cam, "treelet", records=records)


def _flight(sc, num_slots, cams, label):
    from oracle import residency as R
    info = sc.info
    N, tb = info["nodes_per_treelet"], info["treelet_bytes"]
    mm = svo.TreeletMemoryManagerHip(sc, num_slots * tb, device=0)
    model = R.Residency(sc.headers(), num_slots)
    rng = np.random.default_rng(workloads.FLIGHT_SEED)
    r = svo.Renderer(0)
    s = svo.Stream()
    frame = svo.Frame.alloc(cams[0].width, cams[0].height)
    view = [svo.View(mm, cams[0], frame)]
    pages = evicted = failed = 0
    for f, cam in enumerate(cams):
        jit = tuple(float(v) for v in rng.uniform(-1.0, 1.0, 2))
        rec = _oracle(mm, svo.analysis_camera(cam, DIVIDE, jit), records=True)
        old, new = R.aggregate_records(rec, model.slot_gid, model.parent, N, tb)
        ev0 = model.evictions
        inserted = model.update_client_side(old, new)
        mm.performAnalysePass(cam, DIVIDE, jit, "treelet", s)
        st = mm.updateClientSideIncoreBuffer(s)
        assert st["visible_treelets"] == len(old), f"{label} frame {f}"
        assert st["inserted"] == inserted and st["evicted"] == model.evictions - ev0, f"{label} frame {f}: {st}"
        infos = mm.getSlotInfos()
        assert [i[0] for i in infos] == model.slot_gid, f"{label}: slot tables differ after frame {f}"
        assert [i[1] for i in infos] == model.vis, f"{label}: visibilities differ after frame {f}"
        mm.updateDeviceMemory(s)
        view[0].camera = cam
        frame.node.fill(0xAB)
        r.render(view, "treelet", s)
        s.synchronize()
        _same(frame.fetch(), _oracle(mm, cam), f"{label} frame {f}")
        pages += inserted
        evicted += st["evicted"]
        failed += st["insert_failed"]
        _say(f"{label} frame {f}: {inserted} page-ins, {st['evicted']} evictions, {st['insert_failed']} refused, "
             f"{mm.info['resident_treelets']} resident")
    return mm, pages, evicted, failed


def test_flight_8gib_pool_first_frames(terrain13):
    sc = terrain13
    slots = (8 << 30) // (sc.info["nodes_per_treelet"] * 16)  # nodes + colour + normal per node
    assert slots == 32768 and sc.info["num_treelets"] > slots
    cams = workloads.flythrough_cameras(300, lod=1.0)[:12]
    mm, pages, evicted, _ = _flight(sc, slots, cams, "8 GiB pool")
    assert pages > 100 and evicted == 0
    # after the updates the HBM pools equal the host incore buffer (slots are
    # taken from 0 upwards: beyond the highest used slot both are zero)
    used = max(s for s, g in enumerate(mm.getSlots()) if g >= 0) + 1
    n = used * sc.info["nodes_per_treelet"]
    hn = mm.getIncoreBufferView()[:n]
    dn, da = mm.getDeviceIncoreBuffer(0, n)
    assert np.array_equal(hn, dn) and np.array_equal(mm.getAttributeIncoreBuffer(0, n), da)


def test_flight_small_pool_evicts(terrain13):
    cams = workloads.flythrough_cameras(300, lod=1.0)[::12][:24]
    _, pages, evicted, failed = _flight(terrain13, 640, cams, "640-slot pool")
    _say(f"640-slot pool: {pages} page-ins, {evicted} evictions, {failed} refused")
    assert pages > 640 and evicted > 0
