"""GPU: the continuous-batching render server.  Every request's pixels must
be bit-identical alone and in the mix (and to the CPU oracle); HBM-budget
admission of scenes and frame-pool admission of requests are exercised."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (128, 96), (200, 130), (37, 23), (256, 144)]


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


@pytest.fixture(scope="module")
def mms(menger8, terrain8, terrain8_bfs):
    out = []
    for sc, _ in (menger8, terrain8, terrain8_bfs):
        out.append(svo.TreeletMemoryManagerHip(sc, workloads.incore_bytes_for(sc), 0, resident_all=True))
    return out


KINDS = ["menger", "terrain", "terrain"]


def _requests(n, rate, seed):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    out = []
    for i in range(n):
        k = int(rng.integers(len(KINDS)))
        w, h = SIZES[int(rng.integers(len(SIZES)))]
        out.append((k, workloads.scene_camera(KINDS[k], int(rng.integers(1 << 30)), w, h,
                                              lod=float(rng.choice([0.0, 1.0]))), float(t[i])))
    return out


def _alone(renderer, mm, cam, policy="treelet"):
    f = svo.Frame.alloc(cam.width, cam.height)
    renderer.render([svo.View(mm, cam, f)], policy)
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    return f.fetch()


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_each_request_same_alone_and_in_mix(mms, policy):
    r = svo.Renderer(0)
    srv = svo.Server(r, max_batch_views=16, max_inflight=3, policy=policy)
    ids = [srv.add_scene(mm) for mm in mms]
    assert ids == [0, 1, 2]
    reqs = _requests(80, 40000.0, seed=3)
    frames = [svo.Frame.alloc(c.width, c.height) for _, c, _ in reqs]
    for f in frames:
        f.rgba.fill(0xEE)
    _say(f"replay of {len(reqs)} requests, policy={policy}")
    stats, comp = srv.replay([(k, c, t, f) for (k, c, t), f in zip(reqs, frames)])
    _say(f"stats {stats}")
    assert stats["requests"] == len(reqs)
    assert stats["max_batch_views_seen"] > 1, "requests were never batched together"
    assert np.all(comp["done_s"] >= comp["launch_s"]) and np.all(comp["launch_s"] >= comp["arrival_s"])
    assert np.all(comp["batch_views"] >= 1)
    import oracle
    for i, ((k, cam, _), f) in enumerate(zip(reqs, frames)):
        mixed = f.fetch()
        _same(mixed, _alone(r, mms[k], cam, policy), f"request {i} (batch of {comp['batch_views'][i]})")
        if i % 8 == 0:
            ref = oracle.render_buffers(*_pools(mms[k]), cam, policy)
            _same(mixed, ref, f"request {i} vs oracle")


def _pools(mm):
    n = mm.info["pool_nodes"]
    nodes, attr = mm.getIncoreBuffer(0, n)
    return nodes, attr[:, 1], np.asarray(mm.getSlots(), np.int32), mm.getNumNodesPerTreelet(), \
        mm.scene.info["max_depth"]


def test_online_submit_and_poll(mms):
    r = svo.Renderer(0)
    srv = svo.Server(r, max_batch_views=8, max_inflight=2)
    for mm in mms:
        srv.add_scene(mm)
    reqs = _requests(40, 1.0, seed=9)
    frames = {}
    done = []
    for i, (k, cam, _) in enumerate(reqs):
        f = svo.Frame.alloc(cam.width, cam.height)
        rid = srv.submit(k, cam, f)
        frames[rid] = (k, cam, f)
        if i % 3 == 0:
            done += srv.poll(4)  # small cap: completions beyond it are held
    while srv.pending or len(done) < len(reqs):
        done += srv.poll(4)
    assert sorted(c["id"] for c in done) == sorted(frames)
    assert max(c["batch_views"] for c in done) > 1
    for rid, (k, cam, f) in frames.items():
        _same(f.fetch(), _alone(r, mms[k], cam), f"online request {rid}")


def test_frame_pool_admission(mms):
    r = svo.Renderer(0)
    pool = 2 * 200 * 130 * 16 + 512  # two of the largest frames used here
    srv = svo.Server(r, frame_pool=pool, max_batch_views=8, max_inflight=4)
    for mm in mms:
        srv.add_scene(mm)
    reqs = _requests(60, 50000.0, seed=11)
    reqs = [(k, c, t) for k, c, t in reqs if c.width * c.height <= 200 * 130]
    stats, comp = srv.replay(reqs)
    _say(f"pool-limited stats {stats}")
    assert stats["requests"] == len(reqs)
    assert 0 < stats["peak_pool_bytes"] <= pool
    big = workloads.sphere_camera(1, 512, 512)
    with pytest.raises(svo.SvoError):
        srv.replay([(0, big, 0.0)])


def test_hbm_budget_admission(mms):
    r = svo.Renderer(0)
    bytes0 = mms[0].info["device_bytes"]
    pool = 1 << 20
    srv = svo.Server(r, hbm_budget=bytes0 + pool, frame_pool=pool)
    assert srv.add_scene(mms[0]) == 0
    with pytest.raises(svo.SvoError):
        srv.add_scene(mms[1])
    with pytest.raises(svo.SvoError):
        svo.Server(r, hbm_budget=pool, frame_pool=2 * pool)


def test_plain_loop_is_one_view_per_launch(mms):
    r = svo.Renderer(0)
    srv = svo.Server(r, max_batch_views=1, max_inflight=1)
    for mm in mms:
        srv.add_scene(mm)
    reqs = _requests(30, 50000.0, seed=13)
    stats, comp = srv.replay(reqs)
    assert stats["batches"] == len(reqs)
    assert stats["max_batch_views_seen"] == 1 and stats["max_inflight_seen"] == 1
    assert len(set(comp["batch"].tolist())) == len(reqs)


def test_queue_bound_refuses_and_recovers(mms):
    r = svo.Renderer(0)
    srv = svo.Server(r, max_batch_views=1, max_inflight=1, max_waiting=3)
    for mm in mms:
        srv.add_scene(mm)
    cams = [workloads.scene_camera("menger", 70 + i, 64, 64) for i in range(5)]
    ids = [srv.submit(0, c) for c in cams[:3]]
    with pytest.raises(svo.SvoError) as e:
        srv.submit(0, cams[3])  # nothing has been launched yet: the queue holds 3
    assert e.value.status == svo._capi.SVO_ERR_CAPACITY
    done = []
    while srv.pending:
        done += srv.poll()
    assert sorted(c["id"] for c in done) == sorted(ids)
    assert all(c["status"] == svo._capi.SVO_REQ_DONE and c["batch_views"] == 1 for c in done)
    srv.submit(0, cams[4])  # room again
    while srv.pending:
        srv.poll()


def test_timeouts_drop_waiting_requests_unrendered(mms):
    r = svo.Renderer(0)
    srv = svo.Server(r, max_batch_views=1, max_inflight=1)
    for mm in mms:
        srv.add_scene(mm)
    cam = workloads.scene_camera("terrain", 5, 256, 144)
    keep = svo.Frame.alloc(256, 144)
    first = srv.submit(1, cam, keep)  # no timeout
    dropped = []
    for i in range(4):
        f = svo.Frame.alloc(64, 64)
        f.rgba.fill(0xAB)
        dropped.append((srv.submit(0, workloads.scene_camera("menger", 80 + i, 64, 64), f, timeout_us=1), f))
    done = []
    while srv.pending:
        done += srv.poll()
    status = {c["id"]: c for c in done}
    assert status[first]["status"] == svo._capi.SVO_REQ_DONE
    _same(keep.fetch(), _alone(r, mms[1], cam), "request that did not time out")
    for rid, f in dropped:
        c = status[rid]
        assert c["status"] == svo._capi.SVO_REQ_EXPIRED and c["batch_views"] == 0
        assert c["launch_s"] == c["done_s"] >= c["arrival_s"]
        assert np.all(f.fetch()["rgba"] == 0xABABABAB), "an expired request must not be rendered"


def test_replay_under_overload_reports_dropped_requests(mms):
    r = svo.Renderer(0)
    srv = svo.Server(r, max_batch_views=2, max_inflight=1, max_waiting=4)
    for mm in mms:
        srv.add_scene(mm)
    reqs = _requests(80, 200000.0, seed=21)  # all but simultaneous
    reqs = [(k, c, t, None, 300 if i % 2 else 0) for i, (k, c, t) in enumerate(reqs)]
    stats, comp = srv.replay(reqs)
    _say(f"overload stats {stats}")
    st = comp["status"]
    assert stats["requests"] == len(reqs)
    assert stats["rejected"] == int((st == svo._capi.SVO_REQ_REJECTED).sum()) > 0
    assert stats["expired"] == int((st == svo._capi.SVO_REQ_EXPIRED).sum())
    rendered = st == svo._capi.SVO_REQ_DONE
    assert rendered.sum() + stats["rejected"] + stats["expired"] == len(reqs)
    # only requests with a timeout can expire
    assert all(reqs[i][4] > 0 for i in np.flatnonzero(st == svo._capi.SVO_REQ_EXPIRED))
    assert np.all(comp["batch_views"][~rendered] == 0) and np.all(comp["batch_views"][rendered] >= 1)


def test_teardown_in_any_order(mms):
    """A garbage-collection cycle may finalize a Renderer before the Server
    or Batch made from it: their teardown must not read the renderer (it
    used to, leaving an invalid-device HIP error behind)."""
    import ctypes
    import gc
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipPeekAtLastError.restype = ctypes.c_int
    hip.hipGetLastError.restype = ctypes.c_int
    hip.hipGetLastError()
    for _ in range(3):
        r = svo.Renderer(0)
        srv = svo.Server(r, max_batch_views=2, max_inflight=1)
        srv.add_scene(mms[0])
        cam = workloads.scene_camera("menger", 3, 64, 64)
        b = svo.Batch(r, [svo.View(mms[0], cam, svo.Frame.alloc(64, 64))], "treelet")
        cycle = [r, srv, b]
        cycle.append(cycle)  # the objects now live in a reference cycle
        del r, srv, b, cycle
        gc.collect()
        assert hip.hipPeekAtLastError() == 0
