"""GPU: multi-view, multi-scene launches, every launch shape, the autotuner
and hipGraph batches -- all bit-identical to the CPU oracle run over the same
incore buffers (including a partially resident scene, where the oracle reads
the manager's own host incore buffer and slot table)."""
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


def _compare(gpu, ref, what):
    for k in ("node", "treelet", "rgba"):
        bad = np.count_nonzero(gpu[k] != ref[k])
        assert bad == 0, f"{what}: {k} differs on {bad} pixels"
    bad = np.count_nonzero(gpu["t"].view(np.uint32) != ref["t"].view(np.uint32))
    assert bad == 0, f"{what}: t bits differ on {bad} pixels"


def _oracle(mm, cam, policy):
    """Oracle render over the manager's host incore buffer and slot table."""
    import oracle
    n = mm.info["pool_nodes"]
    nodes, attr = mm.getIncoreBuffer(0, n)
    slots = np.asarray(mm.getSlots(), np.int32)
    return oracle.render_buffers(nodes, attr[:, 1], slots, mm.getNumNodesPerTreelet(),
                                 mm.scene.info["max_depth"], cam, policy)


def _partially_resident(sc):
    """Root plus, in gid order, every treelet whose parent is resident, until
    half of the slots are used."""
    mm = svo.TreeletMemoryManagerHip(sc, workloads.incore_bytes_for(sc), device=0)
    hdr = sc.headers()
    slots = mm.info["num_slots"]
    resident = {0}
    for g in range(1, hdr.shape[0]):
        if len(resident) >= slots // 2:
            break
        if int(hdr[g, 4]) in resident:
            mm.insertTreeletIntoIncoreBuffer(g, 0.0)
            resident.add(g)
    mm.updateDeviceMemory()
    assert 1 < len(resident) < hdr.shape[0]
    return mm


@pytest.fixture(scope="module")
def scenes(menger8, terrain8, terrain8_bfs):
    m = svo.TreeletMemoryManagerHip(menger8[0], workloads.incore_bytes_for(menger8[0]), 0, resident_all=True)
    t = svo.TreeletMemoryManagerHip(terrain8[0], workloads.incore_bytes_for(terrain8[0], 2), 0, resident_all=True)
    p = _partially_resident(terrain8_bfs[0])
    return {"menger": m, "terrain": t, "partial": p}


def _views(scenes, seed=0):
    spec = [("menger", (64, 64), 0.0), ("terrain", (320, 200), 0.0), ("partial", (130, 77), 0.0),
            ("menger", (7, 9), 1.0), ("terrain", (97, 61), 1.0), ("partial", (256, 144), 1.0)]
    out = []
    for i, (name, (w, h), lod) in enumerate(spec):
        if name == "menger":
            cam = workloads.sphere_camera(700 + seed + i, w, h, lod=lod)
        else:
            cam = workloads.terrain_camera(40 + seed + i, w, h, lod=lod, terrain_seed=workloads.TERRAIN_SEED)
        out.append(svo.View(scenes[name], cam, svo.Frame.alloc(w, h)))
    return out


def _sync():
    svo._capi.check(svo.lib.svo_device_synchronize(), "svo_device_synchronize")


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_multi_scene_launch_matches_oracle(scenes, policy):
    views = _views(scenes)
    r = svo.Renderer(0)
    _say(f"one launch, {len(views)} views over 3 scenes, policy={policy}")
    r.render(views, policy)
    _sync()
    hit_treelets = set()
    for i, v in enumerate(views):
        g = v.frame.fetch()
        _compare(g, _oracle(v.mm, v.camera, policy), f"view {i} {v.camera.width}x{v.camera.height}")
        if v.mm is scenes["partial"]:
            hit_treelets |= set(np.unique(g["treelet"]).tolist()) - {svo.MISS}
    assert len(hit_treelets) > 1


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_every_launch_shape_bit_identical(scenes, policy):
    views = _views(scenes, seed=50)
    refs = [_oracle(v.mm, v.camera, policy) for v in views]
    r = svo.Renderer(0)
    assert r.config == (8, 1, "lds")
    for shape in svo.LAUNCH_SHAPES:
        r.config = shape
        assert r.config == shape
        for v in views:
            v.frame.rgba.fill(0xAB)
            v.frame.node.fill(0xAB)
        _say(f"launch shape {shape} policy={policy}")
        r.render(views, policy)
        _sync()
        for i, (v, ref) in enumerate(zip(views, refs)):
            _compare(v.frame.fetch(), ref, f"shape {shape} view {i}")


def test_bad_launch_config_rejected():
    r = svo.Renderer(0)
    for bad in [(12, 4, "lds"), (16, 3, "lds"), (16, 4, 7)]:
        with pytest.raises(svo.SvoError):
            r.config = bad
    assert r.config == (8, 1, "lds")
    r.config = (32, 1, "private")
    r.config = None
    assert r.config == (8, 1, "lds")


def test_autotune_installs_a_shape_with_identical_results(scenes):
    views = _views(scenes, seed=90)
    r = svo.Renderer(0)
    _say(f"autotune over {len(svo.LAUNCH_SHAPES)} shapes")
    shape, ms = r.autotune(views, "treelet", iters=2)
    _say(f"autotune picked {shape} at {ms:.4f} ms")
    assert shape in svo.LAUNCH_SHAPES and ms > 0
    r.render(views, "treelet")
    _sync()
    for i, v in enumerate(views):
        _compare(v.frame.fetch(), _oracle(v.mm, v.camera, "treelet"), f"tuned view {i}")


def test_autotune_cache_hit_skips_timing_and_persists(scenes, tmp_path):
    views = _views(scenes, seed=91)
    r = svo.Renderer(0)
    key = r.tune_key(views, "treelet")
    _say(f"tune key {key}")
    assert key.startswith("arch=gfx950;") and ";policy=0;" in key and ";views=" in key
    cache = svo.TuneCache()
    shape, ms = r.autotune(views, "treelet", iters=2, cache=cache)
    assert not cache.last_hit and len(cache) == 1 and shape in svo.LAUNCH_SHAPES
    # a hit sets the recorded shape on another renderer without timing
    r2 = svo.Renderer(0)
    r2.config = (32, 4, "private") if shape != (32, 4, "private") else (8, 1, "lds")
    shape2, ms2 = r2.autotune(views, "treelet", iters=2, cache=cache)
    assert cache.last_hit and shape2 == shape and ms2 == pytest.approx(ms)
    # other views of another class miss
    small = [svo.View(scenes["menger"], workloads.sphere_camera(3, 40, 24), svo.Frame.alloc(40, 24))]
    assert r2.tune_key(small, "treelet") != key
    r2.autotune(small, "treelet", iters=1, cache=cache)
    assert not cache.last_hit and len(cache) == 2
    # through a file, into a new process-wide cache
    path = str(tmp_path / "tune.tsv")
    cache.save(path)
    fresh = svo.TuneCache(path)
    r3 = svo.Renderer(0)
    shape3, _ = r3.autotune(views, "treelet", iters=2, cache=fresh)
    assert fresh.last_hit and shape3 == shape
    r3.render(views, "treelet")
    _sync()
    for i, v in enumerate(views):
        _compare(v.frame.fetch(), _oracle(v.mm, v.camera, "treelet"), f"cached-shape view {i}")


def test_batch_graph_replays_and_camera_updates(scenes):
    views = _views(scenes, seed=120)
    r = svo.Renderer(0)
    s = svo.Stream()
    b = svo.Batch(r, views, "treelet", s)
    refs = [_oracle(v.mm, v.camera, "treelet") for v in views]
    side = svo.View(scenes["menger"], workloads.sphere_camera(5, 48, 40), svo.Frame.alloc(48, 40))
    for rep in range(4):
        for v in views:
            v.frame.rgba.fill(0xCD)
            v.frame.t.fill(0xCD)
        # eager work on the same stream between replays
        r.render([side], "cpu3", s)
        _say(f"batch replay {rep}")
        b.launch(s)
        s.synchronize()
        for i, (v, ref) in enumerate(zip(views, refs)):
            _compare(v.frame.fetch(), ref, f"replay {rep} view {i}")
    _compare(side.frame.fetch(), _oracle(side.mm, side.camera, "cpu3"), "eager side launch")

    # move every camera, keep the frame sizes
    cams = []
    for i, v in enumerate(views):
        w, h, lod = v.camera.width, v.camera.height, 0.5
        if v.mm is scenes["menger"]:
            cams.append(workloads.sphere_camera(900 + i, w, h, lod=lod))
        else:
            cams.append(workloads.terrain_camera(960 + i, w, h, lod=lod, terrain_seed=workloads.TERRAIN_SEED))
    for rep in range(2):
        b.set_cameras(cams, s)
        _say(f"batch after camera update, replay {rep}")
        b.launch(s)
        s.synchronize()
        for i, (v, c) in enumerate(zip(views, cams)):
            _compare(v.frame.fetch(), _oracle(v.mm, c, "treelet"), f"moved camera view {i}")

    wrong = [workloads.sphere_camera(1, 10, 10)] * len(views)
    with pytest.raises(svo.SvoError):
        b.set_cameras(wrong, s)
    with pytest.raises(svo.SvoError):
        b.set_cameras(cams[:-1], s)


def test_view_errors(scenes, menger8):
    r = svo.Renderer(0)
    f = svo.Frame.alloc(8, 8)
    cam = workloads.sphere_camera(1, 8, 8)
    host_only = svo.TreeletMemoryManagerHip(menger8[0], workloads.incore_bytes_for(menger8[0]), device=-1)
    with pytest.raises(svo.SvoError):
        r.render([svo.View(host_only, cam, f)])
    with pytest.raises(ValueError):
        r.render([])
    zero = workloads.sphere_camera(1, 8, 8)
    zero.width = 0
    with pytest.raises(svo.SvoError):
        r.render([svo.View(scenes["menger"], zero, f)])
    with pytest.raises(svo.SvoError):
        r.render([svo.View(scenes["menger"], cam, f)], policy=7)


def _terrain_path(n, w, h, lod, seed=500):
    return [workloads.terrain_camera(seed + i, w, h, lod=lod, terrain_seed=workloads.TERRAIN_SEED) for i in range(n)]


def test_batch_feedback_feeds_every_cycle(terrain8):
    """A graph batch whose view records analysis feedback keeps feeding the
    out-of-core update cycle after cycle (the cycle stamp is read on the
    device at replay time): page-ins, slot tables and pixels equal those of
    the same views rendered with svo_render_views on a second manager."""
    sc, _ = terrain8
    tb = sc.info["treelet_bytes"]
    a = svo.TreeletMemoryManagerHip(sc, 48 * tb, device=0)
    b = svo.TreeletMemoryManagerHip(sc, 48 * tb, device=0)
    cams = _terrain_path(5, 256, 160, lod=0.25)
    r = svo.Renderer(0)
    s = svo.Stream()
    fa, fb = svo.Frame.alloc(256, 160), svo.Frame.alloc(256, 160)
    batch = svo.Batch(r, [svo.View(a, cams[0], fa, feedback_stride=3)], "treelet", s)
    pages = 0
    for c, cam in enumerate(cams):
        batch.set_cameras([cam], s)
        batch.launch(s)
        r.render([svo.View(b, cam, fb, feedback_stride=3)], "treelet", s)
        s.synchronize()
        _compare(fa.fetch(), fb.fetch(), f"cycle {c}")
        sa = a.updateClientSideIncoreBuffer(s)
        sb = b.updateClientSideIncoreBuffer(s)
        _say(f"batch feedback cycle {c}: {sa}")
        assert sa == sb, f"cycle {c}"
        assert sa["visible_slots"] > 0, f"cycle {c}: the batch recorded nothing"
        assert a.getSlots() == b.getSlots(), f"cycle {c}"
        pages += sa["inserted"]
        a.updateDeviceMemory(s)
        b.updateDeviceMemory(s)
    assert pages > 0
    # a replay with slots waiting for upload is refused
    a.removeTreeletFromIncoreBuffer(next(g for g in a.getSlots() if g > 0))
    with pytest.raises(svo.SvoError):
        batch.launch(s)
    a.updateDeviceMemory(s)
    batch.launch(s)
    s.synchronize()


def test_batch_camera_update_switches_view_mode(terrain8):
    """svo_batch_set_cameras checks view modes and, when a view switches to a
    mode that shades with normals, sends the normals that out-of-core
    updates left stale before the next replay."""
    import oracle
    sc, _ = terrain8
    tb = sc.info["treelet_bytes"]
    mm = svo.TreeletMemoryManagerHip(sc, 48 * tb, device=0)
    s = svo.Stream()
    cams = _terrain_path(4, 200, 150, lod=0.25, seed=620)
    for cam in cams[:3]:  # refill slots: their normals are stale on the device
        mm.performAnalysePass(cam, 4, (0.0, 0.0), "treelet", s)
        mm.updateClientSideIncoreBuffer(s)
        mm.updateDeviceMemory(s)
    cam = cams[3]
    f = svo.Frame.alloc(cam.width, cam.height, shade=True)
    r = svo.Renderer(0)
    batch = svo.Batch(r, [svo.View(mm, svo.with_view_mode(cam, "color"), f)], "treelet", s)
    bad = svo.with_view_mode(cam, 99)
    with pytest.raises(svo.SvoError):
        batch.set_cameras([bad], s)
    for mode in ("normal", "phong"):
        vcam = svo.with_view_mode(cam, mode)
        batch.set_cameras([vcam], s)
        batch.launch(s)
        s.synchronize()
        g = f.fetch()
        nodes, attr = mm.getIncoreBuffer()
        ref = oracle.render_buffers(nodes, attr[:, 1], np.asarray(mm.getSlots(), np.int32), mm.getNumNodesPerTreelet(),
                                    mm.scene.info["max_depth"], vcam, "treelet", normals=attr[:, 0])
        _compare(g, ref, f"{mode} after camera update")
        if mode == "normal":
            same = (g["shade"].view(np.uint32) == ref["shade"].view(np.uint32)) | (np.isnan(g["shade"]) &
                                                                                  np.isnan(ref["shade"]))
            assert same.all(), f"{int((~same).sum())} shaded values differ"
        else:
            assert np.allclose(g["shade"], ref["shade"], rtol=0, atol=2e-6)
