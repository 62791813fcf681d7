"""GPU parity at the benchmarked sizes of the batched and serving workloads.

* Batched (64 views over the depth-11 terrain in 4096-node treelets and the
  depth-10 sphere cloud, resolutions uniform over {256^2, 512^2, 1280x720,
  1920x1080}, seed 31): the launch shape comes from the autotuner, the launch
  is captured in a hipGraph and replayed; every view of every replay is
  compared with the oracle over the managers' own host incore buffers.
* Serving (Menger depth 9 in 32 KiB treelets, terrain depth 11, sphere cloud
  depth 10, terrain depth 12, all resident in one server; Poisson arrivals,
  resolutions 256^2 40 %, 720p 40 %, 1080p 15 %, 4K 5 %): every request is
  compared with the same request rendered alone, and one in eight with the
  oracle, for a stream at 1000 req/s and for the same requests all queued at
  once (where launches hold many views).

Node, treelet id and RGBA must be equal and t bit-identical (policy A, the
treelet renderer's sample(), opencl/treeletRenderer_rgba.cl:201-382).
"""
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


def _oracle(mm, cam, policy="treelet"):
    """Oracle render over the manager's own host incore buffer (zero-copy
    views) and slot table."""
    import oracle
    nodes = mm.getIncoreBufferView()
    _, colors = mm.getAttributeIncoreBufferView()
    return oracle.render_buffers(nodes, colors, np.asarray(mm.getSlots(), np.int32), mm.getNumNodesPerTreelet(),
                                 mm.scene.info["max_depth"], cam, policy)


_BUILD = {"terrain11": lambda: workloads.terrain_scene(depth=11),
          "terrain12": lambda: workloads.terrain_scene(depth=12),
          "spheres10": lambda: workloads.sphere_cloud_scene(depth=10),
          "menger9": lambda: workloads.menger_scene(depth=9, treelet_bytes=32768)}


@pytest.fixture(scope="module")
def resident():
    """The benchmark scenes, built once and made resident on demand."""
    cache = {}

    def get(key):
        if key not in cache:
            sc = _BUILD[key]()
            cache[key] = svo.TreeletMemoryManagerHip(sc, workloads.incore_bytes_for(sc), 0, resident_all=True)
            _say(f"{key}: {sc.info['num_treelets']} treelets resident")
        return cache[key]
    yield get
    cache.clear()


def test_batched_64_views_graph_replays_match_oracle(resident):
    mms = [resident("terrain11"), resident("spheres10")]
    spec = workloads.batched_views(["terrain", "spheres"])
    views = [svo.View(mms[k], cam, svo.Frame.alloc(cam.width, cam.height)) for k, cam in spec]
    rays = sum(v.camera.width * v.camera.height for v in views)
    assert len(views) == 64 and rays > 40_000_000
    r = svo.Renderer(0)
    s = svo.Stream()
    shape, ms = r.autotune(views, "treelet", s, iters=1)
    _say(f"batched: {len(views)} views, {rays} rays, autotuned shape {shape} ({ms:.3f} ms)")
    b = svo.Batch(r, views, "treelet", s)
    refs = [_oracle(v.mm, v.camera) for v in views]
    side = svo.View(mms[1], workloads.sphere_camera(3, 96, 80), svo.Frame.alloc(96, 80))
    for rep in range(2):
        for v in views:
            v.frame.node.fill(0xAB)
            v.frame.t.fill(0xAB)
        r.render([side], "treelet", s)  # eager work between replays
        _say(f"batched: graph replay {rep}")
        b.launch(s)
        s.synchronize()
        hits = 0
        for i, (v, ref) in enumerate(zip(views, refs)):
            g = v.frame.fetch()
            _same(g, ref, f"replay {rep} view {i} ({v.camera.width}x{v.camera.height}, "
                          f"{'terrain' if v.mm is mms[0] else 'spheres'})")
            hits += int(np.count_nonzero(g["node"] != svo.MISS))
        assert hits > rays // 4
    _same(side.frame.fetch(), _oracle(side.mm, side.camera), "eager view between replays")


def _serving_mms(resident):
    keys = ["menger9", "terrain11", "spheres10", "terrain12"]
    return [resident(k) for k in keys], ["menger", "terrain", "spheres", "terrain"]


@pytest.mark.parametrize("rate", [1000.0, 1e9])
def test_serving_mix_each_request_same_alone_and_in_mix(resident, rate):
    mms, kinds = _serving_mms(resident)
    budget = 64 << 30
    assert sum(mm.info["device_bytes"] for mm in mms) + (4 << 30) < budget
    r = svo.Renderer(0)
    srv = svo.Server(r, hbm_budget=budget, frame_pool=4 << 30, max_batch_views=64, max_inflight=4,
                     max_batch_rays=16 << 20, policy="treelet")
    assert [srv.add_scene(mm) for mm in mms] == [0, 1, 2, 3]
    reqs = workloads.serving_requests(kinds, rate, 48)
    frames = [svo.Frame.alloc(c.width, c.height) for _, c, _ in reqs]
    for f in frames:
        f.node.fill(0xEE)
    _say(f"serving: {len(reqs)} requests at {rate:g} req/s, sizes "
         f"{sorted({(c.width, c.height) for _, c, _ in reqs})}")
    stats, comp = srv.replay([(k, c, t, f) for (k, c, t), f in zip(reqs, frames)])
    _say(f"serving: {stats['frames_per_s']:.1f} frames/s, p50 {stats['p50_ms']:.3f} ms, p99 {stats['p99_ms']:.3f} ms, "
         f"max views per launch {stats['max_batch_views_seen']}")
    assert stats["requests"] == len(reqs) and stats["expired"] == 0 and stats["rejected"] == 0
    if rate > 1e6:
        assert stats["max_batch_views_seen"] > 1, "requests queued together were never batched"
    for i, ((k, cam, _), f) in enumerate(zip(reqs, frames)):
        mixed = f.fetch()
        alone = svo.Frame.alloc(cam.width, cam.height)
        r.render([svo.View(mms[k], cam, alone)], "treelet")
        svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
        _same(mixed, alone.fetch(), f"request {i} ({kinds[k]}, {cam.width}x{cam.height}, "
                                    f"batch of {comp['batch_views'][i]})")
        if i % 8 == 0:
            _same(mixed, _oracle(mms[k], cam), f"request {i} vs oracle")
