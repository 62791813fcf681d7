"""GPU: svo_render_views_noalloc, the allocation-free launch (descriptors in
a fixed device scratch array), called through the C-ABI, against
svo_render_views on the same views: two calls in a row, the second with
fewer and smaller views, every output bit-identical; more views than the
scratch holds are refused."""
import ctypes

import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import _capi, workloads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def scenes(menger8, terrain8):
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    m = svo.TreeletMemoryManagerHip(menger8[0], workloads.incore_bytes_for(menger8[0]), 0, resident_all=True)
    t = svo.TreeletMemoryManagerHip(terrain8[0], workloads.incore_bytes_for(terrain8[0]), 0, resident_all=True)
    return m, t


def _views(scenes, spec, seed):
    m, t = scenes
    out = []
    for i, (name, w, h, mode) in enumerate(spec):
        if name == "menger":
            cam = workloads.sphere_camera(seed + i, w, h, lod=0.5)
        else:
            cam = workloads.terrain_camera(seed + i, w, h, lod=0.5, terrain_seed=workloads.TERRAIN_SEED)
        cam = svo.with_view_mode(cam, mode)
        out.append(svo.View(m if name == "menger" else t, cam,
                            svo.Frame.alloc(w, h, pushes=True, shade=mode != "color")))
    return out


def _noalloc(views, policy, stream, config=None):
    arr = (_capi.svo_view * len(views))(*[v.c_view() for v in views])
    cfg = None if config is None else ctypes.byref(_capi.svo_launch_config(*config, 0))
    return svo.lib.svo_render_views_noalloc(arr, len(views), svo.POLICIES[policy], cfg, stream.handle)


def _fetch_all(views):
    return [v.frame.fetch() for v in views]


def _same(a, b, what):
    for k in a:
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        bad = np.count_nonzero(x != y)
        assert bad == 0, f"{what}: {k} differs on {bad} values"


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_noalloc_matches_render_views_twice(scenes, policy):
    s = svo.Stream()
    r = svo.Renderer(0)
    calls = [
        [("menger", 200, 150, "color"), ("terrain", 320, 180, "normal"), ("terrain", 77, 53, "phong"),
         ("menger", 128, 128, "iterations")],
        [("terrain", 96, 64, "color")],  # the second call: fewer, smaller views
    ]
    for c, spec in enumerate(calls):
        views = _views(scenes, spec, seed=300 + 10 * c)
        for v in views:
            v.frame.node.fill(0xAB)
            v.frame.t.fill(0xAB)
        _say(f"call {c}: {len(views)} views through svo_render_views_noalloc, policy={policy}")
        _capi.check(_noalloc(views, policy, s), "svo_render_views_noalloc")
        s.synchronize()
        got = _fetch_all(views)
        r.render(views, policy, s)
        s.synchronize()
        want = _fetch_all(views)
        for i, (g, w) in enumerate(zip(got, want)):
            _same(g, w, f"call {c} view {i}")
            assert np.count_nonzero(g["node"] != svo.MISS) > 0


def test_noalloc_shape_and_limits(scenes):
    s = svo.Stream()
    views = _views(scenes, [("menger", 64, 48, "color")], seed=77)
    _capi.check(_noalloc(views, "treelet", s, config=(16, 2, _capi.SVO_STACK_PRIVATE)), "noalloc with a shape")
    s.synchronize()
    got = views[0].frame.fetch()
    svo.Renderer(0).render(views, "treelet", s)
    s.synchronize()
    _same(got, views[0].frame.fetch(), "explicit launch shape")
    too_many = views * (_capi.SVO_NOALLOC_MAX_VIEWS + 1)
    assert _noalloc(too_many, "treelet", s) == _capi.SVO_ERR_ARG
    assert b"at most" in svo.lib.svo_last_error()
    assert _noalloc(views, "treelet", s, config=(12, 1, 0)) == _capi.SVO_ERR_ARG
