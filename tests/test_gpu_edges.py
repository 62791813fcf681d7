"""GPU edge cases against the oracle, bit for bit: the deepest scenes each
policy's 16-level stack allows (a wavy surface patch at depth 15 and 16),
the depth limit refused cleanly, frames of one pixel, one row and one
column through every launch shape, the largest frame sides, eyes exactly
on the octree's planes (signed zeros included), and a scene of a single
leaf."""
import numpy as np
import pytest

import libsvo_amd as svo
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _same(gpu, ref, what):
    for k in ("node", "treelet", "rgba"):
        bad = np.count_nonzero(gpu[k] != ref[k])
        assert bad == 0, f"{what}: {k} differs on {bad} pixels"
    bad = np.count_nonzero(gpu["t"].view(np.uint32) != ref["t"].view(np.uint32))
    assert bad == 0, f"{what}: t bits differ on {bad} pixels"


def _patch_scene(depth, tb):
    """A 256 x 256 wavy surface patch of leaves around the centre of the cube
    at `depth` (deep, narrow subtrees)."""
    side = 1 << depth
    c = side // 2
    i, k = np.meshgrid(np.arange(c - 128, c + 128), np.arange(c - 128, c + 128), indexing="ij")
    j = c + np.floor(6.0 * np.sin(i / 19.0) + 6.0 * np.cos(k / 13.0)).astype(np.int64)
    codes = np.unique(H.morton_np(i.ravel(), j.ravel(), k.ravel(), depth))
    rng = np.random.default_rng(depth)
    attr = np.stack([np.full(codes.size, 0x80FF8000, np.uint32),
                     (rng.integers(0, 2**24, size=codes.size, dtype=np.uint32) << np.uint32(8)) | np.uint32(0xFF)], 1)
    return svo.Scene.from_leaves(depth, codes, attr, tb, "dense")


def _patch_camera(depth, w=96, h=80, lod=0.0):
    s = 1.0 / (1 << depth)
    return svo.look_at((90 * s, 60 * s, 120 * s), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, w, h, far=1.0, lod=lod)


def _oracle(sc, tmp_path, name, cam, policy):
    import oracle
    base = str(tmp_path / name)
    sc.write(base)
    ora = oracle.OracleScene(base, sc.info["treelet_bytes"] * sc.info["num_treelets"])
    return ora.render(cam, policy)


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
@pytest.mark.parametrize("lod", [0.0, 1.0])
def test_depth15_patch(tmp_path, policy, lod):
    sc = _patch_scene(15, 4096)
    _say(f"depth-15 patch: {sc.info['num_treelets']} treelets, {sc.info['tree_nodes']} nodes; policy={policy} lod={lod}")
    mm = svo.TreeletMemoryManagerHip(sc, sc.info["treelet_bytes"] * sc.info["num_treelets"], 0, resident_all=True)
    cam = _patch_camera(15, lod=lod)
    gpu = svo.render(mm, cam, policy)
    ref = _oracle(sc, tmp_path, "d15", cam, policy)
    _same(gpu, ref, f"depth 15 {policy} lod {lod}")
    hits = gpu["node"] != svo.MISS
    assert hits.sum() > 2000 and (~hits).sum() > 100
    if lod == 0.0 and policy == "treelet":
        assert (ref["depth"][hits] == 15).all()


def test_depth15_single_svo_policy(tmp_path):
    import oracle
    sc = _patch_scene(15, 0)
    base = str(tmp_path / "d15_single")
    sc.write_single(base, "v1")
    single = svo.Scene.open_single(base)
    mm = svo.TreeletMemoryManagerHip(single, single.info["treelet_bytes"], 0, resident_all=True)
    cam = _patch_camera(15)
    gpu = svo.render(mm, cam, "single")
    ref = oracle.OracleSingle(base).render(cam)
    for k in ("node", "rgba"):
        assert np.array_equal(gpu[k], ref[k]), k
    assert np.array_equal(gpu["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert (gpu["node"] != svo.MISS).sum() > 2000


def test_depth16_limits(tmp_path):
    sc = _patch_scene(16, 4096)
    mm = svo.TreeletMemoryManagerHip(sc, sc.info["treelet_bytes"] * sc.info["num_treelets"], 0, resident_all=True)
    cam = _patch_camera(16)
    # the treelet policy needs scaleMax = depth + 1 = 17 > 16 stack levels
    with pytest.raises(svo.SvoError):
        svo.render(mm, cam, "treelet")
    # the CPU caster's policy (scaleMax = depth) still fits
    gpu = svo.render(mm, cam, "cpu3")
    ref = _oracle(sc, tmp_path, "d16", cam, "cpu3")
    _same(gpu, ref, "depth 16 cpu3")
    assert (gpu["node"] != svo.MISS).sum() > 2000


def test_degenerate_frames_every_shape(menger8):
    import oracle
    from libsvo_amd import workloads
    sc, base = menger8
    ib = workloads.incore_bytes_for(sc)
    mm = svo.TreeletMemoryManagerHip(sc, ib, 0, resident_all=True)
    ora = oracle.OracleScene(base, ib)
    sizes = [(1, 1), (1, 64), (64, 1), (3, 2), (65, 1), (1, 129)]
    views = [svo.View(mm, workloads.sphere_camera(900 + i, w, h), svo.Frame.alloc(w, h)) for i, (w, h) in enumerate(sizes)]
    refs = [ora.render(v.camera, "treelet") for v in views]
    r = svo.Renderer(0)
    for shape in svo.LAUNCH_SHAPES:
        r.config = shape
        for v in views:
            v.frame.node.fill(0xAB)
            v.frame.rgba.fill(0xAB)
        r.render(views, "treelet")
        svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
        for (w, h), v, ref in zip(sizes, views, refs):
            _same(v.frame.fetch(), ref, f"{w}x{h} shape {shape}")


def test_largest_frame_sides(menger8):
    """The largest frame sides the renderer takes (16384 pixels) as a row of
    16384 x 3 and a column of 5 x 16384, next to each other in one launch,
    bit-identical to the oracle; a side of 16385 is refused."""
    import oracle
    from libsvo_amd import workloads
    sc, base = menger8
    ib = workloads.incore_bytes_for(sc)
    mm = svo.TreeletMemoryManagerHip(sc, ib, 0, resident_all=True)
    ora = oracle.OracleScene(base, ib)
    sizes = [(16384, 3), (5, 16384)]
    views = [svo.View(mm, workloads.sphere_camera(950 + i, w, h), svo.Frame.alloc(w, h)) for i, (w, h) in enumerate(sizes)]
    refs = [ora.render(v.camera, "treelet") for v in views]
    r = svo.Renderer(0)
    for v in views:
        v.frame.node.fill(0xAB)
        v.frame.rgba.fill(0xAB)
    _say("frames 16384x3 and 5x16384 in one launch")
    r.render(views, "treelet")
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    for (w, h), v, ref in zip(sizes, views, refs):
        g = v.frame.fetch()
        _same(g, ref, f"{w}x{h}")
        assert (g["node"] != svo.MISS).sum() > 0, f"{w}x{h}: no hit"
    for w, h in [(16385, 1), (1, 16385)]:
        with pytest.raises(svo.SvoError):
            r.render([svo.View(mm, workloads.sphere_camera(960, w, h), svo.Frame.alloc(1, 1))], "treelet")


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_eyes_on_octree_planes(menger8, policy):
    """Eyes exactly on the octree's planes: on the root's midplanes (0 and
    the negative zero), on deeper dyadic planes, on a face of the cube and
    inside it on a node corner.  Plane parameters of such rays are exact
    zeros (signed as the IEEE rules give them) and entries land exactly on
    edges, so every select and comparison of the traversal decides ties
    there; bit-identical to the oracle."""
    import oracle
    from libsvo_amd import workloads
    sc, base = menger8
    ib = workloads.incore_bytes_for(sc)
    mm = svo.TreeletMemoryManagerHip(sc, ib, 0, resident_all=True)
    ora = oracle.OracleScene(base, ib)
    eyes = [((0.0, 0.0, 2.5), (0.0, 0.0, 0.0)), ((-0.0, 0.5, 2.0), (0.1, 0.05, 0.0)),
            ((0.0, -0.0, -2.0), (0.0, 0.0, 0.0)), ((0.25, 0.125, 1.0), (-0.2, 0.1, -0.5)),
            ((1.0, 0.5, 0.25), (-1.0, 0.0, 0.0)), ((0.25, 0.25, 0.25), (0.0, 0.0, 0.0)),
            ((-0.0, -0.0, -0.0), (0.3, 0.2, 0.1))]
    r = svo.Renderer(0)
    for i, (eye, target) in enumerate(eyes):
        cam = svo.look_at(eye, target, (0.0, 1.0, 0.0), 60.0, 64, 48, far=20.0)
        f = svo.Frame.alloc(64, 48)
        f.node.fill(0xAB)
        f.rgba.fill(0xAB)
        _say(f"eye {eye} -> {target}, policy={policy}")
        r.render([svo.View(mm, cam, f)], policy)
        svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
        _same(f.fetch(), ora.render(cam, policy), f"eye {eye}")


def test_single_leaf_scene(tmp_path):
    code = H.morton_np(np.array([5]), np.array([2]), np.array([6]), 3)
    attr = np.array([[0x80FF8000, 0x11223344]], np.uint32)
    sc = svo.Scene.from_leaves(3, code, attr, 0, "dense")
    mm = svo.TreeletMemoryManagerHip(sc, sc.info["treelet_bytes"], 0, resident_all=True)
    # looking at the leaf's centre, then away from the cube
    centre = tuple(((np.array([5, 2, 6]) + 0.5) / 8.0 - 0.5).tolist())
    for eye, target, want_hits in [((1.2, 0.9, 1.4), centre, True), ((1.2, 0.9, 1.4), (3.0, 3.0, 3.0), False)]:
        cam = svo.look_at(eye, target, (0.0, 1.0, 0.0), 20.0, 48, 40)
        for policy in ("treelet", "cpu3"):
            gpu = svo.render(mm, cam, policy)
            ref = _oracle(sc, tmp_path, "one_leaf", cam, policy)
            _same(gpu, ref, f"single leaf {policy} hits={want_hits}")
            hits = gpu["node"] != svo.MISS
            assert bool(hits.any()) == want_hits
            if want_hits:
                assert np.all(gpu["rgba"][hits] == 0x112233FF)
            else:
                assert np.all(gpu["rgba"] == svo.BACKGROUND_RGBA)


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_far_camera_maxloops_rays(tmp_path, policy):
    """A camera 2000 units away drives rays into maxLoops (the epsilon
    vanishes against t): pixels, push counts and loop counts (the loops view
    mode) match the oracle, including the rays whose cycles the kernel skips
    by whole periods."""
    import oracle
    from libsvo_amd import workloads
    sc = workloads.menger_scene(depth=6, level=3, treelet_bytes=4096)
    base = str(tmp_path / "menger6")
    sc.write(base)
    ib = workloads.incore_bytes_for(sc)
    mm = svo.TreeletMemoryManagerHip(sc, ib, 0, resident_all=True)
    ora = oracle.OracleScene(base, ib)
    eye = np.array([0.7, 0.45, 0.55])
    eye = eye / np.linalg.norm(eye) * 2000.0
    cam = svo.with_view_mode(svo.look_at(tuple(eye), (0.01, 0.02, 0.03), (0.0, 1.0, 0.0), 0.09, 64, 64,
                                         far=4000.0), "loops")
    ref = ora.render(cam, policy, shade=True)
    # every stack layout (the node-holding one finds the node a maxLoops ray
    # reports by walking its path from the root)
    for stack in ("lds", "lds_node", "private"):
        r = svo.Renderer(0)
        r.config = (8, 1, stack)
        f = svo.Frame.alloc(cam.width, cam.height, pushes=True, shade=True)
        r.render([svo.View(mm, cam, f)], policy)
        svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
        gpu = f.fetch()
        _same(gpu, ref, f"far camera {policy} {stack}")
        assert np.array_equal(gpu["pushes"], ref["pushes"]), stack
        assert np.array_equal(gpu["shade"].view(np.uint32), ref["shade"].view(np.uint32)), stack
    max_loops = 14 * 14 * 5 if policy == "treelet" else 7 * 7 * 5
    assert int((ref["loops"] == max_loops).sum()) > 40
