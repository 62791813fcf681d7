"""GPU: the view modes of renderToBuffer in the fused epilogue, against the
oracle's restatement (pinned on the CPU by tests/test_view_modes.py).  Modes
built from +, *, /, sqrt and table lookups are bit-identical; the colour map
and phong call pow, whose device and host implementations may differ in the
last bit: tolerance 2e-6 (phong) and one 1/255 step on at most 0.1 % of the
pixels (colour map, where pow feeds a rounding)."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads

pytestmark = pytest.mark.gpu

EXACT = ["color", "normal", "svo_depth", "t", "loops", "quality", "fill_svo_depth", "fill_loops"]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


@pytest.fixture(scope="module")
def scenes(menger8, terrain8):
    import oracle
    out = {}
    for name, (sc, base) in (("menger", menger8), ("terrain", terrain8)):
        ib = workloads.incore_bytes_for(sc)
        out[name] = (svo.TreeletMemoryManagerHip(sc, ib, 0, resident_all=True), oracle.OracleScene(base, ib))
    return out


def _cam(name, mode, lod):
    cam = (workloads.sphere_camera(21, 200, 150, lod=lod) if name == "menger"
           else workloads.terrain_camera(22, 200, 150, lod=lod, terrain_seed=workloads.TERRAIN_SEED))
    return svo.with_view_mode(cam, mode)


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
@pytest.mark.parametrize("mode", EXACT + ["iterations", "phong"])
def test_view_mode_matches_oracle(scenes, mode, policy):
    for name, (mm, ora) in scenes.items():
        for lod in (0.0, 1.0):
            cam = _cam(name, mode, lod)
            print(f"{name} {mode} {policy} lod={lod}", flush=True)
            g = svo.render(mm, cam, policy, shade=True)
            r = ora.render(cam, policy, shade=True)
            assert np.array_equal(g["node"], r["node"])
            gs, rs = g["shade"], r["shade"]
            if mode in EXACT:
                same = (gs.view(np.uint32) == rs.view(np.uint32)) | (np.isnan(gs) & np.isnan(rs))
                assert same.all(), f"{name} {mode}: {int((~same).sum())} values differ"
            elif mode == "phong":
                assert np.allclose(gs, rs, rtol=0, atol=2e-6)
            else:
                diff = np.abs(gs - rs).max(axis=-1)
                assert diff.max() <= 1.0 / 255 + 1e-6 and (diff > 0).mean() <= 1e-3


def test_unknown_view_mode_rejected(scenes):
    mm, _ = scenes["menger"]
    cam = _cam("menger", 0, 0.0)
    cam.view_mode = 12
    with pytest.raises(svo.SvoError):
        svo.render(mm, cam, "treelet", shade=True)
    # the shadow and reflection modes trace single-SVO rays: other policies refuse them
    for mode in ("fill_shadow", "fill_reflection"):
        for policy in ("treelet", "cpu3"):
            with pytest.raises(svo.SvoError, match="SVO_POLICY_SINGLE"):
                svo.render(mm, _cam("menger", mode, 0.0), policy, shade=True)


def test_single_svo_policy_view_modes(tmp_path, menger8):
    import oracle
    sc, _ = menger8
    base = str(tmp_path / "m_v1")
    n = sc.treelet(0)[0]["num_nodes"]
    sc.write_single(base, "v1", np.random.default_rng(4).permutation(n).astype(np.uint32))
    single = svo.Scene.open_single(base)
    mm = svo.TreeletMemoryManagerHip(single, workloads.incore_bytes_for(single), 0, resident_all=True)
    ora = oracle.OracleSingle(base)
    for mode in EXACT + ["fill_shadow", "fill_reflection"]:
        cam = svo.with_view_mode(workloads.sphere_camera(9, 160, 120, lod=0.5), mode)
        print(f"single-SVO policy, view mode {mode}", flush=True)
        g = svo.render(mm, cam, "single", shade=True)
        r = ora.render(cam, shade=True)
        same = (g["shade"].view(np.uint32) == r["shade"].view(np.uint32)) | (np.isnan(g["shade"]) & np.isnan(r["shade"]))
        assert same.all(), f"single {mode}: {int((~same).sum())} values differ"


def test_single_svo_secondary_modes_in_a_replayed_batch(tmp_path, menger8):
    """The shadow and reflection modes inside a hipGraph batch of the
    single-SVO policy: replays, with eager work between them, and a camera
    update that switches the views' modes, all equal to the oracle."""
    import oracle
    sc, _ = menger8
    base = str(tmp_path / "m_v1b")
    n = sc.treelet(0)[0]["num_nodes"]
    sc.write_single(base, "v1", np.random.default_rng(5).permutation(n).astype(np.uint32))
    single = svo.Scene.open_single(base)
    mm = svo.TreeletMemoryManagerHip(single, workloads.incore_bytes_for(single), 0, resident_all=True)
    ora = oracle.OracleSingle(base)
    modes = ["fill_shadow", "fill_reflection", "color"]
    cams = [svo.with_view_mode(workloads.sphere_camera(30 + i, 96 + 8 * i, 72, lod=0.25 * i), m)
            for i, m in enumerate(modes)]
    views = [svo.View(mm, c, svo.Frame.alloc(c.width, c.height, shade=True)) for c in cams]
    r = svo.Renderer(0)
    s = svo.Stream()
    b = svo.Batch(r, views, "single", s)

    def check(what):
        for v in views:
            g = v.frame.fetch()
            ref = ora.render(v.camera, shade=True)
            assert np.array_equal(g["node"], ref["node"]), what
            same = g["shade"].view(np.uint32) == ref["shade"].view(np.uint32)
            assert same.all(), f"{what}: {int((~same).sum())} values differ"

    side = svo.View(mm, svo.with_view_mode(workloads.sphere_camera(3, 40, 30), "fill_shadow"),
                    svo.Frame.alloc(40, 30, shade=True))
    for rep in range(3):
        r.render([side], "single", s)
        print(f"secondary-mode batch replay {rep}", flush=True)
        b.launch(s)
        s.synchronize()
        check(f"replay {rep}")
    b.set_cameras([svo.with_view_mode(c, m) for c, m in zip(cams, ["fill_reflection", "fill_svo_depth", "fill_shadow"])], s)
    b.launch(s)
    s.synchronize()
    check("after the mode switch")
    # a batch of another policy refuses the secondary modes
    t = svo.Batch(r, [svo.View(mm, svo.with_view_mode(cams[0], "color"), svo.Frame.alloc(96, 72, shade=True))],
                  "treelet", s)
    with pytest.raises(svo.SvoError, match="SVO_POLICY_SINGLE"):
        t.set_cameras([svo.with_view_mode(cams[0], "fill_reflection")], s)
