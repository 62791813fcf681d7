"""GPU parity: the HIP traversal against the CPU oracle, through the C-ABI.

Bar: node index, treelet id, RGBA8 and the FP32 t are bit-identical to the
oracle under the same policy.  Across policies (GPU treelet policy vs the
oracle's CpuRaycasterSingleRay3 policy, LOD off) pixels may differ only where
the oracle flags a boundary ray; their number is reported.
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


def _compare(gpu, ref, what):
    for k in ("node", "treelet", "rgba", "pushes"):
        if k not in gpu or k not in ref:
            continue
        bad = np.count_nonzero(gpu[k] != ref[k])
        assert bad == 0, f"{what}: {k} differs on {bad} pixels"
    tb_g = gpu["t"].view(np.uint32)
    tb_r = ref["t"].view(np.uint32)
    bad = np.count_nonzero(tb_g != tb_r)
    assert bad == 0, f"{what}: t bits differ on {bad} pixels"


@pytest.fixture(scope="module")
def menger_mm(menger8):
    import oracle
    sc, base = menger8
    ib = workloads.incore_bytes_for(sc)
    mm = svo.TreeletMemoryManagerHip(sc, ib, device=0, resident_all=True)
    ora = oracle.OracleScene(base, ib)
    return sc, mm, ora


def test_ray_directions_bit_exact():
    import oracle
    for cam in (workloads.sphere_camera(5, 96, 64), svo.look_at((0, 0, 2.5), (0, 0, 0), (0, 1, 0), 60, 64, 64)):
        _say(f"ray dirs {cam.width}x{cam.height}")
        g = svo.ray_dirs(cam)
        r = oracle.ray_dirs(cam)
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_device_incore_matches_host(menger_mm):
    sc, mm, ora = menger_mm
    _say("device incore readback")
    n = mm.info["pool_nodes"]
    hn, ha = mm.getIncoreBuffer(0, n)
    dn, da = mm.getDeviceIncoreBuffer(0, n)
    assert np.array_equal(hn, dn) and np.array_equal(ha, da)


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
@pytest.mark.parametrize("lod", [0.0, 1.0])
def test_menger_512_matches_oracle(menger_mm, policy, lod):
    sc, mm, ora = menger_mm
    cam = workloads.sphere_camera(workloads.MENGER_CAMERA_SEED, 512, 512, lod=lod)
    _say(f"menger d8 512^2 policy={policy} lod={lod}")
    gpu = svo.render(mm, cam, policy)
    ref = ora.render(cam, policy)
    _compare(gpu, ref, f"menger {policy} lod={lod}")
    assert np.count_nonzero(gpu["node"] != svo.MISS) > 1000


def test_menger_many_cameras_against_cpu3(menger_mm):
    """GPU treelet policy vs the oracle's CPU-caster policy, LOD off: equal
    except on boundary rays."""
    sc, mm, ora = menger_mm
    total_b = 0
    for seed in range(8):
        cam = workloads.sphere_camera(1000 + seed, 256, 256)
        _say(f"menger cross-policy seed={seed}")
        gpu = svo.render(mm, cam, "treelet")
        ref = ora.render(cam, "cpu3")
        diff = (gpu["node"] != ref["node"]) | (gpu["treelet"] != ref["treelet"])
        flagged = ref["boundary"].astype(bool)
        assert not np.any(diff & ~flagged), f"seed {seed}: {np.count_nonzero(diff & ~flagged)} unflagged diffs"
        total_b += int(np.count_nonzero(diff))
    _say(f"boundary pixels differing across policies: {total_b}")


@pytest.mark.parametrize("fixture", ["terrain8", "terrain8_bfs"])
@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_terrain_multi_treelet_matches_oracle(request, fixture, policy):
    import oracle
    sc, base = request.getfixturevalue(fixture)
    ib = workloads.incore_bytes_for(sc, spare_slots=3)
    mm = svo.TreeletMemoryManagerHip(sc, ib, device=0, resident_all=True)
    ora = oracle.OracleScene(base, ib)
    for lod in (0.0, 1.0):
        cam = workloads.terrain_camera(21, 320, 200, lod=lod, terrain_seed=workloads.TERRAIN_SEED)
        _say(f"{fixture} policy={policy} lod={lod}")
        gpu = svo.render(mm, cam, policy)
        ref = ora.render(cam, policy)
        _compare(gpu, ref, f"{fixture} {policy} lod={lod}")
        assert len(set(np.unique(gpu["treelet"]).tolist()) - {svo.MISS}) > 1


def test_multi_camera_launch_equals_single(menger_mm):
    sc, mm, ora = menger_mm
    cams = [workloads.sphere_camera(300 + i, w, h) for i, (w, h) in enumerate([(64, 64), (100, 37), (7, 9), (128, 72)])]
    frames = [svo.Frame.alloc(c.width, c.height) for c in cams]
    _say("multi-camera launch")
    svo.render_to_buffer(mm, cams, frames, "treelet")
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    for c, f in zip(cams, frames):
        alone = svo.render(mm, c, "treelet")
        _compare(f.fetch(), alone, f"batched {c.width}x{c.height}")
        _compare(alone, ora.render(c, "treelet"), f"oracle {c.width}x{c.height}")


def test_axis_aligned_and_inside_cameras(menger_mm):
    """Rays with exactly-zero direction components and an eye inside the
    cube exercise the 1/0 and sign(0) paths."""
    sc, mm, ora = menger_mm
    cams = [svo.look_at((0, 0, 2.5), (0, 0, 0), (0, 1, 0), 60, 64, 64),
            svo.look_at((2.5, 0, 0), (0, 0, 0), (0, 1, 0), 90, 32, 32),
            svo.look_at((0.01, 0.02, 0.03), (0.3, 0.1, -0.2), (0, 1, 0), 90, 64, 48)]
    for i, cam in enumerate(cams):
        d = svo.ray_dirs(cam)
        for policy in ("treelet", "cpu3"):
            _say(f"special camera {i} policy={policy} zero-components={int(np.count_nonzero(d == 0))}")
            _compare(svo.render(mm, cam, policy), ora.render(cam, policy), f"special camera {i} {policy}")


@pytest.fixture(scope="module")
def terrain11(scene_dir):
    sc = workloads.terrain_scene(depth=11)
    base = str(scene_dir / "terrain11")
    sc.write(base)
    return sc, base


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_full_size_terrain_1080p_matches_oracle(terrain11, policy):
    """The bench frame itself (depth-11 terrain, 4096-node treelets, 1920x1080):
    bit-identical to the oracle, including the rays that cycle until maxLoops
    (the kernel skips whole periods of such cycles)."""
    import oracle
    sc, base = terrain11
    ib = workloads.incore_bytes_for(sc)
    mm = svo.TreeletMemoryManagerHip(sc, ib, device=0, resident_all=True)
    ora = oracle.OracleScene(base, ib)
    cam = workloads.terrain_camera(workloads.TERRAIN_CAMERA_SEED, 1920, 1080)
    _say(f"terrain d11 1080p policy={policy}")
    gpu = svo.render(mm, cam, policy)
    ref = ora.render(cam, policy)
    _compare(gpu, ref, f"terrain d11 1080p {policy}")
    max_loops = 14 * 14 * 5 if policy == "treelet" else 12 * 12 * 5
    cycling = int(np.count_nonzero(ref["loops"] >= max_loops))
    _say(f"rays reaching maxLoops: {cycling}; boundary rays: {int(ref['boundary'].sum())}")
    if policy == "treelet":
        assert cycling > 0  # the frame exercises the cycle-skip path
        other = ora.render(cam, "cpu3")
        diff = (gpu["node"] != other["node"]) | (gpu["treelet"] != other["treelet"])
        flagged = (ref["boundary"] | other["boundary"]).astype(bool)
        _say(f"treelet policy vs CPU-caster policy: {int(diff.sum())} differing pixels, all flagged")
        assert not np.any(diff & ~flagged)
