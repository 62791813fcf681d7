"""GPU: legacy single-SVO files rendered with the single-SVO renderer's
policy (fillFrameBuffer_rgba.cl::sample) against the oracle, which reads the
same files with its own parser and traverses them with absolute indexing:
node, treelet, colour (through the v1 attribute indices) and t bit-identical."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("fmt", ["v1", "svoloader"])
def test_single_svo_policy_matches_oracle(tmp_path, menger8, fmt):
    import oracle
    sc, _ = menger8
    n = sc.treelet(0)[0]["num_nodes"]
    base = str(tmp_path / f"menger_{fmt}")
    perm = np.random.default_rng(8).permutation(n).astype(np.uint32) if fmt == "v1" else None
    sc.write_single(base, fmt, perm)
    single = svo.Scene.open_single(base)
    mm = svo.TreeletMemoryManagerHip(single, workloads.incore_bytes_for(single), 0, resident_all=True)
    ora = oracle.OracleSingle(base)
    for i, lod in enumerate((0.0, 1.0, 4.0)):
        cam = workloads.sphere_camera(40 + i, 300, 220, lod=lod)
        print(f"single-SVO {fmt} lod={lod}", flush=True)
        gpu = svo.render(mm, cam, "single")
        ref = ora.render(cam)
        _same(gpu, ref, f"{fmt} lod={lod}")
        assert np.count_nonzero(gpu["node"] != svo.MISS) > 1000


def test_single_policy_needs_one_treelet(terrain8):
    sc, _ = terrain8
    mm = svo.TreeletMemoryManagerHip(sc, workloads.incore_bytes_for(sc), 0, resident_all=True)
    with pytest.raises(svo.SvoError):
        svo.render(mm, workloads.terrain_camera(1, 32, 32), "single")


def test_push_below_the_stack_ends_the_ray(tmp_path):
    """Policy B has no depth stop: in a tree with an inner node at depth 15
    below every leaf (a chain of inner nodes towards the cube centre), a ray
    through the chain pushes from scale 14 down to scale 1 and would push a
    15th time from scale 0, below the 15-level stack.  Both the kernel and the
    oracle end such a ray as a miss (t = the root box exit) after 14 pushes;
    every other pixel matches too.  Regression test of the stack bound."""
    import oracle
    from tests.test_single_svo import _inner_chain, _write_fixture
    nodes = _inner_chain(15)
    packages = np.tile(np.array([[0x808080FF, 0x336699FF]], np.uint32), (nodes.shape[0], 1))
    base = str(tmp_path / "chain15")
    _write_fixture(base, "v1 15 %d " % nodes.shape[0], nodes, packages, aidx=list(range(nodes.shape[0])))
    single = svo.Scene.open_single(base)
    assert single.info["max_depth"] == 15
    mm = svo.TreeletMemoryManagerHip(single, workloads.incore_bytes_for(single), 0, resident_all=True)
    # the centre pixel's ray points at (1e-6,)*3, inside the depth-15 box [0, 2^-15]^3
    cam = svo.look_at((2.0, 2.1, 1.9), (1e-6, 1e-6, 1e-6), (0.0, 1.0, 0.0), 2.0, 64, 64)
    print("policy B, inner chain to depth 15", flush=True)
    gpu = svo.render(mm, cam, "single")
    ref = oracle.OracleSingle(base).render(cam)
    _same(gpu, ref, "inner chain")
    assert np.array_equal(gpu["pushes"], ref["pushes"])
    c = gpu["pushes"][32, 32]
    assert c == 14 and gpu["node"][32, 32] == svo.MISS, (c, gpu["node"][32, 32])
