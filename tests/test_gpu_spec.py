"""GPU: a run spec (libsvo_amd/spec.py) rendered end to end -- generated
scene, out-of-core residency with the analyse pass, autotuned launch shape
through a cache file -- and its images checked against the oracle over the
same scene file (opencl/treeletRenderer_rgba.cl:201-382)."""
import os

import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import spec as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _read_ppm(path):
    with open(path, "rb") as f:
        data = f.read()
    head = data.split(b"\n", 3)
    assert head[0] == b"P6" and head[2] == b"255"
    w, h = (int(x) for x in head[1].split())
    return np.frombuffer(head[3], np.uint8).reshape(h, w, 3)


def test_resident_spec_images_match_oracle(tmp_path):
    import oracle
    spec = {"scene": {"generator": "terrain", "depth": 8, "treelet_bytes": 8192}, "incore_bytes": "all",
            "launch": "auto", "tune_cache": str(tmp_path / "tune.tsv"), "frames": 2,
            "views": [{"eye": [-0.3, 0.2, -0.3], "target": [0.2, -0.1, 0.25], "width": 160, "height": 96},
                      {"eye": [0.4, 0.3, -0.4], "target": [0.0, -0.2, 0.0], "width": 75, "height": 50,
                       "view_mode": "normal"}]}
    print("spec run, resident terrain d8, 2 views", flush=True)
    res = S.run(spec, str(tmp_path / "img"))
    assert res["autotune"]["cache_hit"] is False and os.path.exists(spec["tune_cache"])
    assert len(res["frames"]) == 2 and len(res["images"]) == 2
    res2 = S.run(spec, str(tmp_path / "img2"))
    assert res2["autotune"]["cache_hit"] is True and res2["launch_shape"] == res["launch_shape"]
    # oracle over the same scene written to disk
    sc = S.open_scene(S.load(spec))
    base = str(tmp_path / "scene")
    sc.write(base)
    ora = oracle.OracleScene(base, sc.info["treelet_bytes"] * sc.info["num_treelets"])
    for i, cam in enumerate(S.cameras(S.load(spec))):
        ref = ora.render(cam, "treelet", shade=True)
        mode = spec["views"][i].get("view_mode", "color")
        want = S.image_rgb(ref, mode)
        for r in (res, res2):
            got = _read_ppm(r["images"][i])
            assert np.array_equal(got, want), f"view {i}: {np.count_nonzero(np.any(got != want, axis=-1))} pixels"


def test_out_of_core_spec_pages_in_and_matches_oracle(tmp_path):
    import oracle
    spec = {"scene": {"generator": "terrain", "depth": 8, "treelet_bytes": 4096}, "incore_bytes": 48 * 4096,
            "analyse": {"divide": 8, "rumble": True}, "frames": 4, "launch": [16, 1, "lds_node"],
            "views": [{"eye": [-0.35, 0.05, -0.35], "target": [0.3, -0.15, 0.3], "width": 128, "height": 80,
                       "lod": 1.0}]}
    print("spec run, out-of-core terrain d8, 48 slots", flush=True)
    res = S.run(spec, str(tmp_path / "img"), keep=True)
    assert sum(f["page_ins"] for f in res["frames"]) > 0
    assert res["launch_shape"] == [16, 1, "lds_node"]
    assert all(f["ms_render"] > 0 for f in res["frames"])
    # the last frame against the oracle over the manager's incore buffer
    mm, view = res["mm"], res["views"][0]
    n = mm.info["pool_nodes"]
    nodes, attr = mm.getIncoreBuffer(0, n)
    slots = np.asarray(mm.getSlots(), np.int32)
    ref = oracle.render_buffers(nodes, attr[:, 1], slots, mm.getNumNodesPerTreelet(), mm.scene.info["max_depth"],
                                view.camera, "treelet")
    got = view.frame.fetch()
    for k in ("node", "treelet", "rgba"):
        assert np.array_equal(got[k], ref[k]), f"{k} differs on {np.count_nonzero(got[k] != ref[k])} pixels"
    assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert np.array_equal(_read_ppm(res["images"][0]), S.image_rgb(ref, "color"))
