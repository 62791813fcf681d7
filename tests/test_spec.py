"""Run specs (libsvo_amd/spec.py): checking, defaults, scene and camera
building, on the CPU.  The settings a spec carries are the OpenCL app's
globals (projects/raycast_serialized_opencl/src/raycast_serialized_opencl.cpp:
30-35,76,246-250) and compile-time constants (opencl/treeletRenderer_rgba.cl:7-8)."""
import json
import subprocess
import sys

import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import spec as S

BASE = {"scene": {"generator": "menger", "depth": 5, "level": 2},
        "views": [{"eye": [2.0, 1.0, 2.0], "target": [0.0, 0.0, 0.0], "width": 64, "height": 48}]}


def test_defaults_filled_in():
    s = S.load(BASE)
    assert s["incore_bytes"] == S.DEFAULT_INCORE_BYTES == 512 * 1024 * 1024
    assert s["policy"] == "treelet" and s["analyse"] is None and s["launch"] == "default" and s["frames"] == 1
    assert s["scene"] == {"generator": "menger", "depth": 5, "level": 2, "seed": 1, "treelet_bytes": 0,
                          "format": "treelet"}
    v = s["views"][0]
    assert v["up"] == [0.0, 1.0, 0.0] and v["fov_y_deg"] == 60.0 and v["view_mode"] == "color" and v["lod"] == 0.0
    assert BASE["views"][0].get("up") is None  # the caller's dict is not changed
    # the same spec as a JSON string and as a file
    assert S.load(json.dumps(BASE)) == s


def test_single_files_default_to_the_single_policy():
    s = S.load({"scene": {"path": "/x/branch.svo", "format": "single"}, "views": BASE["views"]})
    assert s["policy"] == "single"


@pytest.mark.parametrize("patch, key", [
    ({"scene": None}, "scene"),
    ({"scene": {"generator": "teapot"}}, "scene.generator"),
    ({"scene": {"generator": "menger", "depth": 0}}, "scene.depth"),
    ({"scene": {"generator": "menger", "colour": 1}}, "scene"),
    ({"scene": {"path": "a", "generator": "menger"}}, "scene"),
    ({"scene": {"path": "a", "format": "octree"}}, "scene.format"),
    ({"incore_bytes": -5}, "incore_bytes"),
    ({"incore_bytes": "most"}, "incore_bytes"),
    ({"policy": "fast"}, "policy"),
    ({"policy": "single"}, "policy"),
    ({"analyse": {"divide": 0}}, "analyse.divide"),
    ({"analyse": {"divide": 8, "jitter": 1}}, "analyse"),
    ({"launch": [12, 1, "lds"]}, "launch"),
    ({"launch": [8, 1, "global"]}, "launch"),
    ({"launch": "fastest"}, "launch"),
    ({"tune_cache": "t.tsv"}, "tune_cache"),
    ({"frames": 0}, "frames"),
    ({"views": []}, "views"),
    ({"views": [{"eye": [0, 0, 1]}]}, "views[0].target"),
    ({"views": [{"eye": [0, 0, 1], "target": [0, 0, 1]}]}, "views[0]"),
    ({"views": [{"eye": [0, 2, 0], "target": [0, 0, 0]}]}, "views[0].up"),
    ({"views": [{"eye": [0, 0, 2], "target": [0, 0, 0], "width": 0}]}, "views[0].width"),
    ({"views": [{"eye": [0, 0, 2], "target": [0, 0, 0], "fov_y_deg": 180}]}, "views[0].fov_y_deg"),
    ({"views": [{"eye": [0, 0, 2], "target": [0, 0, 0], "view_mode": "xray"}]}, "views[0].view_mode"),
    ({"views": [{"eye": [0, 0, 2], "target": [0, 0, 0], "zoom": 2}]}, "views[0]"),
    ({"speed": 1}, "spec"),
])
def test_bad_specs_name_the_key(patch, key):
    s = json.loads(json.dumps(BASE))
    s.update(patch)
    with pytest.raises(S.SpecError) as e:
        S.load(s)
    assert str(e.value).startswith(key + ":"), str(e.value)


def test_scene_and_cameras_match_the_direct_calls():
    s = S.load(dict(BASE, views=[dict(BASE["views"][0], view_mode="phong", lod=1.5, fov_y_deg=45, far=9.0)]))
    sc = S.open_scene(s)
    ref = svo.Scene.menger(2, 5, 1, 0)
    assert sc.info == ref.info
    a, b = sc.treelet(0), ref.treelet(0)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    cam = S.cameras(s)[0]
    direct = svo.with_view_mode(svo.look_at((2, 1, 2), (0, 0, 0), (0, 1, 0), 45, 64, 48, 9.0, 1.5), "phong")
    assert bytes(cam) == bytes(direct)


def test_image_rgb_unpacks_colour_and_clamps_shade():
    rgba = np.array([[0x11223344, 0xFF0080FF]], np.uint32)
    assert image_eq(S.image_rgb({"rgba": rgba}, "color"), [[[0x11, 0x22, 0x33], [0xFF, 0x00, 0x80]]])
    shade = np.array([[[-1.0, 0.5, 2.0, 1.0], [0.0, 1.0, 0.25, 1.0]]], np.float32)
    assert image_eq(S.image_rgb({"rgba": rgba, "shade": shade}, "normal"), [[[0, 128, 255], [0, 255, 64]]])


def image_eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b, np.uint8))


def test_cli_check_prints_the_filled_spec(tmp_path):
    p = tmp_path / "s.json"
    p.write_text(json.dumps(BASE))
    out = subprocess.run([sys.executable, "-m", "libsvo_amd.spec", str(p), "--check"], capture_output=True, text=True,
                         cwd=str(tmp_path.parent), env=_env())
    assert out.returncode == 0, out.stderr
    assert json.loads(out.stdout) == S.load(BASE)
    bad = tmp_path / "b.json"
    bad.write_text(json.dumps(dict(BASE, frames=-1)))
    out = subprocess.run([sys.executable, "-m", "libsvo_amd.spec", str(bad), "--check"], capture_output=True,
                         text=True, env=_env())
    assert out.returncode == 2 and "frames:" in out.stderr


def _env():
    import os
    env = dict(os.environ)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env["PYTHONPATH"] = root + os.pathsep + env.get("PYTHONPATH", "")
    return env
