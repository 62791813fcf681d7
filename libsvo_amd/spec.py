"""Run specifications: a JSON document naming the scene, the residency pool,
the cameras, the traversal policy and the launch shape of a render loop.

The OpenCL app keeps these settings in globals and compile-time constants:
screen size and render-buffer divide (projects/raycast_serialized_opencl/
src/raycast_serialized_opencl.cpp:30-35, 246-250), the incore buffer size
(:247, 512 MB), the analyse-pass divide (:76, 8), the scene's base path
(:315) and the traversal stack depth (opencl/treeletRenderer_rgba.cl:7-8).
A spec carries the same settings as data:

    {
      "scene":  {"path": "/data/terrain"}                      treelet bundle (<base>.svo + <base>.ia)
              | {"path": "/data/branch.svo", "format": "single"}   SvoLoader / SvoBranch v1 file
              | {"generator": "terrain", "depth": 11, "seed": 7, "treelet_bytes": 32768}
              | {"generator": "menger", "depth": 8, "level": 4}
              | {"generator": "spheres", "depth": 10, "count": 2000},
      "incore_bytes": 536870912 | "all",        default 512 MiB, "all" = every treelet resident
      "device": 0,
      "policy": "treelet" | "single" | "cpu3",  default "treelet" ("single" for single files)
      "analyse": {"divide": 8, "rumble": false} | null,   out-of-core feedback per frame
      "launch": "default" | "auto" | [tile_width, waves, "lds" | "private" | "lds_node"],
      "tune_cache": "/path/tune.tsv" | null,    autotune picks kept between runs
      "frames": 1,                              frames rendered per view
      "views": [{"eye": [x, y, z], "target": [x, y, z], "up": [0, 1, 0],
                 "fov_y_deg": 60, "width": 1024, "height": 768, "far": 20,
                 "lod": 0.0, "view_mode": "color"}]
    }

:func:`load` checks a spec and fills in the defaults (errors name the key
that is wrong); :func:`run` builds the scene and the residency handle,
renders every frame of every view in one launch per frame (with the
analyse → feedback → upload steps before it when ``analyse`` is set), and
returns timings and, when asked, writes each view's image as a PPM.

    python -m libsvo_amd.spec my_spec.json --out images/
"""
from __future__ import annotations

import argparse
import copy
import json
import math
import os
import sys
import time
from typing import Any, Dict, List, Optional

import numpy as np

DEFAULT_INCORE_BYTES = 512 * 1024 * 1024
GENERATORS = {
    "terrain": {"depth": 11, "seed": 7, "treelet_bytes": 32768},
    "menger": {"depth": 8, "level": 4, "seed": 1, "treelet_bytes": 0},
    "spheres": {"depth": 10, "count": 2000, "r_min": 0.005, "r_max": 0.05, "seed": 5, "treelet_bytes": 32768},
}
VIEW_DEFAULTS = {"up": [0.0, 1.0, 0.0], "fov_y_deg": 60.0, "width": 1024, "height": 768, "far": 20.0,
                 "lod": 0.0, "view_mode": "color"}
STACK_NAMES = ("lds", "private", "lds_node")


class SpecError(ValueError):
    """A spec that cannot be run; the message names the offending key."""


def _need(cond: bool, where: str, what: str) -> None:
    if not cond:
        raise SpecError(f"{where}: {what}")


def _number(v: Any) -> bool:
    return isinstance(v, (int, float)) and not isinstance(v, bool) and math.isfinite(float(v))


def _int(v: Any) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


def _vec3(v: Any) -> bool:
    return isinstance(v, (list, tuple)) and len(v) == 3 and all(_number(x) for x in v)


def load(source) -> Dict[str, Any]:
    """A checked copy of a spec (a dict, a JSON string or a path to a JSON
    file) with every default filled in."""
    if isinstance(source, dict):
        spec = copy.deepcopy(source)
    elif isinstance(source, str) and source.lstrip().startswith("{"):
        spec = json.loads(source)
    else:
        with open(source) as f:
            spec = json.load(f)
    _need(isinstance(spec, dict), "spec", "must be a JSON object")
    known = {"scene", "incore_bytes", "device", "policy", "analyse", "launch", "tune_cache", "frames", "views"}
    extra = sorted(set(spec) - known)
    _need(not extra, "spec", f"unknown keys {extra}")

    sc = spec.get("scene")
    _need(isinstance(sc, dict), "scene", "required, an object with 'path' or 'generator'")
    if "path" in sc:
        _need("generator" not in sc, "scene", "give 'path' or 'generator', not both")
        _need(isinstance(sc["path"], str) and sc["path"], "scene.path", "must be a non-empty string")
        fmt = sc.setdefault("format", "treelet")
        _need(fmt in ("treelet", "single"), "scene.format", f"'treelet' or 'single', not {fmt!r}")
        _need(set(sc) <= {"path", "format"}, "scene", f"unknown keys {sorted(set(sc) - {'path', 'format'})}")
    else:
        gen = sc.get("generator")
        _need(gen in GENERATORS, "scene.generator", f"one of {sorted(GENERATORS)}, not {gen!r}")
        params = GENERATORS[gen]
        extra = sorted(set(sc) - set(params) - {"generator"})
        _need(not extra, "scene", f"unknown keys {extra} for generator {gen!r}")
        for k, d in params.items():
            v = sc.setdefault(k, d)
            if isinstance(d, int):
                _need(_int(v) and v >= 0, f"scene.{k}", f"must be a non-negative integer, not {v!r}")
            else:
                _need(_number(v) and v > 0, f"scene.{k}", f"must be a positive number, not {v!r}")
        _need(1 <= sc["depth"] <= 15, "scene.depth", f"must lie in 1..15, not {sc['depth']}")
        sc["format"] = "treelet"

    ib = spec.setdefault("incore_bytes", DEFAULT_INCORE_BYTES)
    _need(ib == "all" or (_int(ib) and ib > 0), "incore_bytes", f"a positive byte count or 'all', not {ib!r}")
    dev = spec.setdefault("device", 0)
    _need(_int(dev) and dev >= 0, "device", f"must be a device index, not {dev!r}")
    pol = spec.setdefault("policy", "single" if sc["format"] == "single" else "treelet")
    _need(pol in ("treelet", "single", "cpu3"), "policy", f"'treelet', 'single' or 'cpu3', not {pol!r}")
    _need(pol != "single" or sc["format"] == "single" or ib == "all", "policy",
          "'single' traverses one resident treelet: use a single-SVO file or incore_bytes 'all' on a one-treelet scene")

    an = spec.setdefault("analyse", None)
    if an is not None:
        _need(isinstance(an, dict), "analyse", "an object or null")
        _need(set(an) <= {"divide", "rumble"}, "analyse", f"unknown keys {sorted(set(an) - {'divide', 'rumble'})}")
        dv = an.setdefault("divide", 8)
        _need(_int(dv) and 1 <= dv <= 64, "analyse.divide", f"an integer in 1..64, not {dv!r}")
        rb = an.setdefault("rumble", False)
        _need(isinstance(rb, bool), "analyse.rumble", "true or false")
        _need(pol != "single", "analyse", "the single-SVO policy has no analysis pass")

    ln = spec.setdefault("launch", "default")
    if isinstance(ln, str):
        _need(ln in ("default", "auto"), "launch", f"'default', 'auto' or [tile_width, waves, stack], not {ln!r}")
    else:
        _need(isinstance(ln, (list, tuple)) and len(ln) == 3 and _int(ln[0]) and _int(ln[1]) and ln[2] in STACK_NAMES
              and ln[0] in (8, 16, 32) and ln[1] in (1, 2, 4), "launch",
              f"[8|16|32, 1|2|4, {'|'.join(STACK_NAMES)}], not {ln!r}")
        spec["launch"] = list(ln)
    tc = spec.setdefault("tune_cache", None)
    _need(tc is None or (isinstance(tc, str) and tc), "tune_cache", "a file path or null")
    _need(tc is None or spec["launch"] == "auto", "tune_cache", "only used with launch 'auto'")
    fr = spec.setdefault("frames", 1)
    _need(_int(fr) and fr >= 1, "frames", f"a positive integer, not {fr!r}")

    from ._capi import VIEW_MODES
    views = spec.get("views")
    _need(isinstance(views, list) and views, "views", "required, a non-empty list")
    for i, v in enumerate(views):
        w = f"views[{i}]"
        _need(isinstance(v, dict), w, "must be an object")
        extra = sorted(set(v) - set(VIEW_DEFAULTS) - {"eye", "target"})
        _need(not extra, w, f"unknown keys {extra}")
        for k in ("eye", "target"):
            _need(_vec3(v.get(k)), f"{w}.{k}", "required, three numbers")
        for k, d in VIEW_DEFAULTS.items():
            v.setdefault(k, copy.copy(d))
        _need(_vec3(v["up"]), f"{w}.up", "three numbers")
        _need(float(np.linalg.norm(np.subtract(v["target"], v["eye"]))) > 0, w, "eye and target coincide")
        _need(float(np.linalg.norm(np.cross(np.subtract(v["target"], v["eye"]), v["up"]))) > 0, f"{w}.up",
              "parallel to the viewing direction")
        _need(_number(v["fov_y_deg"]) and 0 < v["fov_y_deg"] < 180, f"{w}.fov_y_deg", "must lie in (0, 180)")
        for k in ("width", "height"):
            _need(_int(v[k]) and 1 <= v[k] <= 16384, f"{w}.{k}", f"an integer in 1..16384, not {v[k]!r}")
        _need(_number(v["far"]) and v["far"] > 0, f"{w}.far", "must be positive")
        _need(_number(v["lod"]) and v["lod"] >= 0, f"{w}.lod", "must be >= 0")
        _need(v["view_mode"] in VIEW_MODES, f"{w}.view_mode", f"one of {sorted(VIEW_MODES)}, not {v['view_mode']!r}")
    return spec


def open_scene(spec: Dict[str, Any]):
    """The spec's scene (opened or generated) as a :class:`libsvo_amd.Scene`."""
    from . import Scene
    sc = spec["scene"]
    if "path" in sc:
        return Scene.open_single(sc["path"]) if sc["format"] == "single" else Scene.open(sc["path"])
    g = sc["generator"]
    if g == "terrain":
        return Scene.terrain(sc["depth"], sc["seed"], sc["treelet_bytes"])
    if g == "menger":
        return Scene.menger(sc["level"], sc["depth"], sc["seed"], sc["treelet_bytes"])
    return Scene.spheres(sc["depth"], sc["count"], sc["r_min"], sc["r_max"], sc["seed"], sc["treelet_bytes"])


def cameras(spec: Dict[str, Any]) -> List:
    """One :class:`libsvo_amd.Camera` per view of the spec."""
    from . import look_at, with_view_mode
    out = []
    for v in spec["views"]:
        c = look_at(v["eye"], v["target"], v["up"], v["fov_y_deg"], v["width"], v["height"], v["far"], v["lod"])
        out.append(with_view_mode(c, v["view_mode"]))
    return out


def _ppm(path: str, rgb: np.ndarray) -> None:
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(rgb, np.uint8).tobytes())


def image_rgb(frame: Dict[str, np.ndarray], view_mode: str) -> np.ndarray:
    """8-bit RGB of a fetched frame: the packed colour image in colour mode,
    else the view mode's FP32 image clamped to [0, 1]."""
    if view_mode == "color" or "shade" not in frame:
        p = frame["rgba"].astype(np.uint32)
        return np.stack([(p >> 24) & 255, (p >> 16) & 255, (p >> 8) & 255], axis=-1).astype(np.uint8)
    s = np.clip(frame["shade"][..., :3], 0.0, 1.0)
    return np.round(s * 255.0).astype(np.uint8)


def run(spec, out_dir: Optional[str] = None, keep: bool = False) -> Dict[str, Any]:
    """Runs a spec (see :func:`load`) on its device and returns what it did:
    the scene, the launch shape, per-frame timings and feedback statistics;
    with ``out_dir`` each view's last frame is written there as
    ``view<i>.ppm``; with ``keep`` the result also holds the residency
    handle and the views (``"mm"``, ``"views"``) for inspection."""
    import libsvo_amd as svo
    spec = load(spec)
    t0 = time.perf_counter()
    scene = open_scene(spec)
    info = scene.info
    t_scene = time.perf_counter() - t0
    from . import workloads
    resident_all = spec["incore_bytes"] == "all"
    incore = workloads.incore_bytes_for(scene) if resident_all else spec["incore_bytes"]
    mm = svo.TreeletMemoryManagerHip(scene, incore, spec["device"], resident_all=resident_all)
    cams = cameras(spec)
    shade = [v["view_mode"] != "color" for v in spec["views"]]
    views = [svo.View(mm, c, svo.Frame.alloc(c.width, c.height, shade=s)) for c, s in zip(cams, shade)]
    stream = svo.Stream()
    renderer = svo.Renderer(spec["device"])
    launch = spec["launch"]
    cache = None
    tuned = None
    if launch == "auto":
        cache = svo.TuneCache(spec["tune_cache"]) if spec["tune_cache"] else svo.TuneCache()
        shape, ms = renderer.autotune(views, spec["policy"], stream, iters=3, cache=cache)
        tuned = {"shape": list(shape), "ms": round(ms, 5), "cache_hit": cache.last_hit}
        if spec["tune_cache"]:
            cache.save(spec["tune_cache"])
    elif launch != "default":
        renderer.config = tuple(launch)
    an = spec["analyse"]
    rng = np.random.default_rng(0)
    frames = []
    for i in range(spec["frames"]):
        rec: Dict[str, Any] = {}
        a = time.perf_counter()
        if an is not None:
            jit = tuple(float(x) for x in rng.uniform(-1.0, 1.0, 2)) if an["rumble"] else (0.0, 0.0)
            for c in cams:
                mm.performAnalysePass(c, an["divide"], jit, spec["policy"], stream)
            stream.synchronize()
            st = mm.updateClientSideIncoreBuffer(stream)
            up = mm.updateDeviceMemory(stream)
            rec.update(page_ins=st["inserted"], evictions=st["evicted"], upload_bytes=up["bytes_moved"])
        b = time.perf_counter()
        renderer.render(views, spec["policy"], stream)
        stream.synchronize()
        c_ = time.perf_counter()
        rec.update(ms_feedback=round((b - a) * 1e3, 3), ms_render=round((c_ - b) * 1e3, 3))
        frames.append(rec)
    images = []
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        for i, (v, s) in enumerate(zip(views, spec["views"])):
            p = os.path.join(out_dir, f"view{i}.ppm")
            _ppm(p, image_rgb(v.frame.fetch(), s["view_mode"]))
            images.append(p)
    out = {
        "scene": {"treelets": info["num_treelets"], "tree_nodes": info["tree_nodes"], "max_depth": info["max_depth"],
                  "seconds_to_open": round(t_scene, 3)},
        "policy": spec["policy"], "launch_shape": list(renderer.config), "autotune": tuned,
        "resident_treelets": mm.info["resident_treelets"], "frames": frames, "images": images,
    }
    if keep:
        out["mm"], out["views"] = mm, views
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m libsvo_amd.spec", description="Render the views of a run spec.")
    ap.add_argument("spec", help="JSON run spec")
    ap.add_argument("--out", default=None, help="directory for view<i>.ppm images")
    ap.add_argument("--check", action="store_true", help="only check the spec and print it with its defaults")
    a = ap.parse_args(argv)
    try:
        spec = load(a.spec)
    except (SpecError, json.JSONDecodeError, OSError) as e:
        print(f"spec error: {e}", file=sys.stderr)
        return 2
    if a.check:
        print(json.dumps(spec, indent=1))
        return 0
    print(json.dumps(run(spec, a.out)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
