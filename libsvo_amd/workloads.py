"""Synthetic scenes and cameras of the benchmark configurations.

All scenes live in the unit cube [-0.5,0.5]^3; every random draw is seeded.
Cameras follow the reference renderer's set-up: pinhole, vertical field of
view, far plane at 20 (raycast_serialized_opencl.cpp:297-304) and no
half-pixel offset.

* Menger      : level-4 Menger sponge on a depth-8 (256^3) grid, one treelet,
                512x512 rays from a seeded point on the sphere of radius 2.5,
                60 degree field of view, looking at the centre.
* Terrain     : depth-11 (2048^3) value-noise terrain in 4096-node treelets,
                1920x1080 rays from 0.1 above the surface at a seeded (x,z),
                yaw uniform in [0, 2pi), pitch uniform in [-30, 0] degrees.
* Sphere cloud: depth-10 (1024^3) grid, 2000 spheres, radius uniform in
                [0.005, 0.05], centres uniform in the unit cube.
* Batched     : 64 views per launch over the depth-11 terrain and the sphere
                cloud, resolutions uniform over {256^2, 512^2, 1280x720,
                1920x1080}.
* Out-of-core : depth-13 terrain in 16384-node treelets behind an 8 GiB
                device pool, 300 frames at 1920x1080 along a Catmull-Rom
                spline through 8 seeded control points 0.1 above the surface.
* Serving     : Poisson arrivals over Menger depth 9, terrain depth 11,
                sphere cloud depth 10 and terrain depth 12; resolutions
                256^2 40 %, 1280x720 40 %, 1920x1080 15 %, 3840x2160 5 %.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from . import Scene, look_at

MENGER_SEED = 1
MENGER_CAMERA_SEED = 123
TERRAIN_SEED = 7
TERRAIN_CAMERA_SEED = 11
SPHERES_SEED = 5
BATCH_SEED = 31
SERVING_SEED = 47
FLIGHT_SEED = 59
FAR = 20.0
FOV = 60.0


def menger_scene(depth: int = 8, level: int = 4, treelet_bytes: int = 0, packing: str = "dense") -> Scene:
    return Scene.menger(level, depth, MENGER_SEED, treelet_bytes, packing)


def sphere_camera(seed: int, width: int = 512, height: int = 512, radius: float = 2.5, lod: float = 0.0):
    """Uniform random eye on a sphere around the cube, redrawn within 1 degree
    of the +-y poles (the up vector would be degenerate)."""
    rng = np.random.default_rng(seed)
    while True:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if abs(v[1]) < math.cos(math.radians(1.0)):
            break
    eye = v * radius
    return look_at(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), FOV, width, height, FAR, lod)


def terrain_scene(depth: int = 11, treelet_bytes: int = 32768, packing: str = "dense",
                  seed: int = TERRAIN_SEED) -> Scene:
    return Scene.terrain(depth, seed, treelet_bytes, packing)


def terrain_camera(seed: int, width: int = 1920, height: int = 1080, lod: float = 0.0,
                   terrain_seed: int = TERRAIN_SEED):
    """Eye 0.1 above the terrain surface at a seeded (x,z) away from the
    walls, yaw ~ U[0,2pi), pitch ~ U[-30,0] degrees."""
    from . import _capi
    rng = np.random.default_rng(seed)
    x, z = rng.uniform(0.1, 0.9, size=2)
    h = _capi.ctypes.c_double()
    _capi.check(_capi.lib.svo_terrain_height(terrain_seed, float(x), float(z), _capi.ctypes.byref(h)),
                "svo_terrain_height")
    eye = np.array([x - 0.5, h.value - 0.5 + 0.1, z - 0.5])
    yaw = rng.uniform(0.0, 2.0 * math.pi)
    pitch = math.radians(rng.uniform(-30.0, 0.0))
    fwd = np.array([math.cos(pitch) * math.cos(yaw), math.sin(pitch), math.cos(pitch) * math.sin(yaw)])
    return look_at(eye, eye + fwd, (0.0, 1.0, 0.0), FOV, width, height, FAR, lod)


def sphere_cloud_scene(depth: int = 10, count: int = 2000, r_min: float = 0.005, r_max: float = 0.05,
                       treelet_bytes: int = 32768, packing: str = "dense", seed: int = SPHERES_SEED) -> Scene:
    return Scene.spheres(depth, count, r_min, r_max, seed, treelet_bytes, packing)


BATCH_RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
SERVING_RESOLUTIONS = [(256, 256), (1280, 720), (1920, 1080), (3840, 2160)]
SERVING_WEIGHTS = [0.40, 0.40, 0.15, 0.05]


def scene_camera(kind: str, seed: int, width: int, height: int, lod: float = 0.0):
    """Seeded camera for a scene kind: terrain cameras fly above the surface,
    the others orbit the cube."""
    if kind == "terrain":
        return terrain_camera(seed, width, height, lod=lod)
    return sphere_camera(seed, width, height, lod=lod)


def batched_views(kinds, n: int = 64, seed: int = BATCH_SEED):
    """n (scene index, camera) pairs: scene uniform over `kinds`, resolution
    uniform over BATCH_RESOLUTIONS."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = int(rng.integers(len(kinds)))
        w, h = BATCH_RESOLUTIONS[int(rng.integers(len(BATCH_RESOLUTIONS)))]
        out.append((k, scene_camera(kinds[k], int(rng.integers(1 << 30)), w, h)))
    return out


def serving_requests(kinds, rate: float, count: int, seed: int = SERVING_SEED, resolutions=None, weights=None):
    """`count` requests of a Poisson process of `rate` per second: (scene
    index, camera, arrival_s), scene uniform over `kinds`, resolution drawn
    from SERVING_RESOLUTIONS with SERVING_WEIGHTS."""
    rng = np.random.default_rng(seed)
    res = resolutions or SERVING_RESOLUTIONS
    wts = np.asarray(weights or SERVING_WEIGHTS, np.float64)
    gaps = rng.exponential(1.0 / rate, size=count)
    arrivals = np.cumsum(gaps) - gaps[0]
    scenes = rng.integers(len(kinds), size=count)
    picks = rng.choice(len(res), size=count, p=wts / wts.sum())
    cam_seeds = rng.integers(1 << 30, size=count)
    out = []
    for i in range(count):
        k = int(scenes[i])
        w, h = res[int(picks[i])]
        out.append((k, scene_camera(kinds[k], int(cam_seeds[i]), w, h), float(arrivals[i])))
    return out


def flythrough_cameras(frames: int = 300, seed: int = FLIGHT_SEED, width: int = 1920, height: int = 1080,
                       lod: float = 1.0, terrain_seed: int = TERRAIN_SEED, points: int = 8):
    """Cameras along a uniform Catmull-Rom spline through `points` seeded
    control points (x, z uniform in [0.15, 0.85], 0.1 above the surface;
    end points repeated), never below 0.05 above the surface, looking along
    the horizontal direction of travel and 15 degrees down."""
    from . import _capi
    rng = np.random.default_rng(seed)

    def ground(x, z):
        h = _capi.ctypes.c_double()
        _capi.check(_capi.lib.svo_terrain_height(terrain_seed, float(x), float(z), _capi.ctypes.byref(h)),
                    "svo_terrain_height")
        return h.value

    ctrl = []
    for _ in range(points):
        x, z = rng.uniform(0.15, 0.85, size=2)
        ctrl.append(np.array([x, ground(x, z) + 0.1, z]))
    P = [ctrl[0]] + ctrl + [ctrl[-1]]

    def spline(u):
        seg = min(int(u), points - 2)
        t = u - seg
        p0, p1, p2, p3 = P[seg], P[seg + 1], P[seg + 2], P[seg + 3]
        return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t +
                      (-p0 + 3 * p1 - 3 * p2 + p3) * t * t * t)

    cams = []
    span = float(points - 1)
    pitch = math.radians(-15.0)
    for f in range(frames):
        u = span * f / max(1, frames - 1)
        p = spline(u)
        q = spline(min(span, u + 1e-3)) if u < span else p + (p - spline(u - 1e-3))
        p[1] = max(p[1], ground(p[0], p[2]) + 0.05)
        fwd = q - p
        yaw = math.atan2(fwd[2], fwd[0])
        look = np.array([math.cos(pitch) * math.cos(yaw), math.sin(pitch), math.cos(pitch) * math.sin(yaw)])
        eye = p - 0.5
        cams.append(look_at(eye, eye + look, (0.0, 1.0, 0.0), FOV, width, height, FAR, lod))
    return cams


def incore_bytes_for(scene: Scene, spare_slots: int = 0) -> int:
    """Incore buffer holding every treelet of the scene (plus spare slots)."""
    i = scene.info
    return i["treelet_bytes"] * (i["num_treelets"] + spare_slots)


def algorithmic_bytes(pushes: np.ndarray, node: np.ndarray) -> int:
    """Bytes a ray must move at minimum: one 8-byte node per distinct parent
    (root + pushes), a 4-byte colour per hit and 16 bytes of output
    (t, node, treelet, rgba)."""
    hits = int(np.count_nonzero(node != 0xFFFFFFFF))
    return int(8 * (pushes.astype(np.int64) + 1).sum() + 4 * hits + 16 * node.size)
