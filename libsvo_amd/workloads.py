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
