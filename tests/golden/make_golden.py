#!/usr/bin/env python3
"""Writes tests/golden/golden_small.npz: small scenes, cameras and the oracle's
per-pixel outputs for both traversal policies, with and without the LOD stop.

The reference ships no golden vectors (it has no tests and no .svo files), so
these fixtures are produced by this project's oracle after it was pinned by
tests/test_oracle_cpu.py; they freeze the builder's bytes (sha256 of the
.svo/.ia files) and the traversal outputs against regressions.

Run from the repository root:  python3 tests/golden/make_golden.py
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import libsvo_amd as svo  # noqa: E402
import oracle  # noqa: E402
from libsvo_amd import workloads  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "golden_small.npz")

SCENES = {
    "menger_l2_d5": lambda: svo.Scene.menger(2, 5, 1, 0),
    "terrain_d6_1k_dense": lambda: svo.Scene.terrain(6, 7, 1024, "dense"),
    "terrain_d6_1k_bfs": lambda: svo.Scene.terrain(6, 7, 1024, "bfs"),
}


def cameras(name):
    if name.startswith("menger"):
        return [workloads.sphere_camera(41, 48, 40), workloads.sphere_camera(42, 33, 17, lod=1.0)]
    return [workloads.terrain_camera(43, 48, 40, terrain_seed=7), workloads.terrain_camera(44, 40, 24, lod=1.0,
                                                                                        terrain_seed=7)]


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def build(name, d):
    sc = SCENES[name]()
    base = os.path.join(d, name)
    sc.write(base)
    return sc, base


def camera_bytes(cam):
    return np.frombuffer(bytes(cam), np.uint8).copy()


def main():
    data = {}
    with tempfile.TemporaryDirectory() as d:
        for name in SCENES:
            sc, base = build(name, d)
            data[f"{name}/svo_sha256"] = np.array(sha(base + ".svo"))
            data[f"{name}/ia_sha256"] = np.array(sha(base + ".ia"))
            ib = workloads.incore_bytes_for(sc, spare_slots=1)
            data[f"{name}/incore_bytes"] = np.array(ib)
            o = oracle.OracleScene(base, ib)
            for ci, cam in enumerate(cameras(name)):
                data[f"{name}/cam{ci}/camera"] = camera_bytes(cam)
                for pol in ("treelet", "cpu3"):
                    r = o.render(cam, pol, threads=1)
                    for k in ("t", "node", "treelet", "rgba", "boundary", "pushes"):
                        data[f"{name}/cam{ci}/{pol}/{k}"] = r[k]
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
