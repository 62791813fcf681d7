"""Pins the CPU oracle (oracle/svo_oracle.cpp) without a GPU.

What each oracle function is checked against:
  read_bundle / make_incore  hand-built byte fixture (test_format.py) and the
                             product manager's stitching (test_residency.py)
  sample_treelet (policy A)  brute-force first hit over every leaf box in
  travers_cpu3   (policy C)  float64 on tiny scenes (LOD off); closed-form LOD
                             stops with a huge tScaleRatio; agreement of A and
                             C except on flagged boundary rays; hit point on
                             the hit voxel's box
  pixel_ray                  finite, unit length, through the far-plane pixel
  maxLoops exits             a pure-Python restatement of both traversals on
                             far-camera rays that cycle to maxLoops
                             (test_maxloops.py)
  shade (view modes)         numpy formulas per mode (test_view_modes.py)
  sample_fill (policy B)     brute-force first hit (test_single_svo.py)
"""
import math

import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads
from tests import helpers as H


def tiny_scene(tmp_path, depth, n, seed, tb=0):
    rng = np.random.default_rng(seed)
    side = 1 << depth
    ijk = rng.integers(0, side, size=(n, 3))
    codes = np.unique(H.morton_np(ijk[:, 0], ijk[:, 1], ijk[:, 2], depth))
    attr = np.stack([np.full(codes.size, 0x80FF8000, np.uint32),
                     (np.arange(codes.size, dtype=np.uint32) << np.uint32(8)) | np.uint32(0xFF)], axis=1)
    sc = svo.Scene.from_leaves(depth, codes, attr, tb)
    base = str(tmp_path / f"tiny{depth}_{seed}")
    sc.write(base)
    return sc, base, codes


def decode(codes, depth):
    i = np.zeros(codes.shape, np.int64)
    j = np.zeros_like(i)
    k = np.zeros_like(i)
    c = codes.astype(np.int64)
    for b in range(depth):
        i |= ((c >> (3 * b + 2)) & 1) << b
        j |= ((c >> (3 * b + 1)) & 1) << b
        k |= ((c >> (3 * b)) & 1) << b
    return np.stack([i, j, k], axis=1)


def brute_first_hit(o, d, boxes_lo, size):
    """Entry t of every leaf box along the ray (float64), inf where missed."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (boxes_lo - o) * inv
        t1 = (boxes_lo + size - o) * inv
    tn = np.max(np.minimum(t0, t1), axis=1)
    tf = np.min(np.maximum(t0, t1), axis=1)
    ok = (tn <= tf) & (tf >= 0)
    return np.where(ok, tn, np.inf), tf


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
@pytest.mark.parametrize("depth,seed", [(3, 1), (4, 2), (5, 3)])
def test_first_hit_matches_brute_force(tmp_path, policy, depth, seed):
    import oracle
    sc, base, codes = tiny_scene(tmp_path, depth, 40 * depth, seed)
    o = oracle.OracleScene(base, sc.info["treelet_bytes"])
    side = 1 << depth
    size = 1.0 / side
    lo = decode(codes, depth) * size - 0.5
    rng = np.random.default_rng(100 + seed)
    compared = skipped = 0
    for _ in range(1500):
        v = rng.normal(size=3)
        org = (v / np.linalg.norm(v) * 2.0).astype(np.float32)
        tgt = rng.uniform(-0.45, 0.45, size=3)
        dd = tgt - org
        dr = (dd / np.linalg.norm(dd)).astype(np.float32)
        r = o.trace_ray(org.tolist(), dr.tolist(), 0.0, policy)
        tn, tf = brute_first_hit(org.astype(np.float64), dr.astype(np.float64), lo, size)
        order = np.argsort(tn)
        best = order[0]
        if r["boundary"]:
            skipped += 1
            continue
        if not np.isfinite(tn[best]):
            assert r["node"] == svo.MISS
            compared += 1
            continue
        second = tn[order[1]] if len(order) > 1 else np.inf
        if second - tn[best] < 1e-4 or tf[best] - tn[best] < 1e-4:
            skipped += 1
            continue
        assert r["node"] != svo.MISS, "oracle missed a voxel the brute force hits"
        c = np.array(r["center"], np.float64)
        vox = np.floor((c + 0.5) * side).astype(np.int64)
        assert np.array_equal(vox, decode(codes[best:best + 1], depth)[0]), (org, dr)
        assert abs(r["t"] - tn[best]) < 2e-5
        assert r["depth"] == depth
        compared += 1
    assert compared > 1300 and skipped < 200


def test_lod_stop_closed_forms(tmp_path):
    """A huge tScaleRatio stops at the first valid child of the root: policy A
    reports depth + 1 (treeletRenderer_rgba.cl:339), policy C the child's
    depth (CpuRaycasterSingleRay3.cpp:348)."""
    import oracle
    sc, base, codes = tiny_scene(tmp_path, 5, 300, 9)
    o = oracle.OracleScene(base, sc.info["treelet_bytes"])
    h, nodes, _ = sc.treelet(0)
    first, masks = int(nodes[0, 0]), int(nodes[0, 1])
    # almost along -z through x,y > 0: the entry child is 4x+2y+z = 7.  (An
    # exactly axis-aligned ray keeps d.x = sign(0) * 1e-4 = 0, whose 1/0
    # planes evaluate to NaN and miss, as in the reference.)
    v = np.array([0.001, 0.002, -1.0])
    d = tuple((v / np.linalg.norm(v)).astype(np.float32).tolist())
    org = (0.01, 0.01, 3.0)
    child = 7
    if not H.valid(masks, child):
        pytest.skip("root child 7 empty for this seed")
    exp = (0 + H.nth_child_loop(masks, first, child)) & 0xFFFFFFFF
    ra = o.trace_ray(org, d, 1e6, "treelet")
    rc = o.trace_ray(org, d, 1e6, "cpu3")
    assert ra["node"] == exp and ra["depth"] == 2
    assert rc["node"] == exp and rc["depth"] == 1
    assert ra["t"] == pytest.approx(2.5, abs=1e-3) and rc["t"] == pytest.approx(2.5, abs=1e-3)
    # the exactly axis-aligned ray: NaN slabs, no hit under either policy
    for pol in ("treelet", "cpu3"):
        assert o.trace_ray(org, (0.0, 0.0, -1.0), 0.0, pol)["node"] == svo.MISS


def test_policies_agree_except_boundary(menger8, terrain8):
    import oracle
    total = flagged = 0
    for sc, base in (menger8, terrain8):
        o = oracle.OracleScene(base, workloads.incore_bytes_for(sc))
        for seed in range(3):
            cam = (workloads.sphere_camera(50 + seed, 160, 120) if sc is menger8[0]
                   else workloads.terrain_camera(60 + seed, 160, 120))
            a = o.render(cam, "treelet")
            c = o.render(cam, "cpu3")
            diff = a["node"] != c["node"]
            assert not np.any(diff & (a["boundary"] == 0) & (c["boundary"] == 0))
            total += a["node"].size
            flagged += int(np.count_nonzero(diff))
    assert flagged < total * 1e-3


def test_hits_lie_on_the_hit_voxel(menger8):
    import oracle
    sc, base = menger8
    o = oracle.OracleScene(base, workloads.incore_bytes_for(sc))
    cam = workloads.sphere_camera(7, 128, 128)
    r = o.render(cam, "treelet")
    dirs = oracle.ray_dirs(cam).astype(np.float64)
    hit = (r["node"] != svo.MISS) & (r["boundary"] == 0)
    assert hit.sum() > 1000
    pos = np.array(cam.position[:3], np.float64)
    p = pos + r["t"][..., None].astype(np.float64) * dirs
    half = 0.5 / (1 << sc.info["max_depth"])
    off = np.abs(p - r["center"].astype(np.float64))[hit]
    assert np.all(off <= half + 2e-5)
    # and the point is on the box surface (the entry face)
    assert np.all(np.max(off, axis=1) >= half - 2e-5)


def test_ray_generation_geometry():
    import oracle
    cam = workloads.sphere_camera(3, 64, 48)
    d = oracle.ray_dirs(cam).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=2), 1.0, atol=1e-6)
    ll = np.array(cam.far_lower_left[:3], np.float64)
    ph = np.array(cam.pixel_h[:3], np.float64)
    pv = np.array(cam.pixel_v[:3], np.float64)
    pos = np.array(cam.position[:3], np.float64)
    for x, y in ((0, 0), (63, 47), (32, 24)):
        p = ll + ph * x + pv * (48 - y) - pos
        assert np.allclose(d[y, x], p / np.linalg.norm(p), atol=1e-6)
    # the vertical field of view spans 60 degrees between the frame's edges
    mid = ll + ph * 32  # centre column of the frame
    top, bottom = mid + pv * 48 - pos, mid - pos
    ang = math.degrees(math.acos(np.dot(top, bottom) / np.linalg.norm(top) / np.linalg.norm(bottom)))
    assert ang == pytest.approx(60.0, abs=1e-3)


def test_miss_outputs(menger8):
    import oracle
    sc, base = menger8
    o = oracle.OracleScene(base, workloads.incore_bytes_for(sc))
    cam = svo.look_at((0, 3, 0), (0, 6, 0), (0, 0, 1), 60, 8, 8)  # looking away from the cube
    r = o.render(cam)
    assert np.all(r["node"] == svo.MISS) and np.all(r["treelet"] == svo.MISS)
    assert np.all(r["rgba"] == svo.BACKGROUND_RGBA) and np.all(np.isinf(r["t"]))
