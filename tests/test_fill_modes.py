"""The single-SVO renderer's view modes (opencl/fillFrameBuffer_rgba.cl:
454-704) in the oracle, pinned on the CPU without restating the oracle's
traversal:

  * FILL_SVO_DEPTH: with LOD off every hit is a leaf of the tree's depth D,
    so every hit pixel reads D/12 and every miss 0 (shade_svoDepth);
  * FILL_LOOPS: 1 - loops/300 with an integer loop count, 1 exactly where the
    ray misses the root box (shade_iterationDepth);
  * FILL_SHADOW: the shadow ray's verdict comes from a float64 line test
    against every leaf box, the normal and colour from the file's attribute
    package of the hit (v1 attribute indices, a random permutation), the
    centre from the hit leaf's grid position (shade_diffuse_color_shadow,
    :494-530).  The reference's sample() traces a secondary ray from where
    its line enters the root box and pops a node only when the node lies
    wholly behind the origin (:310-315), so a leaf counts when its box meets
    the line and its parent's box reaches t >= 0, wherever the leaf itself
    lies;
  * FILL_REFLECTION: the same for the hit and two reflection bounces, the
    next hit being the first leaf box along the reflected line, every bounce
    shaded with attribute package 0 (the reference's case 5 never sets
    attribIndex), strengths 0.75^k (shade_diffuse_color_reflection, :536-594).
Pixels where a float64 decision lies within 5e-5 (more than the policy's
2^-16 step epsilon) of a box boundary are skipped (at most a few per cent)."""
import numpy as np
import pytest

import libsvo_amd as svo
from tests import helpers as H
from tests.test_oracle_cpu import decode
from tests.test_single_svo import _abs_nodes, _leaf_voxels_abs

F = np.float32
BG = np.array([0.2, 0.3, 0.2, 1.0])
MARGIN = 5e-5


def _pack(c):
    c = np.clip(np.round(np.asarray(c, np.float64) * 255.0), 0, 255).astype(np.uint32)
    return (c[..., 0] << 24) | (c[..., 1] << 16) | (c[..., 2] << 8) | c[..., 3]


@pytest.fixture(scope="module")
def single_scene(tmp_path_factory):
    """Depth-6 tree of 2,500 random leaves with random normals and colours,
    written as SvoBranch v1 with a random attribute permutation."""
    import oracle
    depth = 6
    rng = np.random.default_rng(61)
    side = 1 << depth
    ijk = rng.integers(8, side - 8, size=(2500, 3))
    codes = np.unique(H.morton_np(ijk[:, 0], ijk[:, 1], ijk[:, 2], depth))
    nrm = rng.normal(size=(codes.size, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    normals = _pack(np.concatenate([nrm * 0.5 + 0.5, np.zeros((codes.size, 1))], axis=1))
    colors = _pack(np.concatenate([rng.uniform(0.1, 1.0, size=(codes.size, 3)), np.ones((codes.size, 1))], axis=1))
    sc = svo.Scene.from_leaves(depth, codes, np.stack([normals, colors], axis=1), 0)
    n = sc.treelet(0)[0]["num_nodes"]
    perm = rng.permutation(n).astype(np.uint32)
    base = str(tmp_path_factory.mktemp("fill") / "branch")
    sc.write_single(base, "v1", perm)
    nodes_abs, attr = _abs_nodes(sc)
    vox = _leaf_voxels_abs(nodes_abs, depth)
    packages = np.empty((n, 2), np.uint32)
    packages[perm] = attr            # node i's package sits at perm[i]
    grid = decode(codes, depth)
    lo = grid / side - 0.5
    parent_lo = (grid // 2) * 2 / side - 0.5
    return {"ora": oracle.OracleSingle(base), "depth": depth, "vox": vox, "perm": perm, "packages": packages,
            "lo": lo, "size": 1.0 / side, "parent_lo": parent_lo, "codes": codes}


def _cams():
    out = []
    for seed in range(4):
        rng = np.random.default_rng(500 + seed)
        v = rng.normal(size=3)
        v = v / np.linalg.norm(v) * 1.6
        out.append(svo.look_at(tuple(v), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 56, 44))
    return out


def _normal(packed):
    u = np.array([(packed >> 24) & 255, (packed >> 16) & 255, (packed >> 8) & 255], np.float32)
    return ((u * F(0.003921569)) - F(0.5)) * F(2.0)


def _color(packed):
    u = np.array([(packed >> 24) & 255, (packed >> 16) & 255, (packed >> 8) & 255, packed & 255], np.float32)
    return u * F(0.003921569)


def _line_boxes(o, d, lo, size):
    """Entry and exit t of the whole line o + t d through every leaf box."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (lo - o) * inv
        t1 = (lo + size - o) * inv
    return np.max(np.minimum(t0, t1), axis=1), np.min(np.maximum(t0, t1), axis=1)


def _reachable(o, d, S):
    """Per leaf: entry t along the line, and whether the traversal reaches
    it -- its box meets the line and its parent's box extends to t >= 0 --
    with None where either test lies within MARGIN of its threshold."""
    tn, tf = _line_boxes(o, d, S["lo"], S["size"])
    _, ptf = _line_boxes(o, d, S["parent_lo"], 2.0 * S["size"])
    close = (np.abs(tf - tn) <= MARGIN) | ((tf - tn > MARGIN) & (np.abs(ptf) <= MARGIN))
    if np.any(close):
        return tn, None
    return tn, (tf - tn > MARGIN) & (ptf > MARGIN)


def _shadowed(o, light, S):
    """1 / 0 when the shadow ray reaches / reaches no leaf, None if close."""
    _, reach = _reachable(o, light, S)
    if reach is None:
        return None
    return 1.0 if reach.any() else 0.0


def _first_box(o, d, S):
    """Index of the first leaf the ray reaches, -1 if none, None if close."""
    tn, reach = _reachable(o, d, S)
    if reach is None:
        return None
    hit = np.nonzero(reach)[0]
    if hit.size == 0:
        return -1
    order = hit[np.argsort(tn[hit])]
    if order.size > 1 and tn[order[1]] - tn[order[0]] < MARGIN:
        return None
    return int(order[0])


def _centre_of_box(k, S):
    return S["lo"][k] + 0.5 * S["size"]


def _shadow_colour(n, col, c, S):
    light = np.array([-0.5, 1.0, 0.6]) / np.linalg.norm([-0.5, 1.0, 0.6])
    s = _shadowed(c + 0.01 * n.astype(np.float64), light, S)
    if s is None:
        return None
    ndl = max(0.0, float(np.dot(n.astype(np.float64), light)))
    out = col.astype(np.float64) * ndl * (1.0 - s) + col.astype(np.float64) * 0.2
    out[3] = 1.0
    return out


def _hit_centre(node, S):
    i, j, k = S["vox"][int(node)]
    return (np.array([i, j, k], np.float64) + 0.5) * S["size"] - 0.5


def _render(S, cam, mode):
    return S["ora"].render(svo.with_view_mode(cam, mode), shade=True)


def test_fill_svo_depth_and_loops_closed_forms(single_scene):
    S = single_scene
    for cam in _cams():
        r = _render(S, cam, "fill_svo_depth")
        hit = r["node"] != svo.MISS
        assert hit.sum() > 100 and (~hit).sum() > 100
        assert np.all(r["shade"][hit][:, :3] == F(S["depth"]) / F(12.0))
        assert np.all(r["shade"][~hit][:, :3] == 0.0) and np.all(r["shade"][..., 3] == 1.0)
        lp = _render(S, cam, "fill_loops")["shade"][..., 0].astype(np.float64)
        k = (1.0 - lp) * 300.0
        assert np.allclose(k, np.round(k), atol=1e-3)
        assert np.all(np.round(k)[hit] >= S["depth"])   # at least one step per level down to the leaf


def test_fill_shadow_against_line_tests(single_scene):
    S = single_scene
    compared = skipped = shadowed = 0
    for cam in _cams():
        r = _render(S, cam, "fill_shadow")
        for y, x in np.argwhere(np.ones((cam.height, cam.width), bool)):
            node = int(r["node"][y, x])
            got = r["shade"][y, x].astype(np.float64)
            if node == svo.MISS:
                assert np.array_equal(got, BG.astype(np.float32).astype(np.float64))
                continue
            if r["boundary"][y, x]:
                skipped += 1
                continue
            pk = S["packages"][S["perm"][node]]   # attribIndices[nodeIndex] -> package
            want = _shadow_colour(_normal(int(pk[0])), _color(int(pk[1])), _hit_centre(node, S), S)
            if want is None:
                skipped += 1
                continue
            shadowed += want[0] < 0.2 * _color(int(pk[1]))[0] + 1e-6
            assert np.allclose(got, want, rtol=0, atol=2e-6), (y, x, got, want)
            compared += 1
    assert compared > 1000 and skipped < compared // 20
    assert 0 < shadowed < compared   # both verdicts occur


def test_fill_reflection_against_line_tests(single_scene):
    S = single_scene
    import oracle
    n0 = _normal(int(S["packages"][0][0])).astype(np.float64)
    c0 = _color(int(S["packages"][0][1]))
    compared = skipped = bounced = 0
    for cam in _cams():
        r = _render(S, cam, "fill_reflection")
        dirs = oracle.ray_dirs(cam).astype(np.float64)
        for y, x in np.argwhere(np.ones((cam.height, cam.width), bool)):
            node = int(r["node"][y, x])
            got = r["shade"][y, x].astype(np.float64)
            if node == svo.MISS:
                assert np.array_equal(got, BG.astype(np.float32).astype(np.float64))
                continue
            if r["boundary"][y, x]:
                skipped += 1
                continue
            c = _hit_centre(node, S)
            view = dirs[y, x]
            acc = np.zeros(4)
            strength = 0.75
            ok = True
            for bounce in range(3):
                sh = _shadow_colour(n0.astype(np.float32), c0, c, S)
                if sh is None:
                    ok = False
                    break
                acc = acc * (1.0 - strength) + sh * strength
                strength *= 0.75
                if bounce == 2:
                    break
                rd = view - 2.0 * n0 * np.dot(n0, view)
                rd /= np.linalg.norm(rd)
                k = _first_box(c + 0.01 * n0, rd, S)
                if k is None:
                    ok = False
                    break
                if k < 0:
                    acc = acc * (1.0 - strength) + BG * strength
                    break
                bounced += 1
                c = _centre_of_box(k, S)
                view = rd
            if not ok:
                skipped += 1
                continue
            assert np.allclose(got, acc, rtol=0, atol=1e-5), (y, x, got, acc)
            compared += 1
    assert compared > 1000 and skipped < compared // 10 and bounced > compared
