"""Legacy single-SVO files (SvoLoader, SvoBranch v1) and the single-SVO
renderer's policy (fillFrameBuffer_rgba.cl::sample), on the CPU.

  * the library's reader against hand-built byte fixtures (written here from
    the file layout, not by the library) and through write/read round trips;
  * the oracle's policy-B traversal (its own reader, absolute indexing)
    against a float64 brute-force first hit over every leaf box, LOD off;
  * policy B against policy A on the same tree: equal away from boundary
    rays, up to the rays between the two policies' epsilons.
"""
import struct

import numpy as np
import pytest

import libsvo_amd as svo
from tests import helpers as H
from tests.test_oracle_cpu import brute_first_hit, decode, tiny_scene


def _abs_nodes(sc):
    """The scene's one treelet with absolute first-child indices."""
    h, nodes, attr = sc.treelet(0)
    out = nodes.copy()
    for i in range(h["num_nodes"]):
        if (out[i, 1] >> 8) & 0xFF:
            out[i, 0] = (int(out[i, 0]) + i) & 0xFFFFFFFF
    return out, attr


def _write_fixture(base, header, nodes_abs, packages, aidx=None):
    with open(base + ".svo", "wb") as f:
        f.write(header.encode())
        f.write(nodes_abs.astype(np.uint32).tobytes())
        if aidx is not None:
            f.write(np.asarray(aidx, np.uint32).tobytes())
    with open(base + ".ia", "wb") as f:
        f.write(struct.pack("<III", 1, packages.shape[0], 2))
        f.write(packages.astype(np.uint32).tobytes())


def test_hand_built_v1_and_svoloader_fixtures(tmp_path):
    # root with two leaf children (child 0 and child 7) at depth 1
    nodes_abs = np.array([[1, (0x81 << 8) | 0x81], [0, 0], [0, 0]], np.uint32)
    packages = np.array([[0x11, 0xAA0000FF], [0x22, 0x00BB00FF], [0x33, 0x0000CCFF]], np.uint32)
    base = str(tmp_path / "fx_v1")
    _write_fixture(base, "v1 1 3 ", nodes_abs, packages, aidx=[2, 0, 1])
    sc = svo.Scene.open_single(base)
    h, nodes, attr = sc.treelet(0)
    assert sc.info["num_treelets"] == 1 and sc.info["max_depth"] == 1 and h["num_nodes"] == 3
    assert nodes[0].tolist() == [1, (0x81 << 8) | 0x81]  # absolute 1 at index 0 -> relative 1
    assert attr.tolist() == [packages[2].tolist(), packages[0].tolist(), packages[1].tolist()]
    base2 = str(tmp_path / "fx_loader")
    _write_fixture(base2, "1 3 ", nodes_abs, packages)
    sc2 = svo.Scene.open_single(base2)
    assert np.array_equal(sc2.treelet(0)[2], packages)
    # deeper node: relative offset = absolute - own index
    nodes_abs = np.array([[1, (0x01 << 8)], [2, (0x01 << 8) | 0x01], [0, 0]], np.uint32)
    base3 = str(tmp_path / "fx_chain")
    _write_fixture(base3, "v1 2 3 ", nodes_abs, packages, aidx=[0, 1, 2])
    n3 = svo.Scene.open_single(base3).treelet(0)[1]
    assert n3[:, 0].tolist() == [1, 1, 0]


@pytest.mark.parametrize("bad", ["depth", "count", "child", "attr", "leafbits", "truncated"])
def test_malformed_single_files_rejected(tmp_path, bad):
    nodes_abs = np.array([[1, (0x81 << 8) | 0x81], [0, 0], [0, 0]], np.uint32)
    packages = np.zeros((3, 2), np.uint32)
    header, aidx = "v1 1 3 ", [0, 1, 2]
    if bad == "depth":
        header = "v1 x 3 "
    elif bad == "count":
        header = "v1 1 "
    elif bad == "child":
        nodes_abs[0, 0] = 2  # children 2..3 of 3 nodes
    elif bad == "attr":
        aidx = [0, 1, 9]
    elif bad == "leafbits":
        nodes_abs[0, 1] = (0x01 << 8) | 0x03
    base = str(tmp_path / f"bad_{bad}")
    _write_fixture(base, header, nodes_abs, packages, aidx=aidx)
    if bad == "truncated":
        with open(base + ".svo", "r+b") as f:
            f.truncate(12)
    with pytest.raises(svo.SvoError):
        svo.Scene.open_single(base)


@pytest.mark.parametrize("fmt", ["v1", "svoloader"])
def test_round_trip(tmp_path, menger8, fmt):
    sc, _ = menger8
    rng = np.random.default_rng(5)
    n = sc.treelet(0)[0]["num_nodes"]
    perm = rng.permutation(n).astype(np.uint32) if fmt == "v1" else None
    base = str(tmp_path / f"menger_{fmt}")
    sc.write_single(base, fmt, perm)
    back = svo.Scene.open_single(base)
    h0, n0, a0 = sc.treelet(0)
    h1, n1, a1 = back.treelet(0)
    assert np.array_equal(n0, n1) and np.array_equal(a0, a1)
    assert back.info["max_depth"] == sc.info["max_depth"] and back.info["tree_leaves"] == sc.info["tree_leaves"]
    with open(base + ".svo", "rb") as f:
        head = f.read(24)
    assert head.startswith(b"v1 8 " if fmt == "v1" else b"8 ")
    # the nodes on disk carry absolute indices
    nabs, _ = _abs_nodes(sc)
    raw = np.fromfile(base + ".svo", np.uint8)
    off = len(("v1 8 %d " if fmt == "v1" else "8 %d ") % n)
    assert np.array_equal(raw[off:off + 8 * n].view(np.uint32).reshape(-1, 2), nabs)


def test_multi_treelet_scene_is_not_a_single_svo(tmp_path, terrain8):
    sc, _ = terrain8
    with pytest.raises(svo.SvoError):
        sc.write_single(str(tmp_path / "no"), "v1")


def _leaf_voxels_abs(nodes_abs, depth):
    """node index -> (i, j, k) of every leaf at `depth`, walking absolute
    first-child indices from the root."""
    out = {}
    stack = [(0, 0, 0, 0, 0)]
    while stack:
        n, d, i, j, k = stack.pop()
        first, masks = int(nodes_abs[n, 0]), int(nodes_abs[n, 1])
        rank = 0
        for c in range(8):
            if not (masks >> (8 + c)) & 1:
                continue
            ci, cj, ck = (i << 1) | ((c >> 2) & 1), (j << 1) | ((c >> 1) & 1), (k << 1) | (c & 1)
            child = first + rank
            rank += 1
            if (masks >> c) & 1:
                if d + 1 == depth:
                    out[child] = (ci, cj, ck)
            else:
                stack.append((child, d + 1, ci, cj, ck))
    return out


@pytest.mark.parametrize("depth,seed", [(3, 1), (4, 2), (5, 3)])
def test_policy_b_first_hit_matches_brute_force(tmp_path, depth, seed):
    import oracle
    sc, _, codes = tiny_scene(tmp_path, depth, 40 * depth, seed)
    base = str(tmp_path / f"single{depth}")
    sc.write_single(base, "v1", np.arange(sc.treelet(0)[0]["num_nodes"], dtype=np.uint32)[::-1].copy())
    o = oracle.OracleSingle(base)
    nodes_abs, _ = _abs_nodes(sc)
    vox = _leaf_voxels_abs(nodes_abs, depth)
    side = 1 << depth
    size = 1.0 / side
    lo = decode(codes, depth) * size - 0.5
    compared = skipped = 0
    for cseed in range(4):
        cam = svo.look_at(*_orbit(cseed), fov_y_deg=50, width=48, height=40)
        r = o.render(cam)
        dirs = oracle.ray_dirs(cam)
        org = np.array(cam.position[:3], np.float64)
        for y in range(cam.height):
            for x in range(cam.width):
                if r["boundary"][y, x]:
                    skipped += 1
                    continue
                tn, tf = brute_first_hit(org, dirs[y, x].astype(np.float64), lo, size)
                order = np.argsort(tn)
                best = order[0]
                node = int(r["node"][y, x])
                if not np.isfinite(tn[best]):
                    assert node == svo.MISS
                    compared += 1
                    continue
                second = tn[order[1]] if len(order) > 1 else np.inf
                if second - tn[best] < 1e-4 or tf[best] - tn[best] < 1e-4:
                    skipped += 1
                    continue
                assert node != svo.MISS, (cseed, x, y)
                assert vox[node] == tuple(decode(codes[best:best + 1], depth)[0]), (cseed, x, y)
                assert abs(r["t"][y, x] - tn[best]) < 2e-5
                compared += 1
    assert compared > 6000 and skipped < compared // 10


def _orbit(seed):
    rng = np.random.default_rng(1000 + seed)
    v = rng.normal(size=3)
    v = v / np.linalg.norm(v) * 2.2
    return tuple(v), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)


def test_policy_b_colours_follow_the_attribute_indices(tmp_path, menger8):
    import oracle
    sc, _ = menger8
    n = sc.treelet(0)[0]["num_nodes"]
    base = str(tmp_path / "menger_perm")
    sc.write_single(base, "v1", np.random.default_rng(2).permutation(n).astype(np.uint32))
    r = oracle.OracleSingle(base).render(svo.look_at((1.5, 1.2, 1.8), (0, 0, 0), (0, 1, 0), 55, 64, 64))
    _, attr = _abs_nodes(sc)
    hit = r["node"] != svo.MISS
    assert hit.sum() > 500
    assert np.array_equal(r["rgba"][hit], attr[r["node"][hit], 1] | 0xFF)
    assert np.all(r["treelet"][hit] == 0) and np.all(r["rgba"][~hit] == svo.BACKGROUND_RGBA)


def test_policy_b_agrees_with_policy_a_except_boundary(tmp_path, menger8):
    import oracle
    sc, base_a = menger8
    base = str(tmp_path / "menger_b")
    sc.write_single(base, "svoloader")
    b = oracle.OracleSingle(base)
    a = oracle.OracleScene(base_a, sc.info["treelet_bytes"])
    cam = svo.look_at((1.7, 0.9, -1.4), (0, 0, 0), (0, 1, 0), 60, 160, 120)
    rb, ra = b.render(cam), a.render(cam, "treelet")
    diff = rb["node"] != ra["node"]
    flagged = (rb["boundary"] | ra["boundary"]).astype(bool)
    assert (rb["node"] != svo.MISS).sum() > 3000
    # the two epsilons differ (1e-5 and 2^-16 = 1.53e-5): a ray whose interval
    # or entry offset falls between them takes different steps without being
    # within 1e-6 of either policy's own threshold -- rare, not zero
    assert int((diff & ~flagged).sum()) <= cam.width * cam.height // 1000


def _inner_chain(depth):
    """Absolute nodes of a tree whose root has a leaf child 0 (depth 1) and an
    inner child 7 that starts a chain of inner nodes, each the child 0 of the
    one before, down to a childless inner node at `depth` (below every
    leaf): the octants around the centre of the cube on the positive side."""
    nodes = [[1, (0x81 << 8) | 0x01], [0, 0]]  # root: children 0 (leaf) and 7
    for d in range(1, depth):  # inner node of depth d at index d+1, its child 0 next
        nodes.append([d + 2, 0x01 << 8])
    nodes.append([0, 0])  # the childless inner node at `depth`
    return np.array(nodes, np.uint32)


def test_inner_chain_below_every_leaf_counts_for_the_depth(tmp_path):
    """The depth of a single-SVO tree counts every valid node: a chain of
    inner nodes below the deepest leaf makes the tree deeper than its leaves,
    so a header that names only the leaf depth is refused, and the policy-B
    stack bound (depth 15) covers every node a traversal can push."""
    nodes = _inner_chain(2)
    packages = np.zeros((nodes.shape[0], 2), np.uint32)
    base = str(tmp_path / "chain_leafdepth")
    _write_fixture(base, "v1 1 %d " % nodes.shape[0], nodes, packages, aidx=list(range(nodes.shape[0])))
    with pytest.raises(svo.SvoError) as e:
        svo.Scene.open_single(base)
    assert e.value.status == svo._capi.SVO_ERR_FORMAT
    base = str(tmp_path / "chain_fulldepth")
    _write_fixture(base, "v1 2 %d " % nodes.shape[0], nodes, packages, aidx=list(range(nodes.shape[0])))
    sc = svo.Scene.open_single(base)
    assert sc.info["max_depth"] == 2 and sc.info["tree_leaves"] == 1
    # a chain to depth 16 loads with depth 16, which the single-SVO policy
    # (15 levels) refuses when a render is staged
    nodes = _inner_chain(16)
    packages = np.zeros((nodes.shape[0], 2), np.uint32)
    base = str(tmp_path / "chain16")
    _write_fixture(base, "v1 16 %d " % nodes.shape[0], nodes, packages, aidx=list(range(nodes.shape[0])))
    assert svo.Scene.open_single(base).info["max_depth"] == 16
