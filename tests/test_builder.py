"""Scene builder: treelet format invariants, leaf-set preservation for both
packings, inner attributes, and the procedural generators against
independent numpy restatements."""
import numpy as np
import pytest

import libsvo_amd as svo
from tests import helpers as H


def random_leaves(depth, n, seed):
    rng = np.random.default_rng(seed)
    side = 1 << depth
    ijk = rng.integers(0, side, size=(n, 3))
    codes = np.unique(H.morton_np(ijk[:, 0], ijk[:, 1], ijk[:, 2], depth))
    attr = np.stack([np.full(codes.size, 0x80FF8000, np.uint32),
                     (rng.integers(0, 2**24, size=codes.size, dtype=np.uint32) << np.uint32(8)) | np.uint32(0xFF)],
                    axis=1)
    return codes, attr


def check_invariants(scene, packing):
    info = scene.info
    N = info["nodes_per_treelet"]
    tls = H.treelets(scene)
    roots_seen = {0}
    for gid, (h, nodes, attr) in enumerate(tls):
        n = h["num_nodes"]
        assert h["gid"] == gid and h["mem_size"] == info["treelet_bytes"]
        assert 1 <= n <= N
        leaf_bits = 0
        expected_next = 1
        for i in range(n):
            first, masks = int(nodes[i, 0]), int(nodes[i, 1])
            assert masks >> 16 == 0
            vm, lm = (masks >> 8) & 255, masks & 255
            assert lm & ~vm == 0, "leaf bit without valid bit"
            leaf_bits += bin(lm).count("1")
            if vm:
                # children contiguous, in child-index order, breadth-first
                ch = [(i + H.nth_child_loop(masks, first, c)) & 0xFFFFFFFF for c in range(8) if vm >> c & 1]
                assert ch == list(range(ch[0], ch[0] + len(ch)))
                assert ch[0] == expected_next, "treelet not stored breadth-first"
                expected_next += len(ch)
            for c in range(8):
                if lm >> c & 1:
                    child = (i + H.nth_child_loop(masks, first, c)) & 0xFFFFFFFF
                    sub = int(nodes[child, 0])
                    if sub:
                        sh = tls[sub][0]
                        assert sh["parent_gid"] == gid and sh["parent_leaf_pos"] == child
                        assert sh["parent_leaf_idx"] == c and sh["parent_leaf_parent_pos"] == i
                        assert i != 0, "sub-treelet anchor directly under a treelet root"
                        assert int(nodes[child, 1]) == 0
                        # the sub-treelet root carries the anchor's attribute
                        assert np.array_equal(tls[sub][2][0], attr[child])
                        roots_seen.add(sub)
        assert expected_next == n
        assert leaf_bits == h["num_leaves"]
        if packing == "bfs":
            # reference rule: leaves are the last num_leaves nodes; a treelet
            # that anchors sub-treelets stopped for lack of room
            first_leaf = n - h["num_leaves"]
            has_sub = any(int(nodes[j, 0]) for j in range(first_leaf, n))
            if has_sub:
                assert n >= N - 8
    assert roots_seen == set(range(len(tls))), "unreachable treelets"
    # FIFO gid order: parents precede children, siblings by anchor position
    parents = [(tls[g][0]["parent_gid"], tls[g][0]["parent_leaf_pos"]) for g in range(1, len(tls))]
    assert parents == sorted(parents)


@pytest.mark.parametrize("packing", ["dense", "bfs"])
@pytest.mark.parametrize("tb", [1024, 4096])
def test_random_scene_round_trips_leaves(packing, tb):
    depth = 6
    codes, attr = random_leaves(depth, 3000, seed=tb)
    sc = svo.Scene.from_leaves(depth, codes, attr, tb, packing)
    info = sc.info
    assert info["tree_leaves"] == codes.size and info["max_depth"] == depth
    check_invariants(sc, packing)
    got = H.walk_leaves(H.treelets(sc), depth)
    assert sorted(got) == codes.tolist()
    colour = dict(zip(codes.tolist(), attr[:, 1].tolist()))
    assert all(got[c] == colour[c] for c in got)


def test_dense_needs_fewer_treelets_than_bfs():
    sc_d = svo.Scene.terrain(8, 3, 4096, "dense")
    sc_b = svo.Scene.terrain(8, 3, 4096, "bfs")
    assert sc_d.info["tree_nodes"] == sc_b.info["tree_nodes"]
    assert sc_d.info["num_treelets"] * 10 < sc_b.info["num_treelets"]
    check_invariants(sc_d, "dense")


def test_single_treelet_when_treelet_bytes_zero():
    sc = svo.Scene.menger(2, 5, 9, 0)
    info = sc.info
    assert info["num_treelets"] == 1
    assert info["total_nodes"] == info["tree_nodes"]
    check_invariants(sc, "dense")


def unpack(c):
    return np.array([(c >> 24) & 255, (c >> 16) & 255, (c >> 8) & 255], np.float32) / np.float32(255)


def test_inner_colour_is_mean_of_children():
    depth = 4
    codes, attr = random_leaves(depth, 300, seed=5)
    sc = svo.Scene.from_leaves(depth, codes, attr, 0)
    h, nodes, at = sc.treelet(0)
    checked = 0
    for i in range(h["num_nodes"]):
        first, masks = int(nodes[i, 0]), int(nodes[i, 1])
        vm = (masks >> 8) & 255
        if not vm:
            continue
        kids = [i + H.nth_child_loop(masks, first, c) for c in range(8) if vm >> c & 1]
        acc = np.zeros(3, np.float32)
        for k in kids:
            acc += unpack(int(at[k, 1]))
        mean = np.clip(acc / np.float32(len(kids)), 0, 1)
        exp = (np.floor(mean * np.float32(255) + np.float32(0.5))).astype(np.uint32)
        got = np.array([(int(at[i, 1]) >> s) & 255 for s in (24, 16, 8)])
        assert np.array_equal(got, exp), (i, got, exp)
        assert int(at[i, 1]) & 255 == 255
        checked += 1
    assert checked > 50


# ------------------------------------------------------------ generators

def shell_of(filled):
    pad = np.pad(filled, 1, constant_values=False)
    inner = pad[1:-1, 1:-1, 1:-1]
    all_n = (pad[:-2, 1:-1, 1:-1] & pad[2:, 1:-1, 1:-1] & pad[1:-1, :-2, 1:-1] & pad[1:-1, 2:, 1:-1]
             & pad[1:-1, 1:-1, :-2] & pad[1:-1, 1:-1, 2:])
    return inner & ~all_n


def leaves_from_mask(mask, depth):
    i, j, k = np.nonzero(mask)
    return sorted(H.morton_np(i, j, k, depth).tolist())


def test_menger_matches_float_restatement():
    depth, level = 5, 2
    n = 1 << depth
    c = (np.arange(n) + 0.5) / n  # centres in [0,1]
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    inside = np.ones_like(X, dtype=bool)
    for k in range(1, level + 1):
        ones = sum(((np.floor(A * 3 ** k) % 3) == 1).astype(int) for A in (X, Y, Z))
        inside &= ones < 2
    sc = svo.Scene.menger(level, depth, 1, 0)
    got = H.walk_leaves(H.treelets(sc), depth)
    assert sorted(got) == leaves_from_mask(shell_of(inside), depth)
    # solid count of the sponge: 20^level of (3^-level)^3 cells at the grid
    assert inside.sum() > 0


def test_terrain_matches_restatement():
    depth, seed = 5, 7
    n = 1 << depth
    Hh = np.zeros((n, n), np.int64)
    for i in range(n):
        for k in range(n):
            h = svo._capi.ctypes.c_double()
            svo._capi.lib.svo_terrain_height(seed, (i + 0.5) / n, (k + 0.5) / n, svo._capi.ctypes.byref(h))
            assert 0.2 <= h.value < 0.8
            Hh[i, k] = min(n, max(0, int(np.ceil(h.value * n - 0.5))))
    j = np.arange(n)
    filled = j[None, :, None] < Hh[:, None, :]
    sc = svo.Scene.terrain(depth, seed, 1024)
    got = H.walk_leaves(H.treelets(sc), depth)
    assert sorted(got) == leaves_from_mask(shell_of(filled), depth)


def test_spheres_match_restatement():
    depth, count, seed = 6, 12, 4
    sc = svo.Scene.spheres(depth, count, 0.05, 0.2, seed, 2048)
    got = H.walk_leaves(H.treelets(sc), depth)
    # rebuild the sphere list exactly as the generator draws it
    def u(s, f):
        z = seed
        def mix(z):
            z = (z + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
            return z ^ (z >> 31)
        return (mix(mix(z) ^ (4 * s + f)) >> 11) * (1.0 / 9007199254740992.0)
    n = 1 << depth
    c = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    filled = np.zeros_like(X, dtype=bool)
    for s in range(count):
        cx, cy, cz = u(s, 0), u(s, 1), u(s, 2)
        r = 0.05 + (0.2 - 0.05) * u(s, 3)
        filled |= (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 < r * r
    assert sorted(got) == leaves_from_mask(shell_of(filled), depth)


def test_tiny_tree_in_one_treelet(tmp_path):
    # a tree that fits one treelet needs no cut, however small the treelet
    code = H.morton_np(np.array([5]), np.array([2]), np.array([6]), 3)
    sc = svo.Scene.from_leaves(3, code, np.array([[0x80FF8000, 0x11223344]], np.uint32), 0, "dense")
    assert sc.info["num_treelets"] == 1 and sc.info["tree_nodes"] == 4 and sc.info["tree_leaves"] == 1
    assert sc.info["treelet_bytes"] == 128  # the smallest treelet a .svo file holds
    check_invariants(sc, "dense")
    assert H.walk_leaves(H.treelets(sc), 3) == {int(code[0]): 0x11223344}
    sc.write(str(tmp_path / "tiny"))
    back = svo.Scene.open(str(tmp_path / "tiny"))
    assert back.info == sc.info


# Hashes of the depth-9 terrain (seed 3, 4096-byte treelets) as written by the
# single-threaded builder; the generator and the radix sort split their work
# over threads from about 65k leaves on, and the bytes must not change.
TERRAIN_D9_SVO = "ba8c00bcb445ea4b8bf86aef309485cf51ac97391b996dd96a1af543ecb1a0a6"
TERRAIN_D9_IA = "da105f4afa083bed3d2be66a503f35d263d4975d0f04540af42e447834e44cca"


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_terrain_bytes_do_not_depend_on_thread_count(tmp_path, threads):
    import hashlib
    import subprocess
    import sys
    import os
    base = str(tmp_path / "t")
    code = ("import libsvo_amd as svo, sys; svo.Scene.terrain(9, 3, 4096).write(sys.argv[1])")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([sys.executable, "-c", code, base], check=True, env=env, cwd=root, timeout=120)
    assert hashlib.sha256(open(base + ".svo", "rb").read()).hexdigest() == TERRAIN_D9_SVO
    assert hashlib.sha256(open(base + ".ia", "rb").read()).hexdigest() == TERRAIN_D9_IA
