"""Node masks and the .svo/.ia bundle: a hand-built byte fixture, round trips
and malformed files."""
import os
import struct

import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import _capi
from tests import helpers as H


def test_rank_closed_form_matches_reference_loop_exhaustively():
    """getNthChildIndex adds one per valid bit below the child index
    (src/CpuSvoNode.cpp:145-152); the kernel uses popcount of the masked
    valid bits.  Checked over every 16-bit mask and child index."""
    masks = np.arange(1 << 16, dtype=np.uint32)
    valid = (masks >> 8) & 0xFF
    for c in range(8):
        loop = np.zeros_like(masks)
        for i in range(c):
            loop += (masks >> (8 + i)) & 1
        closed = np.array([bin(v & ((1 << c) - 1)).count("1") for v in range(256)], np.uint32)[valid]
        assert np.array_equal(loop, closed)
    # relative indices wrap modulo 2^32 (stitched offsets may be negative)
    assert H.nth_child_loop(0xFF00, 0xFFFFFFFE, 3) == 1


def u32s(*v):
    return struct.pack("<%dI" % len(v), *v)


def node(first, valid, leaf):
    return u32s(first & 0xFFFFFFFF, (valid << 8) | leaf)


def write_fixture(base):
    """Two treelets of 128 nodes.  Treelet 0: root with children 0 and 7;
    child 0 has child 3, an anchor for treelet 1; child 7 has leaves 1, 2.
    Treelet 1: root (depth 2) with one leaf child 5 at depth 3."""
    t0 = [node(1, 0b10000001, 0), node(2, 0b00001000, 0b00001000), node(2, 0b00000110, 0b00000110),
          node(1, 0, 0), node(0, 0, 0), node(0, 0, 0)]
    t1 = [node(1, 0b00100000, 0b00100000), node(0, 0, 0)]
    with open(base + ".svo", "wb") as f:
        f.write(u32s(2, 1024))
        f.write(u32s(6, 3, 1024, 0, 0, 0, 0, 0) + b"".join(t0))
        f.write(u32s(2, 1, 1024, 1, 0, 3, 3, 1) + b"".join(t1))
    with open(base + ".ia", "wb") as f:
        f.write(u32s(2))
        f.write(u32s(6, 2) + b"".join(u32s(0x80FF8000, 0x10203000 + i) for i in range(6)))
        f.write(u32s(2, 2) + b"".join(u32s(0x80FF8000, 0x40506000 + i) for i in range(2)))


def test_hand_built_fixture_loads(tmp_path):
    base = str(tmp_path / "fixture")
    write_fixture(base)
    sc = svo.Scene.open(base)
    info = sc.info
    assert info["num_treelets"] == 2 and info["treelet_bytes"] == 1024 and info["nodes_per_treelet"] == 128
    assert info["max_depth"] == 3
    assert info["tree_nodes"] == 7 and info["tree_leaves"] == 3
    h, nodes, attr = sc.treelet(1)
    assert h == {"num_nodes": 2, "num_leaves": 1, "mem_size": 1024, "gid": 1, "parent_gid": 0,
                 "parent_leaf_pos": 3, "parent_leaf_idx": 3, "parent_leaf_parent_pos": 1}
    assert nodes.tolist() == [[1, 0x2020], [0, 0]]
    assert attr[:, 1].tolist() == [0x40506000, 0x40506001]
    # the oracle's own reader sees the same bundle
    import oracle
    o = oracle.OracleScene(base, 2 * 1024)
    assert (o.num_treelets, o.nodes_per_treelet, o.max_depth) == (2, 128, 3)
    # and stitching treelet 1 into slot 1 gives the reference's relative offset
    inc = o.incore()
    assert inc["nodes"][3].tolist() == [(128 + 1) - 3, 0x2020]
    assert inc["nodes"][1, 1] == 0x0800  # leaf bit 3 cleared


def test_written_bytes_follow_reference_layout(tmp_path):
    sc = svo.Scene.terrain(5, 3, 1024)
    base = str(tmp_path / "t5")
    sc.write(base)
    raw = open(base + ".svo", "rb").read()
    nt, tb = struct.unpack_from("<2I", raw, 0)
    info = sc.info
    assert (nt, tb) == (info["num_treelets"], 1024)
    off = 8
    for g in range(nt):
        hdr = struct.unpack_from("<8I", raw, off)
        h, nodes, _ = sc.treelet(g)
        assert hdr == tuple(h[k] for k in ("num_nodes", "num_leaves", "mem_size", "gid", "parent_gid",
                                           "parent_leaf_pos", "parent_leaf_idx", "parent_leaf_parent_pos"))
        off += 32
        body = np.frombuffer(raw, np.uint32, 2 * hdr[0], off).reshape(-1, 2)
        assert np.array_equal(body, nodes)
        off += 8 * hdr[0]
    assert off == len(raw)


def test_round_trip_is_byte_identical(tmp_path):
    sc = svo.Scene.terrain(6, 11, 2048, "bfs")
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    sc.write(a)
    svo.Scene.open(a).write(b)
    for ext in (".svo", ".ia"):
        assert open(a + ext, "rb").read() == open(b + ext, "rb").read()
    assert svo.Scene.open(a).info == sc.info


def corrupt(base, ext, mutate):
    path = base + ext
    data = bytearray(open(path, "rb").read())
    data = mutate(data)
    open(path, "wb").write(bytes(data))


@pytest.mark.parametrize("case", ["truncated_svo", "bad_offset", "gid_mismatch", "missing_ia", "bad_ia_words",
                                  "treelet_overflow", "truncated_ia"])
def test_malformed_bundles_are_rejected(tmp_path, case):
    base = str(tmp_path / "fx")
    write_fixture(base)
    want = _capi.SVO_ERR_FORMAT
    if case == "truncated_svo":
        corrupt(base, ".svo", lambda d: d[:-4])
    elif case == "bad_offset":   # root's first child points past the treelet
        corrupt(base, ".svo", lambda d: d[:40] + u32s(1000) + d[44:])
    elif case == "gid_mismatch":
        corrupt(base, ".svo", lambda d: d[:20] + u32s(5) + d[24:])
    elif case == "missing_ia":
        os.remove(base + ".ia")
        want = _capi.SVO_ERR_IO
    elif case == "bad_ia_words":
        corrupt(base, ".ia", lambda d: d[:8] + u32s(3) + d[12:])
    elif case == "treelet_overflow":  # num_nodes larger than mem_size / 8
        corrupt(base, ".svo", lambda d: d[:8] + u32s(1000) + d[12:])
    elif case == "truncated_ia":
        corrupt(base, ".ia", lambda d: d[:-1])
    with pytest.raises(svo.SvoError) as e:
        svo.Scene.open(base)
    assert e.value.status == want
