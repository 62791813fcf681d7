"""The C++ wrappers of include/svo/svo_hip.hpp compile with a plain C++
compiler against libsvo_hip.so and behave like the Python API (host-only
manager, no GPU needed)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def wrapper_bin(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    out = str(tmp_path_factory.mktemp("cpp") / "wrapper_check")
    lib_dir = os.path.dirname(_capi.LIB_PATH)
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "wrapper_check.cpp"), "-L", lib_dir, "-lsvo_hip",
           "-Wl,-rpath," + lib_dir, "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_wrappers_match_python_api(wrapper_bin, terrain8_bfs):
    sc, base = terrain8_bfs
    slots = 6
    run = subprocess.run([wrapper_bin, base, str(slots)], check=True, capture_output=True, text=True)
    lines = dict(l.split(" ", 1) if " " in l else (l, "") for l in run.stdout.strip().splitlines())
    assert int(lines["error_status"]) == _capi.SVO_ERR_ARG
    assert int(lines["missing_status"]) == _capi.SVO_ERR_IO
    assert int(lines["treelets"]) == sc.info["num_treelets"]
    assert int(lines["nodes_per_treelet"]) == sc.info["nodes_per_treelet"]
    # the same steps through the Python API
    mm = svo.TreeletMemoryManagerHip(sc, slots * sc.info["treelet_bytes"], device=-1)
    hdr = sc.headers()
    inserted = 0
    for g in range(1, hdr.shape[0]):
        if hdr[g, 4] == 0:
            try:
                mm.insertTreeletIntoIncoreBuffer(g, 1.0)
                inserted += 1
            except svo.SvoError:
                pass
    assert int(lines["inserted"]) == inserted == slots - 1
    assert [int(v) for v in lines["slots"].split()] == mm.getSlots()
    assert lines["reinsert"] == "0"
    removed, ok = (int(v) for v in lines["remove"].split())
    assert ok == 1 and removed == mm.getSlots()[1]
    mm.removeTreeletFromIncoreBuffer(removed)
    nodes, _ = mm.getIncoreBuffer()
    checksum = int((nodes[:, 0].astype(np.uint64) * 3 + nodes[:, 1].astype(np.uint64)).sum())
    assert lines["incore_checksum"] == f"{checksum} {nodes.shape[0]}"
    assert int(lines["feedback_status"]) == _capi.SVO_ERR_STATE
    assert int(lines["after_reset"]) == 1


def test_wrappers_residency_surface(wrapper_bin, terrain8_bfs):
    """getTreelets, getFreeIncoreSlotStack, getIncoreSlotsToUpload, getSlots
    (SlotInfo), freeIncorePosition and the borrowed incore buffers through the
    C++ classes, against the same steps through the Python API and the
    restated reference bookkeeping (oracle/residency.py)."""
    from oracle import residency as R
    sc, base = terrain8_bfs
    slots = 12
    run = subprocess.run([wrapper_bin, base, str(slots)], check=True, capture_output=True, text=True)
    lines = dict(l.split(" ", 1) if " " in l else (l, "") for l in run.stdout.strip().splitlines())
    mm = svo.TreeletMemoryManagerHip(sc, slots * sc.info["treelet_bytes"], device=-1)
    hdr, tslots = mm.getTreelets()
    assert hdr.shape == (sc.info["num_treelets"], 8) and np.array_equal(hdr, sc.headers())
    assert tslots[0] == 0 and (tslots[1:] == -1).all()
    model = R.Residency(hdr, slots)
    kids = []
    for g in range(1, hdr.shape[0]):
        if len(kids) < 2 and hdr[g, 4] == 0:
            mm.insertTreeletIntoIncoreBuffer(g, 1.0)
            assert model.insert(g, R.MAX_VISIBILITY, 1.0)
            kids.append(g)
    for g in range(1, hdr.shape[0]):
        if hdr[g, 4] in kids:
            try:
                mm.insertTreeletIntoIncoreBuffer(g, 2.0)
                ok = True
            except svo.SvoError:
                ok = False
            # a refused insert still probes for a freeable slot (and moves the cursor)
            assert model.insert(g, R.MAX_VISIBILITY, 2.0) == ok
    assert mm.getSlots() == model.slot_gid
    assert mm.getFreeIncoreSlotStack() == model.free == [int(v) for v in lines["free_stack"].split()]
    dirty = [int(v) for v in lines["to_upload"].split()]
    assert mm.getIncoreSlotsToUpload() == dirty == sorted(dirty)
    assert set(dirty) == {s for s, g in enumerate(model.slot_gid) if g >= 0}
    _, tslots = mm.getTreelets()
    pairs = [f"{g}:{s}" for g, s in enumerate(tslots) if s >= 0]
    assert lines["treelet_slots"].split() == pairs
    for _ in range(9):
        mm.applyRecords([], [], [], [])
        model.update_client_side([], [])
    infos = mm.getSlotInfos()
    assert [i[0] for i in infos] == model.slot_gid and [i[1] for i in infos] == model.vis
    assert lines["slot_infos"].split() == [f"{g}/{v}/{e:g}" for g, v, e in infos]
    before = set(model.slot_gid)
    assert model.free_position()
    freed = mm.freeIncorePosition()
    assert freed > 0 and freed in before and freed not in model.slot_gid
    assert lines["freed"] == f"1 {freed}"
    assert mm.getSlots() == model.slot_gid == [int(v) for v in lines["slots2"].split()]
    assert mm.getFreeIncoreSlotStack() == model.free == [int(v) for v in lines["free_stack2"].split()]
    assert mm.getIncoreSlotsToUpload() == [int(v) for v in lines["to_upload2"].split()]
    nodes, attr = mm.getIncoreBuffer()
    assert np.array_equal(mm.getIncoreBufferView(), nodes)
    assert np.array_equal(mm.getAttributeIncoreBuffer(), attr)
    assert lines["borrowed_equal"] == f"1 {nodes.shape[0]}"
    assert int(lines["pools_status"]) == _capi.SVO_ERR_STATE
    with pytest.raises(svo.SvoError):
        mm.getDevicePools()
