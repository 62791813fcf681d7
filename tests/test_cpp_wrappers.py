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
    lines = dict(l.split(" ", 1) for l in run.stdout.strip().splitlines())
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
