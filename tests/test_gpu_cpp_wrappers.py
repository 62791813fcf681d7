"""GPU: the C++ classes of include/svo/svo_hip.hpp drive the out-of-core
frame loop (analyse pass, visible sets, client update, device update,
renderToBuffer) from a program built with a plain host compiler
(tests/cpp/wrapper_gpu.cpp, built by tests/cpp/Makefile); every step must
match the same loop through the Python mirror, and the last frame must
match it pixel for pixel."""
import math
import os
import subprocess

import numpy as np
import pytest

import libsvo_amd as svo

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "bin", "wrapper_gpu")
W, H, DIVIDE = 160, 100, 4


def _camera(f):
    a = 0.4 * f
    return svo.look_at((1.6 * math.cos(a), 0.9, 1.6 * math.sin(a)), (0.0, -0.1, 0.0), (0.0, 1.0, 0.0), 55.0, W, H,
                       far=20.0, lod=0.25)


def test_cpp_frame_loop_matches_python(terrain8, tmp_path):
    if not os.path.exists(PROG):
        pytest.fail(f"{PROG} is not built: run `make -C tests/cpp` (part of __graft_entry__.build())")
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    sc, base = terrain8
    slots, frames = 48, 12
    out = str(tmp_path / "frame.bin")
    print(f"running {PROG} on {base}, {slots} slots, {frames} frames", flush=True)
    run = subprocess.run([PROG, base, str(slots), str(frames), out], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    lines = run.stdout.strip().splitlines()

    mm = svo.TreeletMemoryManagerHip(sc, slots * sc.info["treelet_bytes"], device=0)
    s = svo.Stream()
    want = []
    pages = 0
    for f in range(frames):
        cam = _camera(f)
        jx = np.float32(0.25) * np.float32(f % 3) - np.float32(0.25)
        jy = np.float32(0.5) - np.float32(0.25) * np.float32(f % 5)
        mm.performAnalysePass(cam, DIVIDE, (float(jx), float(jy)), "treelet", s)
        vs = mm.readFeedback(s)
        eo = 0.0
        for g, e in vs["old"]:
            eo += g * float(np.float32(e))
        en = 0.0
        for g, e in vs["new"]:
            en += g * float(np.float32(e))
        want.append(f"visible {f} {len(vs['old'])} {len(vs['new'])} {eo:.9g} {en:.9g}")
        st = mm.updateClientSideIncoreBuffer(s)
        up = mm.updateDeviceMemory(s)
        want.append(f"update {f} {st['inserted']} {st['evicted']} {st['resident_treelets']} {up['bytes']}")
        pages += st["inserted"]
    want.append("slots " + " ".join(str(g) for g in mm.getSlots()))
    want.append(f"frame {W} {H}")
    assert lines[:len(want)] == want
    # the device pools taken through the C++ class equal the host buffers,
    # stay equal through one more update, and initClBuffer restores them
    mm.performAnalysePass(_camera(frames), DIVIDE, (0.0, 0.0), "treelet", s)
    st = mm.updateClientSideIncoreBuffer(s)
    assert lines[len(want):] == ["pools_equal 1", f"pools_after_update {st['inserted']} 1", "pools_clobbered 0",
                                 "pools_after_init 1"]
    mm.updateDeviceMemory(s)
    print(f"{pages} page-ins over {frames} frames", flush=True)
    assert pages > 0
    raw = np.fromfile(out, np.uint32).reshape(3, H, W)
    # the frame file holds the last frame of the loop, before the extra update
    ref_mm = svo.TreeletMemoryManagerHip(sc, slots * sc.info["treelet_bytes"], device=0)
    for f in range(frames):
        cam = _camera(f)
        jx = np.float32(0.25) * np.float32(f % 3) - np.float32(0.25)
        jy = np.float32(0.5) - np.float32(0.25) * np.float32(f % 5)
        ref_mm.performAnalysePass(cam, DIVIDE, (float(jx), float(jy)), "treelet", s)
        ref_mm.updateClientSideIncoreBuffer(s)
        ref_mm.updateDeviceMemory(s)
    ref = svo.render(ref_mm, _camera(frames - 1), "treelet")
    assert np.array_equal(raw[0], ref["t"].view(np.uint32))
    assert np.array_equal(raw[1], ref["node"])
    assert np.array_equal(raw[2], ref["rgba"])
    assert np.count_nonzero(ref["node"] != svo.MISS) > 1000
