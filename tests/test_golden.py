"""Committed golden vectors (tests/golden/golden_small.npz, written by
tests/golden/make_golden.py): builder bytes and oracle outputs on the CPU,
and the HIP path against the same stored outputs on the GPU."""
import os

import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import _capi
from tests.golden import make_golden as G

GOLDEN = np.load(G.OUT, allow_pickle=False)


def camera_from(arr):
    cam = _capi.svo_camera()
    _capi.ctypes.memmove(_capi.ctypes.addressof(cam), arr.tobytes(), _capi.ctypes.sizeof(cam))
    return cam


def golden_cameras(name):
    i = 0
    while f"{name}/cam{i}/camera" in GOLDEN:
        yield i, camera_from(GOLDEN[f"{name}/cam{i}/camera"])
        i += 1


@pytest.mark.parametrize("name", list(G.SCENES))
def test_builder_bytes_and_oracle_outputs(tmp_path, name):
    import oracle
    sc, base = G.build(name, str(tmp_path))
    assert G.sha(base + ".svo") == str(GOLDEN[f"{name}/svo_sha256"])
    assert G.sha(base + ".ia") == str(GOLDEN[f"{name}/ia_sha256"])
    o = oracle.OracleScene(base, int(GOLDEN[f"{name}/incore_bytes"]))
    n = 0
    for ci, cam in golden_cameras(name):
        for pol in ("treelet", "cpu3"):
            r = o.render(cam, pol, threads=2)
            for k in ("node", "treelet", "rgba", "boundary", "pushes"):
                assert np.array_equal(r[k], GOLDEN[f"{name}/cam{ci}/{pol}/{k}"]), (name, ci, pol, k)
            assert np.array_equal(r["t"].view(np.uint32), GOLDEN[f"{name}/cam{ci}/{pol}/t"].view(np.uint32))
            n += 1
    assert n == 4


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(G.SCENES))
def test_gpu_matches_golden(tmp_path, name):
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible")
    sc, base = G.build(name, str(tmp_path))
    mm = svo.TreeletMemoryManagerHip(sc, int(GOLDEN[f"{name}/incore_bytes"]), device=0, resident_all=True)
    for ci, cam in golden_cameras(name):
        for pol in ("treelet", "cpu3"):
            print(f"golden {name} cam{ci} {pol}", flush=True)
            r = svo.render(mm, cam, pol)
            for k in ("node", "treelet", "rgba", "pushes"):
                assert np.array_equal(r[k], GOLDEN[f"{name}/cam{ci}/{pol}/{k}"]), (name, ci, pol, k)
            assert np.array_equal(r["t"].view(np.uint32), GOLDEN[f"{name}/cam{ci}/{pol}/t"].view(np.uint32))
