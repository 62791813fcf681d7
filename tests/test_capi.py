"""The C-ABI library loads, exports every declared symbol and reports errors
with status codes (no GPU needed)."""
import ctypes
import os
import re

import pytest

import libsvo_amd as svo
from libsvo_amd import _capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "svo_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(svo_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert len(names) > 30
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    unbound = [n for n in names if n not in _capi.SIGNATURES]
    assert not unbound, f"declared but without ctypes signature: {unbound}"


def test_ctypes_layouts_match_the_c_compiler(tmp_path):
    """Every struct of include/svo_hip.h has the C compiler's size and field
    offsets in the ctypes binding."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    structs = [n for n in dir(_capi) if n.startswith("svo_") and isinstance(getattr(_capi, n), type)
               and issubclass(getattr(_capi, n), ctypes.Structure)]
    assert len(structs) >= 12
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "svo_hip.h"', "int main(void) {"]
    for n in structs:
        lines.append(f'  printf("{n} size %zu\\n", sizeof({n}));')
        for f, _ in getattr(_capi, n)._fields_:
            lines.append(f'  printf("{n} {f} %zu\\n", offsetof({n}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = {}
    for line in filter(None, out):
        name, field, value = line.split()
        want[(name, field)] = int(value)
    for n in structs:
        cls = getattr(_capi, n)
        assert ctypes.sizeof(cls) == want[(n, "size")], n
        for f, _ in cls._fields_:
            assert getattr(cls, f).offset == want[(n, f)], f"{n}.{f}"


def test_struct_sizes_match_header_layout():
    assert ctypes.sizeof(_capi.svo_node) == 8
    assert ctypes.sizeof(_capi.svo_treelet_header) == 32
    assert ctypes.sizeof(_capi.svo_camera) == 80
    assert ctypes.sizeof(_capi.svo_frame_out) == 6 * 8


def test_errors_are_status_codes(tmp_path):
    with pytest.raises(svo.SvoError) as e:
        svo.Scene.open(str(tmp_path / "does_not_exist"))
    assert e.value.status == _capi.SVO_ERR_IO
    assert "could not open" in str(e.value)
    with pytest.raises(svo.SvoError) as e:
        svo.Scene.terrain(1, 0)
    assert e.value.status == _capi.SVO_ERR_ARG
    with pytest.raises(ValueError):
        svo.Scene.terrain(4, 0, packing="zigzag")
    c = svo.Camera()
    assert _capi.lib.svo_camera_look_at(None, None, None, 60.0, 4, 4, 20.0, 0.0, ctypes.byref(c)) == _capi.SVO_ERR_ARG


def test_host_only_handle_refuses_render():
    sc = svo.Scene.menger(1, 3, 0)
    mm = svo.TreeletMemoryManagerHip(sc, sc.info["treelet_bytes"], device=-1)
    cam = svo.look_at((0, 0, 2), (0, 0, 0), (0, 1, 0), 60, 8, 8)
    out = (_capi.svo_frame_out * 1)()
    arr = (_capi.svo_camera * 1)(cam)
    assert _capi.lib.svo_render(mm.handle, arr, 1, out, 0, None) == _capi.SVO_ERR_STATE
    assert "no device pools" in _capi.lib.svo_last_error().decode()
    # the feedback readbacks need the device tables as well
    assert _capi.lib.svo_mm_fetch_feedback(mm.handle, None) == _capi.SVO_ERR_STATE
    assert "no device pools" in _capi.lib.svo_last_error().decode()
    assert _capi.lib.svo_mm_fetch_feedback(None, None) == _capi.SVO_ERR_ARG
