"""CPU oracle of the serialized-SVO ray caster (test infrastructure).

Loads ``oracle/libsvo_oracle.so`` (built from ``oracle/svo_oracle.cpp`` by
``make -C oracle``).  Only the tests, ``__graft_entry__.smoke()`` (to check
the GPU result) and ``bench.py``'s CPU-baseline leg use this package; the
product library never does.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SVO_ORACLE_LIBRARY") or os.path.join(_HERE, "libsvo_oracle.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run make -C oracle")
_lib = ctypes.CDLL(LIB_PATH)
vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
_lib.oracle_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
_lib.oracle_free.argtypes = [vp]
_lib.oracle_free.restype = None
_lib.oracle_info.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]
_lib.oracle_make_incore.argtypes = [vp, u64]
_lib.oracle_incore.argtypes = [vp] + [ctypes.POINTER(vp)] * 4 + [ctypes.POINTER(u64), ctypes.POINTER(u32)]
_lib.oracle_render_buffers.argtypes = [vp, u64, vp, vp, u32, u32, u32, vp, ctypes.c_int] + [vp] * 14 + [ctypes.c_int]
_lib.oracle_render.argtypes = [vp, vp, ctypes.c_int] + [vp] * 10 + [ctypes.c_int]
_lib.oracle_trace_ray.argtypes = [vp, vp, vp, ctypes.c_float, ctypes.c_int, vp, vp, vp, vp, vp]
_lib.oracle_max_threads.restype = ctypes.c_int
_lib.oracle_ray_dirs.argtypes = [vp, vp]

_lib.oracle_open_single.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
_lib.oracle_free_single.argtypes = [vp]
_lib.oracle_free_single.restype = None
_lib.oracle_render_single.argtypes = [vp, vp] + [vp] * 7 + [ctypes.c_int]

POLICY = {"treelet": 0, "single": 1, "cpu3": 2}


def max_threads() -> int:
    return _lib.oracle_max_threads()


def data_point_color(value: float, vmin: float, vmax: float) -> np.ndarray:
    """The iteration view mode's colour map at one value (FP32 RGB)."""
    out = np.empty(3, np.float32)
    _lib.oracle_data_point_color(ctypes.c_float(value), ctypes.c_float(vmin), ctypes.c_float(vmax),
                                 out.ctypes.data_as(ctypes.c_void_p))
    return out


def ray_dirs(cam) -> np.ndarray:
    out = np.empty((cam.height, cam.width, 3), np.float32)
    if _lib.oracle_ray_dirs(ctypes.byref(cam), out.ctypes.data) != 0:
        raise RuntimeError("oracle_ray_dirs failed")
    return out


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class OracleScene:
    """A treelet bundle read from <base>.svo/.ia by the oracle's own reader."""

    def __init__(self, base_path: str, incore_bytes: Optional[int] = None):
        self._h = vp()
        r = _lib.oracle_open(base_path.encode(), ctypes.byref(self._h))
        if r != 0:
            raise RuntimeError(f"oracle_open({base_path}) failed with {r}")
        nt, n, d = u32(), u32(), u32()
        _lib.oracle_info(self._h, ctypes.byref(nt), ctypes.byref(n), ctypes.byref(d))
        self.num_treelets, self.nodes_per_treelet, self.max_depth = nt.value, n.value, d.value
        if incore_bytes is not None:
            self.make_incore(incore_bytes)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.oracle_free(self._h)
            self._h = None

    def make_incore(self, incore_bytes: int) -> None:
        r = _lib.oracle_make_incore(self._h, int(incore_bytes))
        if r != 0:
            raise RuntimeError(f"oracle_make_incore failed with {r}")

    def incore(self) -> Dict[str, np.ndarray]:
        pn, pc, pm, ps = vp(), vp(), vp(), vp()
        pool, slots = u64(), u32()
        r = _lib.oracle_incore(self._h, ctypes.byref(pn), ctypes.byref(pc), ctypes.byref(pm), ctypes.byref(ps),
                               ctypes.byref(pool), ctypes.byref(slots))
        if r != 0:
            raise RuntimeError("oracle incore buffer not built")
        n, s = pool.value, slots.value
        u32p = ctypes.POINTER(ctypes.c_uint32)
        return {
            "nodes": np.ctypeslib.as_array(ctypes.cast(pn, u32p), (n, 2)).copy(),
            "colors": np.ctypeslib.as_array(ctypes.cast(pc, u32p), (n,)).copy(),
            "normals": np.ctypeslib.as_array(ctypes.cast(pm, u32p), (n,)).copy(),
            "slot_gid": np.ctypeslib.as_array(ctypes.cast(ps, ctypes.POINTER(ctypes.c_int32)), (s,)).copy(),
        }

    def render(self, cam, policy="treelet", threads: int = 0, extras: bool = True,
               shade: bool = False) -> Dict[str, np.ndarray]:
        W, H = cam.width, cam.height
        out = {"t": np.empty((H, W), np.float32), "node": np.empty((H, W), np.uint32),
               "treelet": np.empty((H, W), np.uint32), "rgba": np.empty((H, W), np.uint32)}
        if extras:
            out.update({"boundary": np.empty((H, W), np.uint8), "pushes": np.empty((H, W), np.uint32),
                        "depth": np.empty((H, W), np.uint32), "loops": np.empty((H, W), np.uint32),
                        "center": np.empty((H, W, 3), np.float32)})
        if shade:
            out["shade"] = np.empty((H, W, 4), np.float32)
        g = out.get
        r = _lib.oracle_render(self._h, ctypes.byref(cam), POLICY.get(policy, policy), _ptr(g("t")), _ptr(g("node")),
                               _ptr(g("treelet")), _ptr(g("rgba")), _ptr(g("boundary")), _ptr(g("pushes")),
                               _ptr(g("depth")), _ptr(g("loops")), _ptr(g("center")), _ptr(g("shade")),
                               threads if threads > 0 else max_threads())
        if r != 0:
            raise RuntimeError(f"oracle_render failed with {r}")
        return out

    def trace_ray(self, origin, direction, ratio: float = 0.0, policy="treelet"):
        o = (ctypes.c_float * 3)(*origin)
        d = (ctypes.c_float * 3)(*direction)
        node, t, depth = u32(), ctypes.c_float(), u32()
        c = (ctypes.c_float * 3)()
        b = ctypes.c_uint8()
        r = _lib.oracle_trace_ray(self._h, o, d, ctypes.c_float(ratio), POLICY.get(policy, policy),
                                  ctypes.byref(node), ctypes.byref(t), ctypes.byref(depth), c, ctypes.byref(b))
        if r != 0:
            raise RuntimeError("oracle_trace_ray failed")
        return {"node": node.value, "t": t.value, "depth": depth.value, "center": tuple(c), "boundary": bool(b.value)}


class OracleSingle:
    """A legacy single-SVO file (SvoLoader or SvoBranch v1) read by the
    oracle's own parser and traversed with absolute indexing (policy B,
    fillFrameBuffer_rgba.cl::sample)."""

    def __init__(self, base_path: str):
        self._h = vp()
        r = _lib.oracle_open_single(base_path.encode(), ctypes.byref(self._h))
        if r != 0:
            raise RuntimeError(f"oracle_open_single({base_path}) failed with {r}")

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.oracle_free_single(self._h)
            self._h = None

    def render(self, cam, threads: int = 0, shade: bool = False) -> Dict[str, np.ndarray]:
        W, H = cam.width, cam.height
        out = {"t": np.empty((H, W), np.float32), "node": np.empty((H, W), np.uint32),
               "treelet": np.empty((H, W), np.uint32), "rgba": np.empty((H, W), np.uint32),
               "boundary": np.empty((H, W), np.uint8), "pushes": np.empty((H, W), np.uint32)}
        if shade:
            out["shade"] = np.empty((H, W, 4), np.float32)
        r = _lib.oracle_render_single(self._h, ctypes.byref(cam), *[_ptr(out.get(k)) for k in
                                                                    ("t", "node", "treelet", "rgba", "boundary",
                                                                     "pushes", "shade")],
                                      threads if threads > 0 else max_threads())
        if r != 0:
            raise RuntimeError("oracle_render_single failed")
        return out


def render_buffers(nodes: np.ndarray, colors: np.ndarray, slot_gid: np.ndarray, nodes_per_treelet: int, depth: int,
                   cam, policy="treelet", threads: int = 0, records: bool = False,
                   normals: Optional[np.ndarray] = None) -> Dict[str, np.ndarray]:
    """Oracle traversal over caller-provided incore arrays.  records=True adds
    the per-ray analysis record (rec_kind, rec_err, rec_sub); normals (per
    node) adds the view-mode image "shade" of the camera's view mode."""
    nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1, 2)
    colors = np.ascontiguousarray(colors, np.uint32)
    slot_gid = np.ascontiguousarray(slot_gid, np.int32)
    W, H = cam.width, cam.height
    out = {"t": np.empty((H, W), np.float32), "node": np.empty((H, W), np.uint32),
           "treelet": np.empty((H, W), np.uint32), "rgba": np.empty((H, W), np.uint32),
           "boundary": np.empty((H, W), np.uint8), "pushes": np.empty((H, W), np.uint32)}
    if records:
        out.update({"rec_kind": np.empty((H, W), np.uint8), "rec_err": np.empty((H, W), np.float32),
                    "rec_sub": np.empty((H, W), np.uint32)})
    if normals is not None:
        normals = np.ascontiguousarray(normals, np.uint32)
        out["shade"] = np.empty((H, W, 4), np.float32)
    g = out.get
    r = _lib.oracle_render_buffers(_ptr(nodes), nodes.shape[0], _ptr(colors), _ptr(slot_gid), nodes_per_treelet,
                                   slot_gid.size, depth, ctypes.byref(cam), POLICY.get(policy, policy),
                                   _ptr(out["t"]), _ptr(out["node"]), _ptr(out["treelet"]), _ptr(out["rgba"]),
                                   _ptr(out["boundary"]), _ptr(out["pushes"]), None, None, None,
                                   _ptr(g("rec_kind")), _ptr(g("rec_err")), _ptr(g("rec_sub")),
                                   _ptr(normals), _ptr(g("shade")), threads if threads > 0 else max_threads())
    if r != 0:
        raise RuntimeError(f"oracle_render_buffers failed with {r}")
    return out
