"""MI355X serialized-SVO treelet ray caster: host-side Python mirror.

Everything here is a thin layer over the C-ABI of ``libsvo_hip.so``
(include/svo_hip.h); the traversal runs only in the HIP kernels.  Class and
method names follow the reference's C++ API so that code written against
libsvo reads the same:

* :class:`Scene`                   -- the treelet bundle (``.svo`` + ``.ia``)
  as loaded by ``TreeletMemoryManager::loadFromFile``
  (src/TreeletMemoryManager.cpp:229-291) or built procedurally.
* :class:`TreeletMemoryManagerHip` -- ``TreeletMemoryManagerCl`` with HBM
  pools (src/TreeletMemoryManager.h:97-157, src/TreeletMemoryManagerCl.h:69-93).
* :func:`render_to_buffer`         -- the ``renderToBuffer`` launch
  (opencl/treeletRenderer_rgba.cl:610-655) for one or many cameras.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi
from ._capi import (BACKGROUND_RGBA, MISS, SVO_BUILD_PACK_BFS, SVO_POLICY_CPU3, SVO_POLICY_TREELET, SVO_STACK_LDS,
                    SVO_STACK_PRIVATE, SvoError, check, lib, svo_camera, svo_frame_out, svo_launch_config, svo_view)

__all__ = [
    "Scene", "TreeletMemoryManagerHip", "Camera", "DeviceBuffer", "Stream", "Frame", "render_to_buffer",
    "View", "Renderer", "TuneCache", "Batch", "LaunchConfig", "LAUNCH_SHAPES", "Server",
    "device_count", "look_at", "SvoError", "POLICIES", "MISS", "BACKGROUND_RGBA",
]

POLICIES = {"treelet": SVO_POLICY_TREELET, "single": _capi.SVO_POLICY_SINGLE, "cpu3": SVO_POLICY_CPU3}
Camera = svo_camera


def _policy(p) -> int:
    return POLICIES[p] if isinstance(p, str) else int(p)


def _packing(packing: str) -> int:
    if packing not in ("dense", "bfs"):
        raise ValueError(f"packing must be 'dense' or 'bfs', not {packing!r}")
    return SVO_BUILD_PACK_BFS if packing == "bfs" else 0


def device_count() -> int:
    n = ctypes.c_int(0)
    if lib.svo_device_count(ctypes.byref(n)) != 0:
        return 0
    return n.value


def with_view_mode(cam: svo_camera, mode) -> svo_camera:
    """Copy of `cam` with another view mode (name or SVO_VIEW_* number)."""
    c = svo_camera()
    ctypes.pointer(c)[0] = cam
    c.view_mode = _capi.VIEW_MODES[mode] if isinstance(mode, str) else int(mode)
    return c


def look_at(eye, target, up=(0.0, 1.0, 0.0), fov_y_deg: float = 60.0, width: int = 512, height: int = 512,
            far: float = 20.0, lod: float = 0.0) -> svo_camera:
    """Pinhole camera uniforms; ``lod`` is the reference's tScaleRatio
    multiplier (0 disables the level-of-detail stop)."""
    c = svo_camera()
    arr = lambda v: (ctypes.c_double * 3)(*[float(x) for x in v])
    check(lib.svo_camera_look_at(arr(eye), arr(target), arr(up), float(fov_y_deg), int(width), int(height),
                                 float(far), float(lod), ctypes.byref(c)), "svo_camera_look_at")
    return c


# ------------------------------------------------------------------ scenes

class Scene:
    """A treelet bundle held in host memory."""

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.svo_scene_destroy(h)
            self._h = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @classmethod
    def _build(cls, fn, *args) -> "Scene":
        h = ctypes.c_void_p()
        check(fn(*args, ctypes.byref(h)), fn.__name__)
        return cls(h)

    @classmethod
    def open(cls, base_path: str) -> "Scene":
        return cls._build(lib.svo_scene_open, base_path.encode())

    @classmethod
    def open_single(cls, base_path: str) -> "Scene":
        """A legacy single-SVO file (SvoLoader or SvoBranch v1) as a scene of
        one treelet."""
        return cls._build(lib.svo_scene_open_single, base_path.encode())

    def write_single(self, base_path: str, fmt: str = "v1", attrib_index=None) -> None:
        """Writes a one-treelet scene as a single-SVO file ("v1" or
        "svoloader"); attrib_index permutes the attribute packages (v1)."""
        code = {"svoloader": _capi.SVO_SINGLE_SVOLOADER, "v1": _capi.SVO_SINGLE_V1}[fmt]
        idx = None if attrib_index is None else np.ascontiguousarray(attrib_index, np.uint32)
        check(lib.svo_scene_write_single(self._h, base_path.encode(), code,
                                         None if idx is None else idx.ctypes.data), "svo_scene_write_single")

    @classmethod
    def menger(cls, level: int, depth: int, seed: int, treelet_bytes: int = 0, packing: str = "dense") -> "Scene":
        return cls._build(lib.svo_scene_build_menger, level, depth, seed, treelet_bytes, _packing(packing))

    @classmethod
    def terrain(cls, depth: int, seed: int, treelet_bytes: int = 32768, packing: str = "dense") -> "Scene":
        return cls._build(lib.svo_scene_build_terrain, depth, seed, treelet_bytes, _packing(packing))

    @classmethod
    def spheres(cls, depth: int, count: int, r_min: float, r_max: float, seed: int, treelet_bytes: int = 32768,
                packing: str = "dense") -> "Scene":
        return cls._build(lib.svo_scene_build_spheres, depth, count, r_min, r_max, seed, treelet_bytes,
                          _packing(packing))

    @classmethod
    def from_leaves(cls, depth: int, morton: np.ndarray, attribs: np.ndarray, treelet_bytes: int = 0,
                    packing: str = "dense") -> "Scene":
        m = np.ascontiguousarray(morton, dtype=np.uint64)
        a = np.ascontiguousarray(attribs, dtype=np.uint32).reshape(-1)
        if a.size != 2 * m.size:
            raise ValueError("attribs must hold two words per leaf")
        return cls._build(lib.svo_scene_build_from_leaves, depth, m.ctypes.data, a.ctypes.data, m.size,
                          treelet_bytes, _packing(packing))

    def write(self, base_path: str) -> None:
        check(lib.svo_scene_write(self._h, base_path.encode()), "svo_scene_write")

    @property
    def info(self) -> Dict[str, int]:
        i = _capi.svo_scene_info()
        check(lib.svo_scene_get_info(self._h, ctypes.byref(i)), "svo_scene_get_info")
        return {k: getattr(i, k) for k, _ in _capi.svo_scene_info._fields_}

    def treelet(self, gid: int):
        """(header dict, nodes (n,2) uint32 copy, attribs (n,2) uint32 copy)."""
        h = _capi.svo_treelet_header()
        pn, pa = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib.svo_scene_get_treelet(self._h, gid, ctypes.byref(h), ctypes.byref(pn), ctypes.byref(pa)),
              "svo_scene_get_treelet")
        n = h.num_nodes
        nodes = np.ctypeslib.as_array(ctypes.cast(pn, ctypes.POINTER(ctypes.c_uint32)), (n, 2)).copy()
        attr = np.ctypeslib.as_array(ctypes.cast(pa, ctypes.POINTER(ctypes.c_uint32)), (n, 2)).copy()
        return {k: getattr(h, k) for k, _ in _capi.svo_treelet_header._fields_}, nodes, attr

    def headers(self) -> np.ndarray:
        """(num_treelets, 8) uint32 array of every treelet header, in file order
        (num_nodes, num_leaves, mem_size, gid, parent_gid, parent_leaf_pos,
        parent_leaf_idx, parent_leaf_parent_pos)."""
        n = self.info["num_treelets"]
        out = np.empty((n, 8), np.uint32)
        h = _capi.svo_treelet_header()
        for g in range(n):
            check(lib.svo_scene_get_treelet(self._h, g, ctypes.byref(h), None, None), "svo_scene_get_treelet")
            out[g] = [getattr(h, k) for k, _ in _capi.svo_treelet_header._fields_]
        return out

    def getTreeletSizeInByte(self) -> int:
        return self.info["treelet_bytes"]

    def getNumNodesPerTreelet(self) -> int:
        return self.info["nodes_per_treelet"]


# ------------------------------------------------------------ device memory

class DeviceBuffer:
    """Raw HBM allocation owned by Python."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(lib.svo_device_malloc(ctypes.byref(self.ptr), max(1, self.nbytes)), "svo_device_malloc")

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib.svo_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def to_numpy(self, dtype, count: Optional[int] = None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        check(lib.svo_memcpy_d2h(out.ctypes.data, self.ptr, n * dt.itemsize), "svo_memcpy_d2h")
        return out

    def fill(self, byte: int) -> None:
        check(lib.svo_memset_device(self.ptr, byte, self.nbytes), "svo_memset_device")


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib.svo_stream_create(ctypes.byref(self.handle)), "svo_stream_create")

    def synchronize(self) -> None:
        check(lib.svo_stream_synchronize(self.handle), "svo_stream_synchronize")

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            lib.svo_stream_destroy(self.handle)
            self.handle = ctypes.c_void_p()


# ------------------------------------------------------------ residency

class TreeletMemoryManagerHip:
    """Incore buffer of ``incore_bytes`` split into treelet slots, mirrored in
    HBM.  Construction inserts the root treelet and uploads the whole buffer
    (``TreeletMemoryManagerCl`` ctor + ``initClBuffer``)."""

    def __init__(self, scene: Scene, incore_bytes: int, device: int = 0, resident_all: bool = False):
        self.scene = scene  # keeps the host bundle alive
        self._h = ctypes.c_void_p()
        check(lib.svo_mm_create(scene.handle, int(incore_bytes), int(device), ctypes.byref(self._h)),
              "svo_mm_create")
        if resident_all:
            self.make_resident_all()
            self.updateDeviceMemory()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.svo_mm_destroy(h)
            self._h = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def info(self) -> Dict[str, int]:
        i = _capi.svo_mm_info()
        check(lib.svo_mm_get_info(self._h, ctypes.byref(i)), "svo_mm_get_info")
        return {k: getattr(i, k) for k, _ in _capi.svo_mm_info._fields_}

    def resetIncoreBuffer(self) -> None:
        check(lib.svo_mm_reset(self._h), "svo_mm_reset")

    def insertTreeletIntoIncoreBuffer(self, gid: int, error: float = 0.0) -> None:
        check(lib.svo_mm_insert(self._h, gid, error), "svo_mm_insert")

    def removeTreeletFromIncoreBuffer(self, gid: int) -> None:
        check(lib.svo_mm_remove(self._h, gid), "svo_mm_remove")

    def make_resident_all(self) -> None:
        check(lib.svo_mm_make_resident_all(self._h), "svo_mm_make_resident_all")

    def updateDeviceMemory(self, stream: Optional[Stream] = None) -> Dict[str, int]:
        s = _capi.svo_upload_stats()
        check(lib.svo_mm_update_device(self._h, stream.handle if stream else None, ctypes.byref(s)),
              "svo_mm_update_device")
        return {k: getattr(s, k) for k, _ in _capi.svo_upload_stats._fields_}

    def getNumNodesPerTreelet(self) -> int:
        return self.info["nodes_per_treelet"]

    def performAnalysePass(self, cam: svo_camera, divide: int = 8, jitter=(0.0, 0.0), policy="treelet",
                           stream: Optional["Stream"] = None) -> None:
        """Analyse pass over a (W/divide) x (H/divide) frame of `cam`; its
        records accumulate on the device until updateClientSideIncoreBuffer."""
        check(lib.svo_mm_analyse(self._h, ctypes.byref(cam), int(divide), float(jitter[0]), float(jitter[1]),
                                 _policy(policy), stream.handle if stream else None), "svo_mm_analyse")

    def updateClientSideIncoreBuffer(self, stream: Optional["Stream"] = None) -> Dict[str, int]:
        """Applies the recorded feedback: visibility update and page-ins
        (call updateDeviceMemory afterwards)."""
        st = _capi.svo_feedback_stats()
        check(lib.svo_mm_apply_feedback(self._h, stream.handle if stream else None, ctypes.byref(st)),
              "svo_mm_apply_feedback")
        return {k: getattr(st, k) for k, _ in _capi.svo_feedback_stats._fields_}

    def fetchFeedback(self, stream: Optional["Stream"] = None) -> None:
        """Enqueues the readback of the records gathered since the last
        update on `stream` (behind the launches that recorded them) and
        returns at once; the next updateClientSideIncoreBuffer waits for it
        instead of reading and builds the visible sets then.  Lets a caller
        enqueue its render behind the readback and update the residency
        while the render runs."""
        check(lib.svo_mm_fetch_feedback(self._h, stream.handle if stream else None), "svo_mm_fetch_feedback")

    def readFeedback(self, stream: Optional["Stream"] = None) -> Dict[str, List[Tuple[int, float]]]:
        """The visible sets of the records gathered since the last update
        (RenderPassAnalyse::getVisibleOldTreeletsGids / getVisibleNewTreeletsGids):
        ``old`` (gid, error) in ascending gid, ``new`` (gid, error) by
        descending error, capped.  The next updateClientSideIncoreBuffer
        applies these records."""
        sh = stream.handle if stream else None
        n_old, n_new = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib.svo_mm_read_feedback(self._h, sh, None, None, 0, ctypes.byref(n_old), None, None, 0,
                                       ctypes.byref(n_new)), "svo_mm_read_feedback")
        og = np.zeros(max(1, n_old.value), np.uint32)
        oe = np.zeros(max(1, n_old.value), np.float32)
        ng = np.zeros(max(1, n_new.value), np.uint32)
        ne = np.zeros(max(1, n_new.value), np.float32)
        check(lib.svo_mm_read_feedback(self._h, sh, og.ctypes.data, oe.ctypes.data, og.size, ctypes.byref(n_old),
                                       ng.ctypes.data, ne.ctypes.data, ng.size, ctypes.byref(n_new)),
              "svo_mm_read_feedback")
        return {"old": [(int(g), float(e)) for g, e in zip(og[:n_old.value], oe[:n_old.value])],
                "new": [(int(g), float(e)) for g, e in zip(ng[:n_new.value], ne[:n_new.value])]}

    def getVisibleOldTreeletsGids(self, stream: Optional["Stream"] = None) -> List[Tuple[int, float]]:
        return self.readFeedback(stream)["old"]

    def getVisibleNewTreeletsGids(self, stream: Optional["Stream"] = None) -> List[Tuple[int, float]]:
        return self.readFeedback(stream)["new"]

    def applyRecords(self, slots, slot_errors, gids, gid_errors) -> Dict[str, int]:
        """The same update from aggregated records given by the caller."""
        s = np.ascontiguousarray(slots, np.uint32)
        se = np.ascontiguousarray(slot_errors, np.float32)
        g = np.ascontiguousarray(gids, np.uint32)
        ge = np.ascontiguousarray(gid_errors, np.float32)
        if s.size != se.size or g.size != ge.size:
            raise ValueError("one error per slot and per gid")
        st = _capi.svo_feedback_stats()
        check(lib.svo_mm_apply_records(self._h, s.ctypes.data, se.ctypes.data, s.size, g.ctypes.data, ge.ctypes.data,
                                       g.size, ctypes.byref(st)), "svo_mm_apply_records")
        return {k: getattr(st, k) for k, _ in _capi.svo_feedback_stats._fields_}

    def getSlots(self) -> List[int]:
        """getSlots()[i]._treeletGid: the gid held by each slot, -1 empty."""
        return [g for g, _, _ in self.getSlotInfos()]

    def getSlotGid(self, gid: int) -> int:
        s = ctypes.c_int32()
        check(lib.svo_mm_treelet_slot(self._h, gid, ctypes.byref(s)), "svo_mm_treelet_slot")
        return s.value

    def getIncoreBuffer(self, first: int = 0, count: Optional[int] = None):
        """(nodes (n,2) uint32, attribs (n,2) uint32) of the host incore buffer."""
        n = self.info["pool_nodes"] - first if count is None else count
        nodes = np.empty((n, 2), np.uint32)
        attr = np.empty((n, 2), np.uint32)
        check(lib.svo_mm_read_host_incore(self._h, first, n, nodes.ctypes.data, attr.ctypes.data),
              "svo_mm_read_host_incore")
        return nodes, attr

    def getIncoreBufferView(self) -> np.ndarray:
        """The host incore buffer itself (getIncoreBuffer), as a zero-copy
        (pool_nodes, 2) uint32 view {first_child, masks}; valid while this
        manager lives, changed by its updates."""
        p, n = ctypes.c_void_p(), ctypes.c_uint64()
        check(lib.svo_mm_get_incore_buffer(self._h, ctypes.byref(p), ctypes.byref(n)), "svo_mm_get_incore_buffer")
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), (n.value, 2))

    def getAttributeIncoreBufferView(self) -> Tuple[np.ndarray, np.ndarray]:
        """The attribute incore buffer itself (getAttributeIncoreBuffer):
        zero-copy uint32 views (normals, colors), one word per incore node."""
        pn, pc, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_uint64()
        check(lib.svo_mm_get_attribute_incore_buffer(self._h, ctypes.byref(pn), ctypes.byref(pc), ctypes.byref(n)),
              "svo_mm_get_attribute_incore_buffer")
        u = ctypes.POINTER(ctypes.c_uint32)
        return (np.ctypeslib.as_array(ctypes.cast(pn, u), (n.value,)),
                np.ctypeslib.as_array(ctypes.cast(pc, u), (n.value,)))

    def getAttributeIncoreBuffer(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        """(n, 2) uint32 copy of the attribute packages {packed_normal,
        packed_color} of incore nodes [first, first+count)."""
        nrm, col = self.getAttributeIncoreBufferView()
        end = nrm.size if count is None else first + count
        return np.stack([nrm[first:end], col[first:end]], axis=1)

    def getTreelets(self) -> Tuple[np.ndarray, np.ndarray]:
        """getTreelets(): ((num_treelets, 8) uint32 headers in file order,
        (num_treelets,) int32 slot of each treelet, -1 = not resident)."""
        n = self.scene.info["num_treelets"]
        hdr = (_capi.svo_treelet_header * n)()
        slots = np.empty(n, np.int32)
        check(lib.svo_mm_get_treelets(self._h, 0, n, hdr, slots.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
              "svo_mm_get_treelets")
        h = np.ctypeslib.as_array(ctypes.cast(hdr, ctypes.POINTER(ctypes.c_uint32)), (n, 8)).copy()
        return h, slots

    def getTreelet(self, gid: int) -> Dict[str, int]:
        """Header of treelet `gid` plus its slot (-1 = not resident)."""
        hdr = _capi.svo_treelet_header()
        slot = ctypes.c_int32()
        check(lib.svo_mm_get_treelets(self._h, gid, 1, ctypes.byref(hdr), ctypes.byref(slot)), "svo_mm_get_treelets")
        out = {k: getattr(hdr, k) for k, _ in _capi.svo_treelet_header._fields_}
        out["slot"] = slot.value
        return out

    def freeIncorePosition(self) -> int:
        """freeIncorePosition(): frees one invisible treelet (and everything
        resident below it) found by the reference's probe rule; returns its
        gid, -1 if none was found."""
        g = ctypes.c_int32()
        check(lib.svo_mm_free_incore_position(self._h, ctypes.byref(g)), "svo_mm_free_incore_position")
        return g.value

    def _u32_list(self, fn, name) -> List[int]:
        n = ctypes.c_uint32()
        check(fn(self._h, None, 0, ctypes.byref(n)), name)
        buf = (ctypes.c_uint32 * max(1, n.value))()
        check(fn(self._h, buf, n.value, ctypes.byref(n)), name)
        return list(buf[:n.value])

    def getFreeIncoreSlotStack(self) -> List[int]:
        """Free slots, bottom of the stack first (the last one is taken next)."""
        return self._u32_list(lib.svo_mm_get_free_slot_stack, "svo_mm_get_free_slot_stack")

    def getIncoreSlotsToUpload(self) -> List[int]:
        """Slots changed since the last updateDeviceMemory, ascending."""
        return self._u32_list(lib.svo_mm_get_slots_to_upload, "svo_mm_get_slots_to_upload")

    def getSlotInfos(self) -> List[Tuple[int, int, float]]:
        """getSlots() with the whole SlotInfo: (treelet gid, visibility,
        error) per slot."""
        n = self.info["num_slots"]
        buf = np.empty(n, dtype=[("g", np.int32), ("v", np.int32), ("e", np.float32), ("r", np.uint32)])
        check(lib.svo_mm_get_slots(self._h, 0, n, buf.ctypes.data_as(ctypes.POINTER(_capi.svo_slot_info))),
              "svo_mm_get_slots")
        return list(zip(buf["g"].tolist(), buf["v"].tolist(), buf["e"].tolist()))

    def getDevicePools(self) -> Dict[str, int]:
        """Device pointers of the HBM pools (getClIncoreBufferGid /
        getClAttributeIncoreBufferGid); from this call on every update also
        sends the normals."""
        p = _capi.svo_device_pools()
        check(lib.svo_mm_get_device_pools(self._h, ctypes.byref(p)), "svo_mm_get_device_pools")
        return {k: getattr(p, k) for k, _ in _capi.svo_device_pools._fields_ if k != "reserved"}

    def initDeviceBuffer(self, stream: Optional["Stream"] = None) -> None:
        """initClBuffer(): uploads the whole incore buffer again."""
        check(lib.svo_mm_init_device_buffer(self._h, stream.handle if stream else None), "svo_mm_init_device_buffer")

    def getDeviceIncoreBuffer(self, first: int = 0, count: Optional[int] = None):
        n = self.info["pool_nodes"] - first if count is None else count
        nodes = np.empty((n, 2), np.uint32)
        attr = np.empty((n, 2), np.uint32)
        check(lib.svo_mm_read_device_incore(self._h, first, n, nodes.ctypes.data, attr.ctypes.data),
              "svo_mm_read_device_incore")
        return nodes, attr


# ------------------------------------------------------------ rendering

@dataclass
class Frame:
    """Device outputs of one camera (width*height each; shade 4 floats per
    pixel)."""
    width: int
    height: int
    t: DeviceBuffer
    node: DeviceBuffer
    treelet: DeviceBuffer
    rgba: DeviceBuffer
    pushes: Optional[DeviceBuffer] = None
    shade: Optional[DeviceBuffer] = None

    @classmethod
    def alloc(cls, width: int, height: int, pushes: bool = False, shade: bool = False) -> "Frame":
        n = width * height
        return cls(width, height, DeviceBuffer(4 * n), DeviceBuffer(4 * n), DeviceBuffer(4 * n), DeviceBuffer(4 * n),
                   DeviceBuffer(4 * n) if pushes else None, DeviceBuffer(16 * n) if shade else None)

    def out(self) -> svo_frame_out:
        return svo_frame_out(self.t.ptr, self.node.ptr, self.treelet.ptr, self.rgba.ptr,
                             self.pushes.ptr if self.pushes is not None else None,
                             self.shade.ptr if self.shade is not None else None)

    def fetch(self) -> Dict[str, np.ndarray]:
        shp = (self.height, self.width)
        out = {"t": self.t.to_numpy(np.float32).reshape(shp), "node": self.node.to_numpy(np.uint32).reshape(shp),
               "treelet": self.treelet.to_numpy(np.uint32).reshape(shp),
               "rgba": self.rgba.to_numpy(np.uint32).reshape(shp)}
        if self.pushes is not None:
            out["pushes"] = self.pushes.to_numpy(np.uint32).reshape(shp)
        if self.shade is not None:
            out["shade"] = self.shade.to_numpy(np.float32).reshape(self.height, self.width, 4)
        return out


def _cam_array(cams: Sequence[svo_camera]):
    arr = (svo_camera * len(cams))()
    for i, c in enumerate(cams):
        arr[i] = c
    return arr


def render_to_buffer(mm: TreeletMemoryManagerHip, cams: Sequence[svo_camera], frames: Sequence[Frame],
                     policy="treelet", stream: Optional[Stream] = None) -> None:
    """Enqueues one launch that renders every camera into its frame."""
    if len(cams) != len(frames) or not cams:
        raise ValueError("one frame per camera required")
    outs = (svo_frame_out * len(frames))(*[f.out() for f in frames])
    check(lib.svo_render(mm.handle, _cam_array(cams), len(cams), outs, _policy(policy),
                         stream.handle if stream else None), "svo_render")


def render_timed(mm: TreeletMemoryManagerHip, cams: Sequence[svo_camera], frames: Sequence[Frame], policy="treelet",
                 stream: Optional[Stream] = None, iters: int = 10) -> float:
    """Mean milliseconds per launch, timed with HIP events on the launch stream."""
    outs = (svo_frame_out * len(frames))(*[f.out() for f in frames])
    ms = ctypes.c_float()
    check(lib.svo_render_timed(mm.handle, _cam_array(cams), len(cams), outs, _policy(policy),
                               stream.handle if stream else None, iters, ctypes.byref(ms)), "svo_render_timed")
    return ms.value


def render(mm: TreeletMemoryManagerHip, cam: svo_camera, policy="treelet", pushes: bool = True,
           shade: bool = False) -> Dict[str, np.ndarray]:
    """Renders one camera and returns host copies of t, node, treelet, rgba
    (and the per-ray push counts, and the view-mode image when shade)."""
    f = Frame.alloc(cam.width, cam.height, pushes, shade)
    render_to_buffer(mm, [cam], [f], policy)
    check(lib.svo_device_synchronize(), "svo_device_synchronize")
    return f.fetch()


def analysis_camera(cam: svo_camera, divide: int = 8, jitter=(0.0, 0.0)) -> svo_camera:
    """The reduced frame an analyse pass traces for `cam`."""
    out = svo_camera()
    check(lib.svo_analysis_camera(ctypes.byref(cam), int(divide), float(jitter[0]), float(jitter[1]),
                                  ctypes.byref(out)), "svo_analysis_camera")
    return out


def ray_dirs(cam: svo_camera) -> np.ndarray:
    buf = DeviceBuffer(12 * cam.width * cam.height)
    check(lib.svo_debug_ray_dirs(ctypes.byref(cam), buf.ptr, None), "svo_debug_ray_dirs")
    check(lib.svo_device_synchronize(), "svo_device_synchronize")
    return buf.to_numpy(np.float32).reshape(cam.height, cam.width, 3)


# ------------------------------------------------------------ multi-view launches

LaunchConfig = svo_launch_config
STACKS = {"lds": SVO_STACK_LDS, "private": SVO_STACK_PRIVATE, "lds_node": _capi.SVO_STACK_LDS_NODE}
STACK_NAMES = {v: k for k, v in STACKS.items()}
#: every (tile_width, waves_per_group, stack_layout) the kernels are built for
LAUNCH_SHAPES = [(t, w, s) for t in (8, 16, 32) for w in (1, 2, 4) for s in ("lds", "private", "lds_node")]


@dataclass
class View:
    """One camera of one resident scene rendered into one frame (frame None:
    no outputs, e.g. a feedback-only view).  feedback_stride > 0 records
    analysis feedback for every feedback_stride-th pixel in x and y."""
    mm: TreeletMemoryManagerHip
    camera: svo_camera
    frame: Optional[Frame]
    feedback_stride: int = 0

    def c_view(self) -> svo_view:
        out = self.frame.out() if self.frame is not None else svo_frame_out()
        return svo_view(self.mm.handle, self.camera, out, int(self.feedback_stride), 0)


def _view_array(views: Sequence[View]):
    if not views:
        raise ValueError("at least one view required")
    return (svo_view * len(views))(*[v.c_view() for v in views])


class Renderer:
    """Launch state of multi-view rendering on one device: the launch shape
    and a ring of device-side view descriptors."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        check(lib.svo_renderer_create(device, ctypes.byref(self._h)), "svo_renderer_create")

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.svo_renderer_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def config(self):
        c = svo_launch_config()
        check(lib.svo_renderer_get_config(self._h, ctypes.byref(c)), "svo_renderer_get_config")
        return (c.tile_width, c.waves_per_group, STACK_NAMES[c.stack_layout])

    @config.setter
    def config(self, shape) -> None:
        if shape is None:
            check(lib.svo_renderer_set_config(self._h, None), "svo_renderer_set_config")
            return
        t, w, s = shape
        c = svo_launch_config(t, w, STACKS[s] if isinstance(s, str) else s, 0)
        check(lib.svo_renderer_set_config(self._h, ctypes.byref(c)), "svo_renderer_set_config")

    def render(self, views: Sequence[View], policy="treelet", stream: Optional[Stream] = None) -> None:
        """Enqueues ONE launch rendering every view."""
        check(lib.svo_render_views(self._h, _view_array(views), len(views), _policy(policy),
                                   stream.handle if stream else None), "svo_render_views")

    def timed(self, views: Sequence[View], policy="treelet", stream: Optional[Stream] = None, iters: int = 10) -> float:
        ms = ctypes.c_float()
        check(lib.svo_render_views_timed(self._h, _view_array(views), len(views), _policy(policy),
                                         stream.handle if stream else None, iters, ctypes.byref(ms)),
              "svo_render_views_timed")
        return ms.value

    def autotune(self, views: Sequence[View], policy="treelet", stream: Optional[Stream] = None, iters: int = 3,
                 cache: Optional["TuneCache"] = None):
        """Times every launch shape on these views, keeps the fastest and
        returns ((tile_width, waves, stack), ms).  With a :class:`TuneCache`
        a pick recorded for the same workload class is taken without timing
        (``cache.last_hit`` tells which happened) and a new pick is
        recorded."""
        best, ms = svo_launch_config(), ctypes.c_float()
        if cache is not None:
            hit = ctypes.c_int()
            check(lib.svo_autotune_cached(self._h, cache.handle, _view_array(views), len(views), _policy(policy),
                                          stream.handle if stream else None, iters, ctypes.byref(best),
                                          ctypes.byref(ms), ctypes.byref(hit)), "svo_autotune_cached")
            cache.last_hit = bool(hit.value)
            return self.config, ms.value
        check(lib.svo_autotune(self._h, _view_array(views), len(views), _policy(policy),
                               stream.handle if stream else None, iters, ctypes.byref(best), ctypes.byref(ms)),
              "svo_autotune")
        return self.config, ms.value

    def tune_key(self, views: Sequence[View], policy="treelet") -> str:
        """The autotune cache key of these views on this renderer's device."""
        buf = ctypes.create_string_buffer(4096)
        check(lib.svo_tune_cache_key(self._h, _view_array(views), len(views), _policy(policy), buf, len(buf)),
              "svo_tune_cache_key")
        return buf.value.decode()


class TuneCache:
    """Launch shapes picked by :meth:`Renderer.autotune`, per workload class
    (device architecture and CU count, policy, number of views and rays,
    scene classes), optionally kept in a text file between processes."""

    def __init__(self, path: Optional[str] = None):
        self._h = ctypes.c_void_p()
        check(lib.svo_tune_cache_create(ctypes.byref(self._h)), "svo_tune_cache_create")
        self.path = path
        self.last_hit = False
        if path is not None:
            self.load(path)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.svo_tune_cache_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def __len__(self) -> int:
        return int(lib.svo_tune_cache_size(self._h))

    def load(self, path: str) -> None:
        check(lib.svo_tune_cache_load(self._h, os.fsencode(path)), "svo_tune_cache_load")

    def save(self, path: Optional[str] = None) -> None:
        path = path or self.path
        if path is None:
            raise ValueError("TuneCache.save: no path")
        check(lib.svo_tune_cache_save(self._h, os.fsencode(path)), "svo_tune_cache_save")


class Batch:
    """A multi-view launch captured once into a hipGraph; replay with
    :meth:`launch`, move the cameras with :meth:`set_cameras`."""

    def __init__(self, renderer: Renderer, views: Sequence[View], policy="treelet", stream: Optional[Stream] = None):
        self._renderer = renderer          # keeps the renderer alive
        self._views = list(views)          # keeps the frames and residency handles alive
        self._h = ctypes.c_void_p()
        check(lib.svo_batch_create(renderer.handle, _view_array(views), len(views), _policy(policy),
                                   stream.handle if stream else None, ctypes.byref(self._h)), "svo_batch_create")

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.svo_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def set_cameras(self, cams: Sequence[svo_camera], stream: Optional[Stream] = None) -> None:
        check(lib.svo_batch_set_cameras(self._h, _cam_array(cams), len(cams), stream.handle if stream else None),
              "svo_batch_set_cameras")
        for v, c in zip(self._views, cams):
            v.camera = c

    def launch(self, stream: Optional[Stream] = None) -> None:
        check(lib.svo_batch_launch(self._h, stream.handle if stream else None), "svo_batch_launch")


# ------------------------------------------------------------ serving

class Server:
    """Continuous-batching render server over resident scenes (one device).

    Requests without a :class:`Frame` render into the server's frame pool
    (results stay in HBM); requests with one render into it."""

    def __init__(self, renderer: Renderer, hbm_budget: int = 0, frame_pool: int = 1 << 30, max_batch_rays: int = 0,
                 max_batch_views: int = 64, max_inflight: int = 4, policy="treelet", max_waiting: int = 0):
        self._renderer = renderer
        self._scenes: List[TreeletMemoryManagerHip] = []
        self._frames = {}
        cfg = _capi.svo_server_config(int(hbm_budget), int(frame_pool), int(max_batch_rays), int(max_batch_views),
                                      int(max_inflight), _policy(policy), int(max_waiting))
        self._h = ctypes.c_void_p()
        check(lib.svo_server_create(renderer.handle, ctypes.byref(cfg), ctypes.byref(self._h)), "svo_server_create")

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.svo_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def add_scene(self, mm: TreeletMemoryManagerHip) -> int:
        sid = ctypes.c_uint32()
        check(lib.svo_server_add_scene(self._h, mm.handle, ctypes.byref(sid)), "svo_server_add_scene")
        self._scenes.append(mm)
        return sid.value

    @staticmethod
    def _request(scene: int, cam: svo_camera, arrival_s: float = 0.0, frame: Optional[Frame] = None,
                 timeout_us: int = 0):
        out = frame.out() if frame is not None else svo_frame_out()
        return _capi.svo_request(scene, int(timeout_us), float(arrival_s), cam, out)

    def submit(self, scene: int, cam: svo_camera, frame: Optional[Frame] = None, timeout_us: int = 0) -> int:
        """Queues a request (SvoError with status SVO_ERR_CAPACITY when
        max_waiting requests already wait)."""
        rid = ctypes.c_uint64()
        req = self._request(scene, cam, 0.0, frame, timeout_us)
        check(lib.svo_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)), "svo_server_submit")
        if frame is not None:
            self._frames[rid.value] = frame  # kept alive until completion
        return rid.value

    def poll(self, cap: int = 256) -> List[Dict[str, float]]:
        buf = (_capi.svo_completion * cap)()
        n = ctypes.c_uint32()
        check(lib.svo_server_poll(self._h, buf, cap, ctypes.byref(n)), "svo_server_poll")
        out = []
        for c in buf[:n.value]:
            self._frames.pop(c.id, None)
            out.append({k: getattr(c, k) for k, _ in _capi.svo_completion._fields_})
        return out

    @property
    def pending(self) -> int:
        n = ctypes.c_uint32()
        check(lib.svo_server_pending(self._h, ctypes.byref(n)), "svo_server_pending")
        return n.value

    def replay(self, requests: Sequence[tuple]):
        """Serves (scene, camera, arrival_s[, frame[, timeout_us]]) requests
        in real time.  Returns (stats dict, completions array with fields
        arrival_s, launch_s, done_s, batch, batch_views, status, in request
        order)."""
        n = len(requests)
        reqs = (_capi.svo_request * n)(*[self._request(*r) for r in requests])
        done = (_capi.svo_completion * n)()
        st = _capi.svo_server_stats()
        check(lib.svo_server_replay(self._h, reqs, n, done, ctypes.byref(st)), "svo_server_replay")
        stats = {k: getattr(st, k) for k, _ in _capi.svo_server_stats._fields_}
        comp = np.ctypeslib.as_array(done).copy() if n else np.zeros(0)
        return stats, comp
