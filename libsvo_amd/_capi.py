"""ctypes bindings of libsvo_hip.so (include/svo_hip.h).

The library is built in-tree by ``make -C libsvo_amd`` (``__graft_entry__.build``).
Importing this module without the built library raises ImportError: there is
no fallback path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SVO_HIP_LIBRARY may name another build of the same library (the
# sanitizer build, libsvo_amd/_san/libsvo_hip.so)
LIB_PATH = os.environ.get("SVO_HIP_LIBRARY") or os.path.join(_HERE, "libsvo_hip.so")

u32, u64, i32, f32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float, ctypes.c_double
vp = ctypes.c_void_p
P = ctypes.POINTER

SVO_OK = 0
SVO_ERR_ARG, SVO_ERR_IO, SVO_ERR_FORMAT, SVO_ERR_HIP, SVO_ERR_NOMEM, SVO_ERR_STATE, SVO_ERR_CAPACITY = range(1, 8)
SVO_POLICY_TREELET = 0
SVO_POLICY_SINGLE = 1
SVO_REQ_DONE, SVO_REQ_EXPIRED, SVO_REQ_REJECTED = 0, 1, 2
SVO_POLICY_CPU3 = 2
SVO_BUILD_PACK_BFS = 1
SVO_SINGLE_SVOLOADER, SVO_SINGLE_V1 = 1, 2
SVO_STACK_LDS, SVO_STACK_PRIVATE, SVO_STACK_LDS_NODE = 0, 1, 2
SVO_NOALLOC_MAX_VIEWS = 256
VIEW_MODES = {"color": 0, "normal": 1, "iterations": 2, "svo_depth": 3, "t": 4, "loops": 5, "phong": 6, "quality": 7,
              "fill_svo_depth": 8, "fill_loops": 9, "fill_reflection": 10, "fill_shadow": 11}
# the single-SVO renderer's otherParams.x (opencl/fillFrameBuffer_rgba.cl:633-700) as view modes
FILL_VIEW_MODES = {0: "color", 1: "normal", 2: "fill_svo_depth", 3: "t", 4: "fill_loops", 5: "fill_reflection",
                   6: "fill_shadow"}
MISS = 0xFFFFFFFF
BACKGROUND_RGBA = 0x334D33FF


class svo_node(ctypes.Structure):
    _fields_ = [("first_child", u32), ("masks", u32)]


class svo_treelet_header(ctypes.Structure):
    _fields_ = [(n, u32) for n in ("num_nodes", "num_leaves", "mem_size", "gid", "parent_gid",
                                    "parent_leaf_pos", "parent_leaf_idx", "parent_leaf_parent_pos")]


class svo_scene_info(ctypes.Structure):
    _fields_ = [("num_treelets", u32), ("treelet_bytes", u32), ("nodes_per_treelet", u32),
                ("max_depth", u32), ("total_nodes", u64), ("tree_nodes", u64), ("tree_leaves", u64)]


class svo_camera(ctypes.Structure):
    _fields_ = [("far_lower_left", f32 * 4), ("pixel_h", f32 * 4), ("pixel_v", f32 * 4),
                ("position", f32 * 4), ("t_scale_ratio", f32), ("width", u32), ("height", u32),
                ("view_mode", u32)]


class svo_mm_info(ctypes.Structure):
    _fields_ = [("num_slots", u32), ("nodes_per_treelet", u32), ("free_slots", u32),
                ("dirty_slots", u32), ("resident_treelets", u32), ("device", u32),
                ("pool_nodes", u64), ("device_bytes", u64)]


class svo_upload_stats(ctypes.Structure):
    _fields_ = [("slots_uploaded", u32), ("copy_calls", u32), ("bytes", u64), ("bytes_moved", u64)]


class svo_frame_out(ctypes.Structure):
    _fields_ = [("t", vp), ("node", vp), ("treelet", vp), ("rgba", vp), ("pushes", vp), ("shade", vp)]


class svo_view(ctypes.Structure):
    _fields_ = [("mm", vp), ("camera", svo_camera), ("out", svo_frame_out), ("feedback_stride", u32),
                ("reserved", u32)]


class svo_feedback_stats(ctypes.Structure):
    _fields_ = [(n, u32) for n in ("visible_slots", "visible_treelets", "requests", "requests_kept", "inserted",
                                    "insert_failed", "evicted", "resident_treelets")]


class svo_server_config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", u64), ("frame_pool_bytes", u64), ("max_batch_rays", u64),
                ("max_batch_views", u32), ("max_inflight", u32), ("policy", i32), ("max_waiting", u32)]


class svo_request(ctypes.Structure):
    _fields_ = [("scene", u32), ("timeout_us", u32), ("arrival_s", f64), ("camera", svo_camera),
                ("out", svo_frame_out)]


class svo_completion(ctypes.Structure):
    _fields_ = [("id", u64), ("arrival_s", f64), ("launch_s", f64), ("done_s", f64), ("batch", u32),
                ("batch_views", u32), ("status", u32), ("reserved", u32)]


class svo_server_stats(ctypes.Structure):
    _fields_ = [("requests", u64), ("rays", u64), ("batches", u64), ("peak_pool_bytes", u64),
                ("max_batch_views_seen", u32), ("max_inflight_seen", u32), ("wall_s", f64),
                ("frames_per_s", f64), ("mrays_per_s", f64), ("p50_ms", f64), ("p99_ms", f64),
                ("mean_ms", f64), ("max_ms", f64), ("mean_batch_views", f64), ("expired", u64),
                ("rejected", u64)]


class svo_slot_info(ctypes.Structure):
    _fields_ = [("treelet_gid", i32), ("visibility", i32), ("error", f32), ("reserved", u32)]


class svo_device_pools(ctypes.Structure):
    _fields_ = [("nodes", vp), ("colors", vp), ("normals", vp), ("slot_gid", vp), ("pool_nodes", u64),
                ("num_slots", u32), ("nodes_per_treelet", u32), ("device", i32), ("reserved", u32)]


class svo_launch_config(ctypes.Structure):
    _fields_ = [("tile_width", u32), ("waves_per_group", u32), ("stack_layout", u32), ("reserved", u32)]


# name -> (restype, argtypes); every function declared in include/svo_hip.h
SIGNATURES = {
    "svo_last_error": (ctypes.c_char_p, []),
    "svo_scene_open": (ctypes.c_int, [ctypes.c_char_p, P(vp)]),
    "svo_scene_write": (ctypes.c_int, [vp, ctypes.c_char_p]),
    "svo_scene_destroy": (None, [vp]),
    "svo_scene_get_info": (ctypes.c_int, [vp, P(svo_scene_info)]),
    "svo_scene_get_treelet": (ctypes.c_int, [vp, u32, P(svo_treelet_header), P(vp), P(vp)]),
    "svo_scene_build_from_leaves": (ctypes.c_int, [u32, vp, vp, ctypes.c_size_t, u32, u32, P(vp)]),
    "svo_scene_open_single": (ctypes.c_int, [ctypes.c_char_p, P(vp)]),
    "svo_scene_write_single": (ctypes.c_int, [vp, ctypes.c_char_p, u32, vp]),
    "svo_scene_build_menger": (ctypes.c_int, [u32, u32, u64, u32, u32, P(vp)]),
    "svo_scene_build_terrain": (ctypes.c_int, [u32, u64, u32, u32, P(vp)]),
    "svo_scene_build_spheres": (ctypes.c_int, [u32, u32, f32, f32, u64, u32, u32, P(vp)]),
    "svo_terrain_height": (ctypes.c_int, [u64, f64, f64, P(f64)]),
    "svo_camera_look_at": (ctypes.c_int, [P(f64), P(f64), P(f64), f64, u32, u32, f64, f64, P(svo_camera)]),
    "svo_mm_create": (ctypes.c_int, [vp, u64, ctypes.c_int, P(vp)]),
    "svo_mm_destroy": (None, [vp]),
    "svo_mm_reset": (ctypes.c_int, [vp]),
    "svo_mm_insert": (ctypes.c_int, [vp, u32, f32]),
    "svo_mm_remove": (ctypes.c_int, [vp, u32]),
    "svo_mm_make_resident_all": (ctypes.c_int, [vp]),
    "svo_mm_update_device": (ctypes.c_int, [vp, vp, P(svo_upload_stats)]),
    "svo_mm_get_info": (ctypes.c_int, [vp, P(svo_mm_info)]),
    "svo_mm_slot_treelet": (ctypes.c_int, [vp, u32, P(i32)]),
    "svo_mm_treelet_slot": (ctypes.c_int, [vp, u32, P(i32)]),
    "svo_mm_read_host_incore": (ctypes.c_int, [vp, u64, u64, vp, vp]),
    "svo_mm_read_device_incore": (ctypes.c_int, [vp, u64, u64, vp, vp]),
    "svo_mm_get_treelets": (ctypes.c_int, [vp, u32, u32, P(svo_treelet_header), P(i32)]),
    "svo_mm_get_incore_buffer": (ctypes.c_int, [vp, P(vp), P(u64)]),
    "svo_mm_get_attribute_incore_buffer": (ctypes.c_int, [vp, P(vp), P(vp), P(u64)]),
    "svo_mm_free_incore_position": (ctypes.c_int, [vp, P(i32)]),
    "svo_mm_get_free_slot_stack": (ctypes.c_int, [vp, P(u32), u32, P(u32)]),
    "svo_mm_get_slots_to_upload": (ctypes.c_int, [vp, P(u32), u32, P(u32)]),
    "svo_mm_get_slots": (ctypes.c_int, [vp, u32, u32, P(svo_slot_info)]),
    "svo_mm_get_device_pools": (ctypes.c_int, [vp, P(svo_device_pools)]),
    "svo_mm_init_device_buffer": (ctypes.c_int, [vp, vp]),
    "svo_render": (ctypes.c_int, [vp, P(svo_camera), u32, P(svo_frame_out), ctypes.c_int, vp]),
    "svo_renderer_create": (ctypes.c_int, [ctypes.c_int, P(vp)]),
    "svo_renderer_destroy": (None, [vp]),
    "svo_renderer_set_config": (ctypes.c_int, [vp, P(svo_launch_config)]),
    "svo_renderer_get_config": (ctypes.c_int, [vp, P(svo_launch_config)]),
    "svo_render_views": (ctypes.c_int, [vp, P(svo_view), u32, ctypes.c_int, vp]),
    "svo_render_views_noalloc": (ctypes.c_int, [P(svo_view), u32, ctypes.c_int, P(svo_launch_config), vp]),
    "svo_render_views_timed": (ctypes.c_int, [vp, P(svo_view), u32, ctypes.c_int, vp, u32, P(f32)]),
    "svo_autotune": (ctypes.c_int, [vp, P(svo_view), u32, ctypes.c_int, vp, u32, P(svo_launch_config), P(f32)]),
    "svo_tune_cache_create": (ctypes.c_int, [P(vp)]),
    "svo_tune_cache_destroy": (None, [vp]),
    "svo_tune_cache_size": (u32, [vp]),
    "svo_tune_cache_load": (ctypes.c_int, [vp, ctypes.c_char_p]),
    "svo_tune_cache_save": (ctypes.c_int, [vp, ctypes.c_char_p]),
    "svo_tune_cache_key": (ctypes.c_int, [vp, P(svo_view), u32, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "svo_autotune_cached": (ctypes.c_int, [vp, vp, P(svo_view), u32, ctypes.c_int, vp, u32, P(svo_launch_config),
                                            P(f32), P(ctypes.c_int)]),
    "svo_batch_create": (ctypes.c_int, [vp, P(svo_view), u32, ctypes.c_int, vp, P(vp)]),
    "svo_batch_set_cameras": (ctypes.c_int, [vp, P(svo_camera), u32, vp]),
    "svo_batch_launch": (ctypes.c_int, [vp, vp]),
    "svo_batch_destroy": (None, [vp]),
    "svo_analysis_camera": (ctypes.c_int, [P(svo_camera), u32, f32, f32, P(svo_camera)]),
    "svo_mm_analyse": (ctypes.c_int, [vp, P(svo_camera), u32, f32, f32, ctypes.c_int, vp]),
    "svo_mm_apply_feedback": (ctypes.c_int, [vp, vp, P(svo_feedback_stats)]),
    "svo_mm_apply_records": (ctypes.c_int, [vp, vp, vp, u32, vp, vp, u32, P(svo_feedback_stats)]),
    "svo_mm_fetch_feedback": (ctypes.c_int, [vp, vp]),
    "svo_mm_read_feedback": (ctypes.c_int, [vp, vp, vp, vp, u32, P(u32), vp, vp, u32, P(u32)]),
    "svo_server_create": (ctypes.c_int, [vp, P(svo_server_config), P(vp)]),
    "svo_server_destroy": (None, [vp]),
    "svo_server_add_scene": (ctypes.c_int, [vp, vp, P(u32)]),
    "svo_server_submit": (ctypes.c_int, [vp, P(svo_request), P(u64)]),
    "svo_server_poll": (ctypes.c_int, [vp, P(svo_completion), u32, P(u32)]),
    "svo_server_pending": (ctypes.c_int, [vp, P(u32)]),
    "svo_server_replay": (ctypes.c_int, [vp, P(svo_request), u32, P(svo_completion), P(svo_server_stats)]),
    "svo_device_count": (ctypes.c_int, [P(ctypes.c_int)]),
    "svo_set_device": (ctypes.c_int, [ctypes.c_int]),
    "svo_device_malloc": (ctypes.c_int, [P(vp), ctypes.c_size_t]),
    "svo_device_free": (ctypes.c_int, [vp]),
    "svo_memcpy_h2d": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "svo_memcpy_d2h": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "svo_memset_device": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_size_t]),
    "svo_stream_create": (ctypes.c_int, [P(vp)]),
    "svo_stream_destroy": (ctypes.c_int, [vp]),
    "svo_stream_synchronize": (ctypes.c_int, [vp]),
    "svo_device_synchronize": (ctypes.c_int, []),
    "svo_render_timed": (ctypes.c_int, [vp, P(svo_camera), u32, P(svo_frame_out), ctypes.c_int, vp, u32, P(f32)]),
    "svo_debug_ray_dirs": (ctypes.c_int, [P(svo_camera), vp, vp]),
    "svo_debug_copy": (ctypes.c_int, [vp, vp, ctypes.c_size_t, vp]),
    "svo_debug_valu": (ctypes.c_int, [vp, u32, u32, vp]),
}


class SvoError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() (make -C libsvo_amd)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def check(status: int, where: str) -> None:
    if status != SVO_OK:
        raise SvoError(status, where, lib.svo_last_error().decode(errors="replace"))
