"""ctypes binding of librdu.so, the C ABI declared in include/rdu.h.

The library is loaded from this package directory. There is no fallback: if librdu.so is absent
the import fails with the command that builds it, and queries without a HIP device fail with the
library's own RDU_ERR_HIP message.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librdu.so")

RDU_OK = 0
RDU_ERR_INVALID = -1
RDU_ERR_STATE = -2
RDU_ERR_HIP = -3
RDU_ERR_ALLOC = -4
RDU_ERR_UNSUPPORTED = -5

RDU_DEVICE_POINTERS = 1
RDU_REFERENCE_ORDER = 2
RDU_INCOHERENT = 4

RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"),
                      ("hit", "<u4"), ("mesh_id", "<u4"), ("local_id", "<u4")])


class RduError(RuntimeError):
    """A non-zero status returned by librdu, with rdu_last_error()'s message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"rdu error {code}: {message}")
        self.code = code
        self.message = message


class _Info(ctypes.Structure):
    _fields_ = [("meshes", c_uint64), ("triangles", c_uint64), ("vertices", c_uint64), ("nodes", c_uint64),
                ("internal_nodes", c_uint64), ("max_depth", c_uint64), ("device_bytes", c_uint64),
                ("wide_nodes", c_uint64), ("wide_depth", c_uint64), ("binary_bytes", c_uint64),
                ("wide_bytes", c_uint64)]


class TuneReport(ctypes.Structure):
    _fields_ = [("refill_min", c_uint32), ("best_ms", c_float), ("default_ms", c_float), ("candidates", c_uint32),
                ("chunk", c_uint32), ("blocks_per_cu", c_uint32), ("source", c_uint32)]


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "(or `make -C rendu_amd`) first")
    lib = ctypes.CDLL(LIB_PATH)
    fp = POINTER(c_float)
    sig = {
        "rdu_raycaster_create": (c_int, [POINTER(c_void_p)]),
        "rdu_raycaster_destroy": (None, [c_void_p]),
        "rdu_raycaster_add_mesh": (c_int, [c_void_p, fp, c_size_t, POINTER(c_uint32), c_size_t, fp]),
        "rdu_raycaster_update_hierarchy": (c_int, [c_void_p]),
        "rdu_raycaster_update_hierarchy_gpu": (c_int, [c_void_p, c_void_p]),
        "rdu_scene_set_gpu_build": (c_int, [c_void_p, c_int]),
        "rdu_scene_set_instancing": (c_int, [c_void_p, c_int]),
        "rdu_scene_set_graph": (c_int, [c_void_p, c_int]),
        "rdu_scene_autotune": (c_int, [c_void_p, c_void_p, c_size_t, c_size_t, c_uint32, c_uint32, c_void_p, c_void_p]),
        "rdu_raycaster_intersect": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_uint32, c_void_p]),
        "rdu_raycaster_intersect_any": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_uint32, c_void_p]),
        "rdu_raycaster_intersect_noalloc": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_uint32, c_void_p]),
        "rdu_raycaster_set_per_ray_device": (c_int, [c_void_p, c_int]),
        "rdu_raycaster_intersect1": (c_int, [c_void_p, fp, fp, c_float, c_float, c_void_p]),
        "rdu_raycaster_intersect_any1": (c_int, [c_void_p, fp, fp, c_float, c_float, POINTER(c_int)]),
        "rdu_raycaster_visible": (c_int, [c_void_p, fp, fp, POINTER(c_int)]),
        "rdu_raycaster_info": (c_int, [c_void_p, POINTER(_Info)]),
        "rdu_raycaster_export_nodes": (c_int, [c_void_p, fp, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint8)]),
        "rdu_raycaster_export_triangles": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
        "rdu_raycaster_trace_stats": (c_int, [c_void_p, c_void_p, c_size_t, c_int, c_uint32,
                                              POINTER(c_uint64), POINTER(c_uint64)]),
        "rdu_raycaster_autotune": (c_int, [c_void_p, c_void_p, c_size_t, c_int, c_uint32, c_void_p, c_void_p]),
        "rdu_raycaster_tuning": (c_int, [c_void_p, c_size_t, c_int, c_uint32, c_void_p]),
        "rdu_raycaster_set_tuning": (c_int, [c_void_p, c_int, c_uint32, c_void_p]),
        "rdu_tune_cache_save": (c_int, [c_char_p]),
        "rdu_tune_cache_load": (c_int, [c_char_p]),
        "rdu_tune_cache_clear": (c_int, []),
        "rdu_tune_cache_size": (c_int, [POINTER(c_size_t)]),
        "rdu_raycaster_set_kernel_timing": (c_int, [c_void_p, c_int]),
        "rdu_raycaster_kernel_time": (c_int, [c_void_p, POINTER(ctypes.c_double), POINTER(c_uint64)]),
        "rdu_scene_create": (c_int, [POINTER(c_void_p)]),
        "rdu_scene_destroy": (None, [c_void_p]),
        "rdu_scene_add_texture": (c_int, [c_void_p, c_void_p, POINTER(ctypes.c_int32)]),
        "rdu_scene_add_mesh": (c_int, [c_void_p, c_void_p, POINTER(ctypes.c_int32)]),
        "rdu_scene_add_material": (c_int, [c_void_p, c_void_p, POINTER(ctypes.c_int32)]),
        "rdu_scene_add_object": (c_int, [c_void_p, c_void_p]),
        "rdu_scene_add_light": (c_int, [c_void_p, c_void_p]),
        "rdu_scene_set_background": (c_int, [c_void_p, c_void_p]),
        "rdu_scene_commit": (c_int, [c_void_p]),
        "rdu_scene_raycaster": (c_void_p, [c_void_p]),
        "rdu_pathtracer_render": (c_int, [c_void_p, c_void_p, c_size_t, c_size_t, POINTER(c_float), c_uint32, c_uint32,
                                          c_uint32, c_void_p]),
        "rdu_scene_last_render_stats": (c_int, [c_void_p, c_void_p]),
        "rdu_rcache_create": (c_int, [POINTER(c_void_p), fp, fp, c_uint32]),
        "rdu_rcache_destroy": (None, [c_void_p]),
        "rdu_rcache_set_learning_rate": (c_int, [c_void_p, c_float]),
        "rdu_rcache_infer": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_uint32, c_void_p]),
        "rdu_rcache_train": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_uint32, c_uint32, c_void_p, fp]),
        "rdu_rcache_gradients": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, fp, fp]),
        "rdu_rcache_get_params": (c_int, [c_void_p, fp, fp, POINTER(ctypes.c_int32)]),
        "rdu_rcache_set_params": (c_int, [c_void_p, fp, fp]),
        "rdu_pathtracer_render_cached": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, POINTER(c_float),
                                                 c_uint32, c_uint32, c_uint32, c_void_p, c_void_p]),
        "rdu_scene_last_cache_stats": (c_int, [c_void_p, c_void_p]),
        "rdu_scene_last_cache_pairs": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
        "rdu_pathtracer_render_batch": (c_int, [c_void_p, c_size_t, c_uint32, c_void_p, c_void_p]),
        "rdu_mesh_tangent_frames": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p,
                                            c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
        "rdu_mesh_corner_tangents": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p]),
        "rdu_last_error": (c_char_p, []),
        "rdu_version": (c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(status: int) -> None:
    if status != RDU_OK:
        raise RduError(status, lib.rdu_last_error().decode("utf-8", "replace"))


def fptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_float))


def u32ptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_uint32))


def u64ptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_uint64))


def u8ptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_uint8))


def vptr(a: np.ndarray):
    return c_void_p(a.ctypes.data)


Info = _Info
