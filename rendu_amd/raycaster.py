"""Python mirror of Rendu's Raycaster (src/engine/raycaster/Raycaster.hpp:10-157) over librdu.

Same method names, argument meanings and defaults as the C++ class, plus batch entry points
that the GPU needs for throughput. Every query runs on the GPU through the C ABI.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import ctypes
import numpy as np

from . import _lib
from ._lib import HIT_DTYPE, RAY_DTYPE, check, fptr, lib, u8ptr, u32ptr, u64ptr, vptr

FLT_MAX = float(np.finfo(np.float32).max)


@dataclass
class Hit:
    """Raycaster::Hit (Raycaster.hpp:16-42). Miss: hit False, dist FLT_MAX, all else 0."""
    hit: bool = False
    dist: float = FLT_MAX
    u: float = 0.0
    v: float = 0.0
    w: float = 0.0
    localId: int = 0
    meshId: int = 0

    @staticmethod
    def from_record(rec) -> "Hit":
        if not rec["hit"]:
            return Hit()
        u = np.float32(rec["u"])
        v = np.float32(rec["v"])
        w = np.float32(np.float32(1.0) - u - v)  # Raycaster.cpp:10-12
        return Hit(True, float(rec["t"]), float(u), float(v), float(w), int(rec["local_id"]), int(rec["mesh_id"]))


def make_rays(origins, directions, mini=1e-4, maxi=1e8) -> np.ndarray:
    """Pack origins (n,3), directions (n,3) and scalar or per-ray mini/maxi into rdu_ray records."""
    o = np.asarray(origins, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(directions, dtype=np.float32).reshape(-1, 3)
    rays = np.empty(o.shape[0], dtype=RAY_DTYPE)
    rays["ox"], rays["oy"], rays["oz"] = o[:, 0], o[:, 1], o[:, 2]
    rays["dx"], rays["dy"], rays["dz"] = d[:, 0], d[:, 1], d[:, 2]
    rays["tmin"] = np.float32(mini) if np.isscalar(mini) else np.asarray(mini, np.float32)
    rays["tmax"] = np.float32(maxi) if np.isscalar(maxi) else np.asarray(maxi, np.float32)
    return rays


def _flags(reference_order: bool, incoherent: bool) -> int:
    return (_lib.RDU_REFERENCE_ORDER if reference_order else 0) | (_lib.RDU_INCOHERENT if incoherent else 0)


class Raycaster:
    """Raycaster (Raycaster.hpp:10). Not copyable, like the reference (Raycaster.hpp:100-113)."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib.rdu_raycaster_create(ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and not getattr(self, "_borrowed", False):
            lib.rdu_raycaster_destroy(h)
        self._h = None

    def __copy__(self):
        raise TypeError("Raycaster is not copyable (Raycaster.hpp:100-113)")

    __deepcopy__ = __copy__

    # -- setup -------------------------------------------------------------------------------------
    def addMesh(self, mesh, model: Optional[np.ndarray] = None) -> None:
        """Raycaster::addMesh(const Mesh&, const glm::mat4&) (Raycaster.cpp:14-52).

        mesh: any object with `positions` (n,3) and `indices` (3k,); model: 4x4 matrix in the usual
        row/column math layout (M @ [x,y,z,1]); it is handed to the library column-major as glm holds it.
        """
        pos = np.ascontiguousarray(np.asarray(mesh.positions, dtype=np.float32).reshape(-1, 3))
        idx = np.ascontiguousarray(np.asarray(mesh.indices, dtype=np.uint32).reshape(-1))
        m = None
        if model is not None:
            m = np.ascontiguousarray(np.asarray(model, dtype=np.float32).reshape(4, 4).T.reshape(-1))
        check(lib.rdu_raycaster_add_mesh(self._h, fptr(pos), pos.shape[0], u32ptr(idx), idx.shape[0],
                                         fptr(m) if m is not None else None))

    def updateHierarchy(self) -> None:
        """Raycaster::updateHierarchy() (Raycaster.cpp:54-142)."""
        check(lib.rdu_raycaster_update_hierarchy(self._h))

    def updateHierarchyGpu(self, stream: int = 0) -> None:
        """Build another tree on the GPU instead (rdu_raycaster_update_hierarchy_gpu): same closest
        hits except equal-t ties, no reference visiting order."""
        check(lib.rdu_raycaster_update_hierarchy_gpu(self._h, ctypes.c_void_p(stream) if stream else None))

    # -- per-ray API (reference signatures) ---------------------------------------------------------
    def set_per_ray_device(self, on_device: bool) -> None:
        """Per-ray calls as one-ray GPU launches (True) instead of the host walk of the uploaded
        tree (False, the default on a host-built tree): rdu_raycaster_set_per_ray_device."""
        check(lib.rdu_raycaster_set_per_ray_device(self._h, int(on_device)))

    def intersects(self, origin, direction, mini: float = 0.0001, maxi: float = 1e8) -> Hit:
        """Raycaster::intersects (Raycaster.hpp:65)."""
        o = np.asarray(origin, np.float32).reshape(3).copy()
        d = np.asarray(direction, np.float32).reshape(3).copy()
        out = np.zeros(1, dtype=HIT_DTYPE)
        check(lib.rdu_raycaster_intersect1(self._h, fptr(o), fptr(d), mini, maxi, vptr(out)))
        return Hit.from_record(out[0])

    def intersectsAny(self, origin, direction, mini: float = 0.0001, maxi: float = 1e8) -> bool:
        """Raycaster::intersectsAny (Raycaster.hpp:74)."""
        o = np.asarray(origin, np.float32).reshape(3).copy()
        d = np.asarray(direction, np.float32).reshape(3).copy()
        occ = ctypes.c_int(0)
        check(lib.rdu_raycaster_intersect_any1(self._h, fptr(o), fptr(d), mini, maxi, ctypes.byref(occ)))
        return bool(occ.value)

    def visible(self, p0, p1) -> bool:
        """Raycaster::visible (Raycaster.cpp:226-230)."""
        a = np.asarray(p0, np.float32).reshape(3).copy()
        b = np.asarray(p1, np.float32).reshape(3).copy()
        vis = ctypes.c_int(0)
        check(lib.rdu_raycaster_visible(self._h, fptr(a), fptr(b), ctypes.byref(vis)))
        return bool(vis.value)

    @staticmethod
    def interpolateAttribute(hit: Hit, mesh, attribute: np.ndarray) -> np.ndarray:
        """Raycaster::interpolateAttribute (Raycaster.hpp:90-97): w*A[i0] + u*A[i1] + v*A[i2]."""
        i0, i1, i2 = (int(mesh.indices[hit.localId + k]) for k in range(3))
        a = np.asarray(attribute, np.float32)
        w, u, v = np.float32(hit.w), np.float32(hit.u), np.float32(hit.v)
        return (w * a[i0] + u * a[i1]) + v * a[i2]

    # -- batch API ----------------------------------------------------------------------------------
    def intersect_batch(self, rays: np.ndarray, reference_order: bool = False, incoherent: bool = False) -> np.ndarray:
        """Closest hit of every rdu_ray record (host arrays, synchronous)."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.empty(rays.shape[0], dtype=HIT_DTYPE)
        flags = _flags(reference_order, incoherent)
        check(lib.rdu_raycaster_intersect(self._h, vptr(rays), rays.shape[0], vptr(hits), flags, None))
        return hits

    def intersect_any_batch(self, rays: np.ndarray, reference_order: bool = False,
                            incoherent: bool = False) -> np.ndarray:
        """Any-hit occlusion of every rdu_ray record, as a bool array."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.empty(rays.shape[0], dtype=np.uint8)
        flags = _flags(reference_order, incoherent)
        check(lib.rdu_raycaster_intersect_any(self._h, vptr(rays), rays.shape[0], vptr(occ), flags, None))
        return occ.astype(bool)

    def intersect_device(self, rays_ptr: int, n: int, out_ptr: int, stream: int = 0, any_hit: bool = False,
                         reference_order: bool = False, incoherent: bool = False) -> None:
        """Asynchronous query on device buffers (raw device addresses) on a HIP stream handle."""
        flags = _lib.RDU_DEVICE_POINTERS | _flags(reference_order, incoherent)
        fn = lib.rdu_raycaster_intersect_any if any_hit else lib.rdu_raycaster_intersect
        check(fn(self._h, ctypes.c_void_p(rays_ptr), n, ctypes.c_void_p(out_ptr), flags,
                 ctypes.c_void_p(stream) if stream else None))

    def autotune(self, rays=None, n: int = 0, rays_ptr: int = 0, any_hit: bool = False, incoherent: bool = False,
                 stream: int = 0) -> dict:
        """Pick the persistent traversal's launch parameters (refill threshold, ray tile, workgroups
        per CU) for this class of batch by timing candidates (rdu_raycaster_autotune); the winner
        also enters the process-wide tuning cache under the scene's statistics.
        Pass host `rays` (rdu_ray records) or a device pointer `rays_ptr` with its count `n`."""
        rep = _lib.TuneReport()
        flags = _flags(False, incoherent)
        if rays is not None:
            rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
            ptr, count = vptr(rays), rays.shape[0]
        else:
            ptr, count = ctypes.c_void_p(rays_ptr), n
            flags |= _lib.RDU_DEVICE_POINTERS
        check(lib.rdu_raycaster_autotune(self._h, ptr, count, int(any_hit), flags,
                                         ctypes.c_void_p(stream) if stream else None, ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in _lib.TuneReport._fields_}

    def tuning(self, n: int, any_hit: bool = False, incoherent: bool = False) -> dict:
        """Launch parameters a batch of n rays of this class would use now, and their source
        (0 defaults, 1 this raycaster's autotune, 2 the tuning cache)."""
        rep = _lib.TuneReport()
        check(lib.rdu_raycaster_tuning(self._h, n, int(any_hit), _flags(False, incoherent), ctypes.byref(rep)))
        return {k: getattr(rep, k) for k in ("refill_min", "chunk", "blocks_per_cu", "source")}

    def set_tuning(self, refill_min: int = 0, chunk: int = 0, blocks_per_cu: int = 0, any_hit: bool = False,
                   incoherent: bool = False) -> None:
        """Fix this class's launch parameters without timing (0 = default; rdu_raycaster_set_tuning)."""
        rep = _lib.TuneReport(refill_min, 0.0, 0.0, 0, chunk, blocks_per_cu, 0)
        check(lib.rdu_raycaster_set_tuning(self._h, int(any_hit), _flags(False, incoherent), ctypes.byref(rep)))

    def set_kernel_timing(self, enable: bool) -> None:
        """Record HIP events around each traversal kernel launch (rdu_raycaster_set_kernel_timing)."""
        check(lib.rdu_raycaster_set_kernel_timing(self._h, int(enable)))

    def kernel_time(self):
        """(summed traversal-kernel milliseconds, launches) since the last call; waits for them."""
        ms, k = ctypes.c_double(0.0), ctypes.c_uint64(0)
        check(lib.rdu_raycaster_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(k)))
        return ms.value, k.value

    def trace_stats(self, rays: np.ndarray, any_hit: bool = False, reference_order: bool = False):
        """(internal-node expansions, triangle tests) summed over a batch, counted on the GPU."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        flags = _lib.RDU_REFERENCE_ORDER if reference_order else 0
        check(lib.rdu_raycaster_trace_stats(self._h, vptr(rays), rays.shape[0], int(any_hit), flags,
                                            ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    # -- introspection ------------------------------------------------------------------------------
    def info(self) -> dict:
        inf = _lib.Info()
        check(lib.rdu_raycaster_info(self._h, ctypes.byref(inf)))
        return {k: int(getattr(inf, k)) for k, _ in _lib.Info._fields_}

    def export_nodes(self):
        """Rendu's _hierarchy after updateHierarchy: (boxes (n,6), left, right, leaf)."""
        n = self.info()["nodes"]
        boxes = np.empty((n, 6), np.float32)
        left = np.empty(n, np.uint64)
        right = np.empty(n, np.uint64)
        leaf = np.empty(n, np.uint8)
        check(lib.rdu_raycaster_export_nodes(self._h, fptr(boxes), u64ptr(left), u64ptr(right), u8ptr(leaf)))
        return boxes, left, right, leaf.astype(bool)

    def export_triangles(self):
        """Rendu's _triangles order after the build: (mesh_id, local_id) arrays."""
        n = self.info()["triangles"]
        mesh = np.empty(n, np.uint32)
        local = np.empty(n, np.uint32)
        check(lib.rdu_raycaster_export_triangles(self._h, u32ptr(mesh), u32ptr(local)))
        return mesh, local
