"""Python mirror of Rendu's PathTracer (src/apps/pathtracer/PathTracer.hpp:10-76) over librdu.

PathTracer(scene) builds the raycaster over the scene's objects (PathTracer.cpp:11-21);
render(camera, samples, depth, image) fills a caller-owned float image (PathTracer.cpp:159-299):
like Rendu's Image it is accumulated into and divided by the sample count, so pass zeros.
Rendering runs entirely in the library's wavefront kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import scene as S
from ._lib import TuneReport, check, fptr, lib


class _BorrowedRaycaster:
    """The raycaster a committed scene owns (PathTracer::raycaster()), usable for queries."""

    def __init__(self, handle, owner):
        from .raycaster import Raycaster
        self._owner = owner  # keeps the scene alive
        rc = Raycaster.__new__(Raycaster)
        rc._h = ctypes.c_void_p(handle)
        rc._borrowed = True
        self.rc = rc

    def __getattr__(self, name):
        return getattr(self.rc, name)

    def __del__(self):
        # Never destroy: the scene owns the handle.
        if hasattr(self, "rc"):
            self.rc._h = None


class PathTracer:
    def __init__(self, scene: S.Scene, gpu_build: bool = False, instancing: bool = False):
        """gpu_build: build the tree on the GPU (rdu_scene_set_gpu_build) instead of Rendu's host build.
        instancing: one tree per mesh shared by the objects that use it, under a top level over the
        objects (rdu_scene_set_instancing; an extension: hits differ from Rendu's in the last bits)."""
        h = ctypes.c_void_p()
        check(lib.rdu_scene_create(ctypes.byref(h)))
        self._h = h
        if gpu_build:
            check(lib.rdu_scene_set_gpu_build(h, 1))
        if instancing:
            check(lib.rdu_scene_set_instancing(h, 1))
        keep, tex, meshes, mats, objs, lights, bg = scene.descriptors()
        i = ctypes.c_int32()
        for t in tex:
            check(lib.rdu_scene_add_texture(h, ctypes.byref(t), ctypes.byref(i)))
        for m in meshes:
            check(lib.rdu_scene_add_mesh(h, ctypes.byref(m), ctypes.byref(i)))
        for m in mats:
            check(lib.rdu_scene_add_material(h, ctypes.byref(m), ctypes.byref(i)))
        for o in objs:
            check(lib.rdu_scene_add_object(h, ctypes.byref(o)))
        for l in lights:
            check(lib.rdu_scene_add_light(h, ctypes.byref(l)))
        check(lib.rdu_scene_set_background(h, ctypes.byref(bg)))
        check(lib.rdu_scene_commit(h))
        del keep

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.rdu_scene_destroy(h)
            self._h = None

    def raycaster(self):
        """PathTracer::raycaster() (PathTracer.hpp:29). Instanced scenes have none."""
        p = lib.rdu_scene_raycaster(self._h)
        if not p:
            raise RuntimeError("an instanced scene has no raycaster of its own")
        return _BorrowedRaycaster(p, self)

    def render(self, camera, samples: int, depth: int, image: np.ndarray, seed: int = 0,
               estimator: int = S.ESTIMATOR_REFERENCE, rr_start: int = 3, max_paths: int = 0,
               passes_in_flight: int = 0) -> dict:
        """PathTracer::render(camera, samples, depth, image). image: (h, w, c >= 3) float32.
        max_paths / passes_in_flight: paths per pass and passes kept in flight at once (0 = the
        library's defaults); the image does not depend on them."""
        if image.dtype != np.float32 or not image.flags.c_contiguous or image.ndim != 3:
            raise ValueError("image must be a C-contiguous (height, width, components) float32 array")
        h, w, c = image.shape
        opts = S.RenderOpts(seed, estimator, rr_start, max_paths, passes_in_flight)
        cam = S.camera_desc(camera)
        check(lib.rdu_pathtracer_render(self._h, ctypes.byref(cam), samples, depth, fptr(image), w, h, c,
                                        ctypes.byref(opts)))
        return self.last_stats()

    def set_graph(self, enable: bool) -> None:
        """Replay repeated full passes from a captured HIP graph (rdu_scene_set_graph)."""
        check(lib.rdu_scene_set_graph(self._h, int(enable)))

    def autotune(self, camera, samples: int, depth: int, width: int, height: int, seed: int = 0,
                 estimator: int = S.ESTIMATOR_REFERENCE, rr_start: int = 3, max_paths: int = 0) -> dict:
        """Time one pass of this render configuration per candidate of the traversal launch
        parameters and keep the fastest for this scene's renders (rdu_scene_autotune)."""
        rep = TuneReport()
        opts = S.RenderOpts(seed, estimator, rr_start, max_paths)
        cam = S.camera_desc(camera)
        check(lib.rdu_scene_autotune(self._h, ctypes.byref(cam), samples, depth, width, height, ctypes.byref(opts),
                                     ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in TuneReport._fields_}

    def render_cached(self, cache, camera, samples: int, depth: int, image: np.ndarray, seed: int = 0,
                      estimator: int = S.ESTIMATOR_EXTENDED, rr_start: int = 3, max_paths: int = 0,
                      terminate_bounce: int = 0, train_steps: int = 0, train_batch: int = 0, max_bias: float = 0.0,
                      min_count: float = 0.0, frozen: bool = False, always_trust: bool = False,
                      from_primary: bool = False) -> dict:
        """render() with a rendu_amd.rcache.RadianceCache (rdu_pathtracer_render_cached): paths end in
        the cache where it is trusted, one pixel in 16 trains it. Zero options take the defaults."""
        if image.dtype != np.float32 or not image.flags.c_contiguous or image.ndim != 3:
            raise ValueError("image must be a C-contiguous (height, width, components) float32 array")
        h, w, c = image.shape
        opts = S.RenderOpts(seed, estimator, rr_start, max_paths)
        copts = S.CacheOpts(terminate_bounce, train_steps, train_batch, max_bias, min_count, int(frozen),
                            int(always_trust), int(from_primary))
        cam = S.camera_desc(camera)
        check(lib.rdu_pathtracer_render_cached(self._h, cache.handle, ctypes.byref(cam), samples, depth, fptr(image),
                                               w, h, c, ctypes.byref(opts), ctypes.byref(copts)))
        out = self.last_stats()
        st = S.CacheStats()
        check(lib.rdu_scene_last_cache_stats(self._h, ctypes.byref(st)))
        out.update({k: getattr(st, k) for k, _ in S.CacheStats._fields_})
        return out

    def last_cache_pairs(self):
        """(queries (n, 8), targets (n, 4)) of the last pass of the last cached render."""
        n = ctypes.c_size_t(0)
        check(lib.rdu_scene_last_cache_pairs(self._h, None, None, 0, ctypes.byref(n)))
        q = np.zeros((n.value, 8), np.float32)
        t = np.zeros((n.value, 4), np.float32)
        if n.value:
            check(lib.rdu_scene_last_cache_pairs(self._h, q.ctypes.data_as(ctypes.c_void_p),
                                                 t.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n)))
        return q, t

    def last_stats(self) -> dict:
        st = S.RenderStats()
        check(lib.rdu_scene_last_render_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in S.RenderStats._fields_}
