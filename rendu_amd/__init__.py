"""rendu_amd — MI355X (gfx950) implementation of Rendu's BVH raycaster and CPU path tracer.

The package is a thin host layer over librdu.so (C ABI in include/rdu.h): Raycaster mirrors
src/engine/raycaster/Raycaster.hpp and PathTracer mirrors src/apps/pathtracer/PathTracer.hpp.
All ray queries and rendering run in hand-written HIP kernels; there is no CPU fallback.
"""
from ._lib import RduError, lib  # noqa: F401  (fails loudly if librdu.so is missing)
from .camera import Camera  # noqa: F401
from .raycaster import Hit, Raycaster, make_rays  # noqa: F401

__all__ = ["Camera", "Hit", "Raycaster", "RduError", "make_rays"]
