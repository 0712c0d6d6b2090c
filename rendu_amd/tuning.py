"""Process-wide tuning cache of the persistent traversal (rdu_tune_cache_* in include/rdu.h).

Autotuning (Raycaster.autotune, PathTracer.autotune) stores its winners here under the scene's
statistics: triangle count, 4-wide tree depth and mesh count (power-of-two buckets where counts),
batch size bucket and the kind of query. Saving the cache to a file and loading it in another
process lets later runs on scenes of the same class start tuned without timing anything.
No counterpart in Rendu, whose CPU traversal has no launch parameters.
"""
from __future__ import annotations

import ctypes

from ._lib import check, lib


def save(path: str) -> None:
    """Write every entry to a text file."""
    check(lib.rdu_tune_cache_save(str(path).encode()))


def load(path: str) -> None:
    """Merge a file written by save() (its entries win); a malformed file changes nothing."""
    check(lib.rdu_tune_cache_load(str(path).encode()))


def clear() -> None:
    check(lib.rdu_tune_cache_clear())


def size() -> int:
    n = ctypes.c_size_t(0)
    check(lib.rdu_tune_cache_size(ctypes.byref(n)))
    return int(n.value)
