"""Learned radiance cache (rdu_rcache_* in include/rdu.h).

No Rendu counterpart: an extension of the path tracer that maps a world position and an outgoing
direction to RGB radiance through a 16-level hash-grid encoding of the position (2^19 entries x 2
features per level), degree-3 spherical harmonics of the direction (16 features) and a 5 x 64 ReLU
MLP, evaluated and trained on the GPU with bf16 MFMA kernels (fp32 accumulation, fp32 master
weights, Adam on a relative L2 loss).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

MLP_PARAMS = 20480
TABLE_PARAMS = 16 * (1 << 19) * 2
LEVELS = 16
TABLE_SIZE = 1 << 19
# (rows, cols) of each MLP weight matrix, row-major [out][in], concatenated in this order.
LAYER_SHAPES = [(64, 48), (64, 64), (64, 64), (64, 64), (64, 64), (16, 64)]


def _as4(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, np.float32)
    if a.ndim != 2 or a.shape[1] not in (3, 4):
        raise ValueError("expected an (n, 3) or (n, 4) float array")
    if a.shape[1] == 3:
        a = np.concatenate([a, np.zeros((a.shape[0], 1), np.float32)], axis=1)
    return np.ascontiguousarray(a)


def make_queries(positions: np.ndarray, directions: np.ndarray | None = None) -> np.ndarray:
    """(n, 8) float32 query records: position xyz, 0, outgoing direction xyz, 0 (default +z)."""
    p = _as4(positions)
    d = _as4(directions) if directions is not None else np.tile(np.float32([0, 0, 1, 0]), (p.shape[0], 1))
    if d.shape[0] != p.shape[0]:
        raise ValueError("positions and directions differ in length")
    return np.ascontiguousarray(np.concatenate([p, d], axis=1))


class RadianceCache:
    """Owns one rdu_rcache. `scene_min`/`scene_max` bound the positions it encodes."""

    def __init__(self, scene_min, scene_max, seed: int = 0):
        lo = np.ascontiguousarray(scene_min, np.float32)
        hi = np.ascontiguousarray(scene_max, np.float32)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib.rdu_rcache_create(ctypes.byref(h), _lib.fptr(lo), _lib.fptr(hi), seed))
        self._h = h
        self.scene_min = lo.copy()
        self.scene_max = hi.copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib.lib is not None:  # the library may be gone at interpreter exit
            _lib.lib.rdu_rcache_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_learning_rate(self, lr: float) -> None:
        _lib.check(_lib.lib.rdu_rcache_set_learning_rate(self._h, float(lr)))

    def infer(self, positions: np.ndarray, directions: np.ndarray | None = None) -> np.ndarray:
        q = make_queries(positions, directions)
        out = np.zeros((q.shape[0], 4), np.float32)
        _lib.check(_lib.lib.rdu_rcache_infer(self._h, _lib.vptr(q), q.shape[0], _lib.vptr(out), 0, None))
        return out[:, :3]

    def infer_device(self, queries_ptr: int, n: int, out_ptr: int, stream: int = 0) -> None:
        _lib.check(_lib.lib.rdu_rcache_infer(self._h, ctypes.c_void_p(queries_ptr), n, ctypes.c_void_p(out_ptr),
                                             _lib.RDU_DEVICE_POINTERS, ctypes.c_void_p(stream)))

    def train(self, positions: np.ndarray, targets: np.ndarray, directions: np.ndarray | None = None,
              steps: int = 1) -> float:
        q, t = make_queries(positions, directions), _as4(targets)
        if q.shape[0] != t.shape[0]:
            raise ValueError("queries and targets differ in length")
        loss = ctypes.c_float(0.0)
        _lib.check(_lib.lib.rdu_rcache_train(self._h, _lib.vptr(q), _lib.vptr(t), q.shape[0], steps, 0, None,
                                             ctypes.byref(loss)))
        return float(loss.value)

    def train_device(self, queries_ptr: int, targets_ptr: int, n: int, steps: int = 1, stream: int = 0) -> None:
        _lib.check(_lib.lib.rdu_rcache_train(self._h, ctypes.c_void_p(queries_ptr), ctypes.c_void_p(targets_ptr), n,
                                             steps, _lib.RDU_DEVICE_POINTERS, ctypes.c_void_p(stream), None))

    def gradients(self, positions: np.ndarray, targets: np.ndarray, directions: np.ndarray | None = None,
                  tables: bool = True):
        q, t = make_queries(positions, directions), _as4(targets)
        gm = np.zeros(MLP_PARAMS, np.float32)
        gt = np.zeros(TABLE_PARAMS, np.float32) if tables else None
        _lib.check(_lib.lib.rdu_rcache_gradients(self._h, _lib.vptr(q), _lib.vptr(t), q.shape[0], _lib.fptr(gm),
                                                 _lib.fptr(gt) if tables else None))
        return gm, gt

    def get_params(self):
        mlp = np.zeros(MLP_PARAMS, np.float32)
        tables = np.zeros(TABLE_PARAMS, np.float32)
        res = np.zeros(LEVELS, np.int32)
        _lib.check(_lib.lib.rdu_rcache_get_params(self._h, _lib.fptr(mlp), _lib.fptr(tables),
                                                  res.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return mlp, tables, res

    def set_params(self, mlp: np.ndarray | None = None, tables: np.ndarray | None = None) -> None:
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        t = None if tables is None else np.ascontiguousarray(tables, np.float32)
        if m is not None and m.size != MLP_PARAMS:
            raise ValueError(f"mlp needs {MLP_PARAMS} floats")
        if t is not None and t.size != TABLE_PARAMS:
            raise ValueError(f"tables need {TABLE_PARAMS} floats")
        _lib.check(_lib.lib.rdu_rcache_set_params(self._h, None if m is None else _lib.fptr(m),
                                                  None if t is None else _lib.fptr(t)))


def split_layers(mlp: np.ndarray):
    """The six weight matrices of a flat MLP parameter vector."""
    out, off = [], 0
    for r, c in LAYER_SHAPES:
        out.append(mlp[off:off + r * c].reshape(r, c))
        off += r * c
    return out
