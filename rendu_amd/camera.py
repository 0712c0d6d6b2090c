"""Camera (src/engine/input/Camera.hpp) restricted to what the path tracer reads: the pose and
pixelShifts (Camera.cpp:10-18, 42-49), evaluated in float32 in glm's operation order."""
from __future__ import annotations

import numpy as np

F32 = np.float32


def _dot(a, b):
    t = a * b
    return (t[..., 0] + t[..., 1]) + t[..., 2]


def _normalize(v):
    v = np.asarray(v, F32)
    return v * (F32(1.0) / np.sqrt(_dot(v, v)))


def _cross(x, y):
    return np.array([x[1] * y[2] - y[1] * x[2], x[2] * y[0] - y[2] * x[0], x[0] * y[1] - y[0] * x[1]], F32)


class Camera:
    def __init__(self):
        self.position = np.array([0.0, 0.0, 1.0], F32)
        self.center = np.array([0.0, 0.0, 0.0], F32)
        self.up = np.array([0.0, 1.0, 0.0], F32)
        self.up_in = np.array([0.0, 1.0, 0.0], F32)
        self.right = np.array([1.0, 0.0, 0.0], F32)
        self.fov = F32(1.3)  # Camera.hpp:172
        self.ratio = F32(1.0)  # Camera.hpp:173

    def pose(self, position, center, up):
        """Camera::pose (Camera.cpp:10-18)."""
        self.position = np.asarray(position, F32).copy()
        self.center = np.asarray(center, F32).copy()
        self.up_in = np.asarray(up, F32).copy()  # the pose input, handed to the library as given
        u = _normalize(up)
        view = _normalize(self.center - self.position)
        self.right = _cross(view, u)
        self.up = _cross(self.right, view)

    def projection(self, ratio=None, fov=None):
        if ratio is not None:
            self.ratio = F32(ratio)
        if fov is not None:
            self.fov = F32(fov)

    def pixelShifts(self):
        """Camera::pixelShifts (Camera.cpp:42-49): (corner, dx, dy)."""
        hs = F32(np.tan(F32(0.5) * self.fov))
        ws = F32(self.ratio * hs)
        dv = self.center - self.position
        dist = F32(np.sqrt(_dot(dv, dv)))
        corner = self.center + dist * ((-ws) * self.right + hs * self.up)
        dx = F32(F32(F32(2.0) * ws) * dist) * self.right
        dy = F32(F32(F32(-2.0) * hs) * dist) * self.up
        return corner.astype(F32), dx.astype(F32), dy.astype(F32)
