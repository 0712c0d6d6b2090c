"""ORACLE — test infrastructure only.

CPU restatement of the path-tracer app's output conversion, in float32 as the reference computes it:
the tonemap c -> 1 - exp(-c) of src/apps/pathtracer/main.cpp:109-116 on the colour channels, then
Image::saveAsLDR (src/engine/resources/Image.cpp:138-151): v -> pow(v, 1/2.2) except channel 3,
byte = (unsigned char)min(255, max(0, 255 v)) with std::max/std::min's NaN behaviour (NaN -> 0),
alpha byte 255 under IGNORE_ALPHA. Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np


def to_ldr(image: np.ndarray, tonemap: bool = True, srgb: bool = True, ignore_alpha: bool = True) -> np.ndarray:
    img = np.asarray(image, np.float32)
    a = img if img.ndim == 3 else img[:, :, None]
    comps = a.shape[2]
    out = np.zeros(a.shape, np.uint8)
    with np.errstate(invalid="ignore", over="ignore"):
        for c in range(comps):
            v = a[:, :, c].copy()
            if tonemap and comps >= 3 and c < 3:
                v = np.float32(1.0) - np.exp(-v)
            if srgb and c != 3:
                v = np.power(v, np.float32(1.0 / 2.2))
            s = np.float32(255.0) * v
            s = np.where(np.float32(0.0) < s, s, np.float32(0.0))      # std::max(0.0f, s)
            s = np.where(s < np.float32(255.0), s, np.float32(255.0))  # std::min(255.0f, s)
            q = s.astype(np.uint8)
            if c == 3 and ignore_alpha:
                q[:] = 255
            out[:, :, c] = q
    return out.reshape(img.shape)
