"""ORACLE — test infrastructure only.

CPU reference of the learned radiance cache (rdu_rcache_*), which has no Rendu counterpart: the
path tracer's cache extension follows the published neural radiance cache (Mueller, Rousselle,
Novak, Keller, "Real-time Neural Radiance Caching for Path Tracing", SIGGRAPH 2021) with the
multiresolution hash encoding of Mueller, Evans, Schied, Keller ("Instant Neural Graphics
Primitives", SIGGRAPH 2022, section 3: spatial hash with primes 1, 2654435761, 805459861, trilinear
interpolation of the 8 cell corners). Only tests/ may import this module.

The forward model restates what the GPU computes, including its rounding points: table features
read as fp16, encoding and every hidden activation rounded to bf16 (round to nearest even), weights
as bf16, products accumulated in fp32. Gradients come from torch autograd on that function (the
roundings pass gradients straight through), independent of the GPU's hand-written backward pass.

The outgoing direction enters as the 16 real spherical harmonics of degrees 0..3 (inputs 32..47),
as in the neural radiance cache's direction encoding.

Pinning (tests/test_oracle_rcache.py): trilinear interpolation reproduces a linear function of the
corner coordinates exactly on a dense level; the hash agrees with arbitrary-precision integer
arithmetic; the harmonics are orthonormal on the sphere (quadrature); autograd agrees with central
finite differences in float64; the relative L2 loss treats its denominator as a constant (finite
differences with the denominator frozen).
"""
from __future__ import annotations

import numpy as np
import torch

LEVELS = 16
FEATURES = 2
TABLE = 1 << 19
PRIMES = (1, 2654435761, 805459861)
LAYER_SHAPES = [(64, 48), (64, 64), (64, 64), (64, 64), (64, 64), (16, 64)]

# Real spherical harmonics, degrees 0..3, with the Condon-Shortley phase: the standard normalised
# constants sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!) times the usual factors (for example Sloan,
# "Stupid Spherical Harmonics Tricks", 2008, appendix A2).
SH_C = {
    "c0": 0.28209479177387814, "c1": 0.4886025119029199, "c2a": 1.0925484305920792,
    "c2b": 0.31539156525252005, "c2c": 0.5462742152960396, "c3a": 0.5900435899266435,
    "c3b": 2.890611442640554, "c3c": 0.4570457994644658, "c3d": 0.3731763325901154,
    "c3e": 1.445305721320277,
}


def unit_dirs(directions: np.ndarray) -> np.ndarray:
    """Directions normalised as the GPU does (zero or non-finite length -> +z), float32."""
    d = np.asarray(directions, np.float32)[:, :3]
    with np.errstate(invalid="ignore", over="ignore", divide="ignore"):
        l2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        ok = (l2 > 0) & (l2 < np.float32(3.0e38))
        inv = np.float32(1.0) / np.sqrt(np.where(ok, l2, np.float32(1.0)))
        out = d * inv[:, None]
    out[~ok] = np.float32([0.0, 0.0, 1.0])
    return out.astype(np.float32)


def sh_features(directions: np.ndarray) -> np.ndarray:
    """(n, 16) float32 SH values of the unit directions, in the GPU's order and operation order."""
    d = unit_dirs(directions)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    f = lambda v: np.float32(v)
    xx, yy, zz = x * x, y * y, z * z
    c = {k: f(v) for k, v in SH_C.items()}
    cols = [
        np.full_like(x, c["c0"]), -c["c1"] * y, c["c1"] * z, -c["c1"] * x,
        c["c2a"] * (x * y), -c["c2a"] * (y * z), c["c2b"] * (f(3) * zz - f(1)), -c["c2a"] * (x * z),
        c["c2c"] * (xx - yy), -c["c3a"] * (y * (f(3) * xx - yy)), c["c3b"] * ((x * y) * z),
        -c["c3c"] * (y * (f(5) * zz - f(1))), c["c3d"] * (z * (f(5) * zz - f(3))),
        -c["c3c"] * (x * (f(5) * zz - f(1))), c["c3e"] * (z * (xx - yy)), -c["c3a"] * (x * (xx - f(3) * yy)),
    ]
    return np.stack(cols, axis=1).astype(np.float32)


def level_resolutions() -> np.ndarray:
    """16 * b^l, b = (2048 / 16)^(1/15), floored (geometric progression of Instant-NGP eq. 2-3)."""
    b = np.exp((np.log(2048.0) - np.log(16.0)) / 15.0)
    return np.array([int(np.floor(16.0 * b ** l)) for l in range(LEVELS)], np.int64)


def dense_levels(res: np.ndarray) -> np.ndarray:
    return ((res + 1).astype(np.float64) ** 3 <= TABLE).astype(np.int64)


def normalise(positions: np.ndarray, lo, hi) -> np.ndarray:
    lo = np.asarray(lo, np.float32)
    ext = np.asarray(hi, np.float32) - lo
    inv = np.where(ext > 0, np.float32(1.0) / np.where(ext > 0, ext, 1), np.float32(1.0)).astype(np.float32)
    p = (positions[:, :3].astype(np.float32) - lo) * inv
    return np.where(p > 0, np.where(p < 1, p, 1), 0).astype(np.float32)  # NaN -> 0


def corner_indices(p: np.ndarray, res: int, dense: bool):
    """(n, 8) table indices and fp32 trilinear weights; corner c has offsets (c&1, c>>1&1, c>>2&1)."""
    x = p * np.float32(res)
    ci = np.clip(np.floor(x).astype(np.int64), 0, res - 1)
    fr = (x - ci.astype(np.float32)).astype(np.float32)
    idx = np.zeros((p.shape[0], 8), np.int64)
    w = np.zeros((p.shape[0], 8), np.float32)
    for c in range(8):
        d = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
        cc = (ci + d).astype(np.uint64)
        ws = [np.where(d[a] == 1, fr[:, a], np.float32(1.0) - fr[:, a]).astype(np.float32) for a in range(3)]
        w[:, c] = (ws[0] * ws[1]) * ws[2]
        if dense:
            r1 = np.uint64(res + 1)
            idx[:, c] = (cc[:, 0] + cc[:, 1] * r1 + cc[:, 2] * r1 * r1).astype(np.int64)
        else:
            m = np.uint64(0xffffffff)
            h = ((cc[:, 0] * np.uint64(PRIMES[0])) & m) ^ ((cc[:, 1] * np.uint64(PRIMES[1])) & m) ^ \
                ((cc[:, 2] * np.uint64(PRIMES[2])) & m)
            idx[:, c] = (h & np.uint64(TABLE - 1)).astype(np.int64)
    return idx, w


def hash_int(cx: int, cy: int, cz: int) -> int:
    """The spatial hash in arbitrary-precision integers (for pinning the vectorised form)."""
    m = (1 << 32) - 1
    return (((cx * PRIMES[0]) & m) ^ ((cy * PRIMES[1]) & m) ^ ((cz * PRIMES[2]) & m)) & (TABLE - 1)


def bf16(x: torch.Tensor) -> torch.Tensor:
    """Round to bf16 (nearest even) with a straight-through gradient."""
    return x + (x.to(torch.bfloat16).to(x.dtype) - x).detach()


class CacheReference:
    """Forward model and autograd gradients of the cache for given fp32 parameters."""

    def __init__(self, mlp: np.ndarray, tables: np.ndarray, res: np.ndarray, lo, hi, round_like_gpu: bool = True,
                 dtype=torch.float32):
        self.res = np.asarray(res, np.int64)
        self.dense = dense_levels(self.res)
        self.lo, self.hi = lo, hi
        self.round = round_like_gpu
        self.dtype = dtype
        mlp = np.asarray(mlp, np.float32)
        ws, off = [], 0
        for r, c in LAYER_SHAPES:
            w = torch.tensor(mlp[off:off + r * c].reshape(r, c), dtype=dtype)
            if round_like_gpu:
                w = w.to(torch.bfloat16).to(dtype)
            ws.append(w.requires_grad_(True))
            off += r * c
        self.w = ws
        t = torch.tensor(np.asarray(tables, np.float32).reshape(LEVELS, TABLE, FEATURES), dtype=dtype)
        if round_like_gpu:
            t = t.to(torch.float16).to(dtype)
        self.t = t.requires_grad_(True)

    def encode(self, positions: np.ndarray) -> torch.Tensor:
        p = normalise(positions, self.lo, self.hi)
        feats = []
        for l in range(LEVELS):
            idx, w = corner_indices(p, int(self.res[l]), bool(self.dense[l]))
            g = self.t[l][torch.from_numpy(idx)]                    # (n, 8, 2)
            wt = torch.from_numpy(w).to(self.dtype)                  # (n, 8)
            acc = torch.zeros(p.shape[0], FEATURES, dtype=self.dtype)
            for c in range(8):                                       # corner order of the GPU
                acc = acc + wt[:, c:c + 1] * g[:, c, :]
            feats.append(acc)
        return torch.cat(feats, dim=1)                               # (n, 32): level l -> 2l, 2l+1

    def forward(self, positions: np.ndarray, directions: np.ndarray | None = None) -> torch.Tensor:
        if directions is None:
            directions = np.tile(np.float32([0, 0, 1]), (positions.shape[0], 1))
        sh = torch.from_numpy(sh_features(directions)).to(self.dtype)
        x = torch.cat([self.encode(positions), sh], dim=1)          # (n, 48)
        if self.round:
            x = bf16(x)
        for l in range(5):
            x = torch.relu(x @ self.w[l].t())
            if self.round:
                x = bf16(x)
        return (x @ self.w[5].t())[:, :3]

    def loss(self, positions: np.ndarray, targets: np.ndarray, directions: np.ndarray | None = None) -> torch.Tensor:
        """Mean over samples and channels of (y - t)^2 / (sg(y)^2 + 0.01)."""
        y = self.forward(positions, directions)
        t = torch.tensor(np.asarray(targets, np.float32)[:, :3], dtype=self.dtype)
        den = (y.detach() ** 2 + 0.01)
        return ((y - t) ** 2 / den).mean()

    def gradients(self, positions: np.ndarray, targets: np.ndarray, directions: np.ndarray | None = None):
        for w in self.w:
            w.grad = None
        self.t.grad = None
        L = self.loss(positions, targets, directions)
        L.backward()
        mlp = np.concatenate([w.grad.detach().numpy().reshape(-1) for w in self.w]).astype(np.float32)
        tab = self.t.grad.detach().numpy().reshape(-1).astype(np.float32)
        return float(L.detach()), mlp, tab
