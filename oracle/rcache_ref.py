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

Pinning (tests/test_oracle_rcache.py): trilinear interpolation reproduces a linear function of the
corner coordinates exactly on a dense level; the hash agrees with arbitrary-precision integer
arithmetic; autograd agrees with central finite differences in float64; the relative L2 loss
treats its denominator as a constant (finite differences with the denominator frozen).
"""
from __future__ import annotations

import numpy as np
import torch

LEVELS = 16
FEATURES = 2
TABLE = 1 << 19
PRIMES = (1, 2654435761, 805459861)
LAYER_SHAPES = [(64, 32), (64, 64), (64, 64), (64, 64), (64, 64), (16, 64)]


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

    def forward(self, positions: np.ndarray) -> torch.Tensor:
        x = self.encode(positions)
        if self.round:
            x = bf16(x)
        for l in range(5):
            x = torch.relu(x @ self.w[l].t())
            if self.round:
                x = bf16(x)
        return (x @ self.w[5].t())[:, :3]

    def loss(self, positions: np.ndarray, targets: np.ndarray) -> torch.Tensor:
        """Mean over samples and channels of (y - t)^2 / (sg(y)^2 + 0.01)."""
        y = self.forward(positions)
        t = torch.tensor(np.asarray(targets, np.float32)[:, :3], dtype=self.dtype)
        den = (y.detach() ** 2 + 0.01)
        return ((y - t) ** 2 / den).mean()

    def gradients(self, positions: np.ndarray, targets: np.ndarray):
        for w in self.w:
            w.grad = None
        self.t.grad = None
        L = self.loss(positions, targets)
        L.backward()
        mlp = np.concatenate([w.grad.detach().numpy().reshape(-1) for w in self.w]).astype(np.float32)
        tab = self.t.grad.detach().numpy().reshape(-1).astype(np.float32)
        return float(L.detach()), mlp, tab
