"""Pins the radiance-cache reference (oracle/rcache_ref.py) on the CPU: hash, trilinear encoding,
autograd gradients and the stop-gradient in the relative L2 loss."""
import numpy as np
import pytest
import torch

from oracle import rcache_ref as R


def test_level_resolutions_geometric():
    res = R.level_resolutions()
    assert res[0] == 16 and res[-1] in (2047, 2048)
    assert np.all(np.diff(res) > 0)
    dense = R.dense_levels(res)
    # (res+1)^3 <= 2^19 exactly for the coarse levels 16..58.
    assert list(dense[:5]) == [1, 1, 1, 1, 1] and dense[5:].sum() == 0


def test_hash_matches_integer_arithmetic():
    rng = np.random.default_rng(1)
    p = rng.random((500, 3)).astype(np.float32)
    for res in (80, 2047):
        idx, _ = R.corner_indices(p, res, dense=False)
        x = p * np.float32(res)
        ci = np.clip(np.floor(x).astype(np.int64), 0, res - 1)
        for i in range(0, 500, 7):
            for c in range(8):
                d = (c & 1, (c >> 1) & 1, (c >> 2) & 1)
                assert idx[i, c] == R.hash_int(*(int(ci[i, a]) + d[a] for a in range(3)))


def test_trilinear_weights_partition_of_unity():
    rng = np.random.default_rng(2)
    p = rng.random((1000, 3)).astype(np.float32)
    _, w = R.corner_indices(p, 22, dense=True)
    assert np.all(w >= 0)
    np.testing.assert_allclose(w.sum(axis=1), 1.0, rtol=0, atol=3e-7)


def _linear_tables(res: int, level: int):
    """Dense level filled with f(c) = 0.25 + (cx + 2 cy + 3 cz) / 64 (feature 0) and 1 - cx / 128."""
    t = np.zeros((R.LEVELS, R.TABLE, 2), np.float32)
    r1 = res + 1
    c = np.arange(r1 ** 3)
    cx, cy, cz = c % r1, (c // r1) % r1, c // (r1 * r1)
    t[level, c, 0] = 0.25 + (cx + 2 * cy + 3 * cz) / 64.0
    t[level, c, 1] = 1.0 - cx / 128.0
    return t


def test_trilinear_reproduces_linear_function():
    res = R.level_resolutions()
    t = _linear_tables(int(res[2]), 2)
    mlp = np.zeros(sum(r * c for r, c in R.LAYER_SHAPES), np.float32)
    ref = R.CacheReference(mlp, t, res, [0, 0, 0], [1, 1, 1])
    rng = np.random.default_rng(3)
    pos = rng.random((2000, 3)).astype(np.float32)
    enc = ref.encode(pos).detach().numpy()
    x = pos.astype(np.float64) * float(res[2])
    expect0 = 0.25 + (x[:, 0] + 2 * x[:, 1] + 3 * x[:, 2]) / 64.0
    expect1 = 1.0 - x[:, 0] / 128.0
    np.testing.assert_allclose(enc[:, 4], expect0, rtol=2e-6)
    np.testing.assert_allclose(enc[:, 5], expect1, rtol=2e-6)
    assert np.abs(np.delete(enc, [4, 5], axis=1)).max() == 0.0


def test_spherical_harmonics_orthonormal():
    """Gauss-Legendre in cos(theta) x uniform in phi integrates products of degree <= 6 exactly."""
    xs, ws = np.polynomial.legendre.leggauss(16)
    nphi = 32
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    z = np.repeat(xs, nphi)
    w = np.repeat(ws, nphi) * (2 * np.pi / nphi)
    s = np.sqrt(1 - z * z)
    d = np.stack([s * np.cos(np.tile(phi, 16)), s * np.sin(np.tile(phi, 16)), z], axis=1)
    Y = R.sh_features(d).astype(np.float64)
    gram = (Y * w[:, None]).T @ Y
    np.testing.assert_allclose(gram, np.eye(16), atol=2e-6)


def test_direction_normalisation():
    d = np.array([[0, 0, 5], [3, 4, 0], [0, 0, 0], [np.nan, 1, 0], [np.inf, 0, 0]], np.float32)
    u = R.unit_dirs(d)
    np.testing.assert_allclose(u[:2], [[0, 0, 1], [0.6, 0.8, 0]], rtol=1e-7)
    assert np.all(u[2:] == [0, 0, 1])


def _random_params(seed):
    rng = np.random.default_rng(seed)
    mlp = np.concatenate([rng.uniform(-1, 1, r * c).astype(np.float32) * np.sqrt(6.0 / c)
                          for r, c in R.LAYER_SHAPES])
    tables = rng.uniform(-1, 1, R.LEVELS * R.TABLE * 2).astype(np.float32)
    return mlp, tables


def test_autograd_matches_finite_differences():
    mlp, tables = _random_params(4)
    res = R.level_resolutions()
    rng = np.random.default_rng(5)
    pos = rng.random((24, 3)).astype(np.float32)
    tgt = rng.random((24, 3)).astype(np.float32)
    ref = R.CacheReference(mlp, tables, res, [0, 0, 0], [1, 1, 1], round_like_gpu=False, dtype=torch.float64)
    _, gm, gt = ref.gradients(pos, tgt)
    with torch.no_grad():
        den = ref.forward(pos) ** 2 + 0.01
    t = torch.tensor(tgt, dtype=torch.float64)

    def frozen_loss():
        with torch.no_grad():
            return float((((ref.forward(pos) - t) ** 2) / den).mean())

    h = 1e-6
    # MLP parameters (one per layer) and touched table entries.
    offs = np.cumsum([0] + [r * c for r, c in R.LAYER_SHAPES])
    for l, (r, c) in enumerate(R.LAYER_SHAPES):
        rows = 3 if l == 5 else r
        k = int(rng.integers(rows * c))
        i, j = divmod(k, c)
        w = ref.w[l]
        with torch.no_grad():
            w[i, j] += h
            lp = frozen_loss()
            w[i, j] -= 2 * h
            lm = frozen_loss()
            w[i, j] += h
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gm[offs[l] + k]) <= 1e-6 + 1e-5 * abs(fd), (l, fd, gm[offs[l] + k])
    nz = np.nonzero(gt)[0]
    assert nz.size > 0
    for k in rng.choice(nz, 12, replace=False):
        lv, rem = divmod(int(k), R.TABLE * 2)
        e, f = divmod(rem, 2)
        with torch.no_grad():
            ref.t[lv, e, f] += h
            lp = frozen_loss()
            ref.t[lv, e, f] -= 2 * h
            lm = frozen_loss()
            ref.t[lv, e, f] += h
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gt[k]) <= 1e-8 + 1e-5 * abs(fd), (k, fd, gt[k])


def test_loss_denominator_is_constant():
    """The gradient differs from that of the loss with a live denominator (a dropped detach fails)."""
    mlp, tables = _random_params(6)
    res = R.level_resolutions()
    rng = np.random.default_rng(7)
    pos = rng.random((16, 3)).astype(np.float32)
    tgt = rng.random((16, 3)).astype(np.float32)
    ref = R.CacheReference(mlp, tables, res, [0, 0, 0], [1, 1, 1], round_like_gpu=False, dtype=torch.float64)
    _, gm, _ = ref.gradients(pos, tgt)
    for w in ref.w:
        w.grad = None
    y = ref.forward(pos)
    t = torch.tensor(tgt, dtype=torch.float64)
    live = (((y - t) ** 2) / (y ** 2 + 0.01)).mean()
    live.backward()
    g_live = np.concatenate([w.grad.numpy().reshape(-1) for w in ref.w])
    assert np.linalg.norm(g_live - gm) > 1e-3 * np.linalg.norm(gm)


def test_bf16_rounding_nearest_even():
    x = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 1.0 + 2 ** -8 + 2 ** -20])
    y = R.bf16(x)
    assert y.tolist() == [1.0, 1.0 + 2 ** -6, -1.0, 1.0 + 2 ** -7]
