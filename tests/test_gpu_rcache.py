"""Radiance cache on the GPU against the CPU reference (oracle/rcache_ref.py): inference, encoding
closed forms, gradients of the fused MFMA backward pass and the weight-gradient GEMMs, training."""
import numpy as np
import pytest

from oracle import rcache_ref as R

pytestmark = pytest.mark.gpu

LO, HI = np.array([-1.0, -0.5, 0.0], np.float32), np.array([1.0, 1.5, 3.0], np.float32)
N_MLP = sum(r * c for r, c in R.LAYER_SHAPES)


def _cache(seed=0):
    from rendu_amd.rcache import RadianceCache
    return RadianceCache(LO, HI, seed=seed)


def _positions(n, seed):
    rng = np.random.default_rng(seed)
    return (LO + rng.random((n, 3)) * (HI - LO)).astype(np.float32)


def _directions(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True) * scale).astype(np.float32)


def _set_random(cache, seed, table_scale=1.0):
    rng = np.random.default_rng(seed)
    mlp = np.concatenate([rng.uniform(-1, 1, r * c).astype(np.float32) * np.sqrt(6.0 / c)
                          for r, c in R.LAYER_SHAPES])
    mlp[-(16 - 3) * 64:] = 0.0  # padded output rows
    tables = (rng.uniform(-1, 1, R.LEVELS * R.TABLE * 2) * table_scale).astype(np.float32)
    cache.set_params(mlp, tables)
    return mlp, tables


def _layers(mlp):
    W, off = [], 0
    for r, cc in R.LAYER_SHAPES:
        W.append(mlp[off:off + r * cc].reshape(r, cc))
        off += r * cc
    return W


def test_initial_params(log_case):
    log_case("rcache initial params")
    c = _cache(3)
    mlp, tables, res = c.get_params()
    assert mlp.size == N_MLP
    assert list(res) == list(R.level_resolutions())
    assert np.abs(tables).max() <= 1e-4 and np.abs(tables).max() > 0
    assert np.all(mlp[-(16 - 3) * 64:] == 0)
    w0 = mlp[:64 * 48]
    assert np.abs(w0).max() <= np.sqrt(6 / 48) and w0.std() > 0.1


def test_infer_matches_reference(log_case):
    log_case("rcache inference vs reference, n=5000")
    c = _cache(1)
    mlp, tables = _set_random(c, 11)
    _, _, res = c.get_params()
    pos = _positions(5000, 12)
    dirs = _directions(5000, 13, scale=2.5)  # unnormalised on purpose: the kernel normalises
    y = c.infer(pos, dirs)
    ref = R.CacheReference(mlp, tables, res, LO, HI)
    yr = ref.forward(pos, dirs).detach().numpy()
    scale = np.abs(yr).mean()
    err = np.abs(y - yr)
    # bf16 activations: a rounding that flips between the two accumulation orders moves one
    # activation by 2^-8 relative; the output error stays a small fraction of its scale.
    assert err.mean() <= 5e-3 * scale, (err.mean(), scale)
    assert err.max() <= 5e-2 * np.abs(yr).max(), (err.max(), np.abs(yr).max())


def test_infer_ragged_and_device_pointers(log_case):
    import torch
    from rendu_amd.rcache import make_queries
    log_case("rcache inference ragged sizes, device pointers")
    c = _cache(2)
    _set_random(c, 21)
    pos, dirs = _positions(4099, 22), _directions(4099, 23)
    full = c.infer(pos, dirs)
    for n in (1, 15, 17, 1000):
        np.testing.assert_array_equal(c.infer(pos[:n], dirs[:n]), full[:n])
    dq = torch.from_numpy(make_queries(pos, dirs)).cuda()
    do = torch.zeros((pos.shape[0], 4), dtype=torch.float32, device="cuda")
    c.infer_device(dq.data_ptr(), pos.shape[0], do.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(do.cpu().numpy()[:, :3], full)


def test_encoding_closed_form(log_case):
    """MLP routing one encoding feature to the output; tables linear on a dense level."""
    log_case("rcache encoding closed form")
    c = _cache(4)
    res = R.level_resolutions()
    lvl = 3
    r1 = int(res[lvl]) + 1
    tables = np.zeros((R.LEVELS, R.TABLE, 2), np.float32)
    ci = np.arange(r1 ** 3)
    cx, cy, cz = ci % r1, (ci // r1) % r1, ci // (r1 * r1)
    tables[lvl, ci, 0] = 0.25 + (cx + 2 * cy + 3 * cz) / 64.0
    mlp = np.zeros(N_MLP, np.float32)
    W = _layers(mlp)
    W[0][0, 2 * lvl] = 1.0
    for l in range(1, 6):
        W[l][0, 0] = 1.0
    c.set_params(mlp, tables.reshape(-1))
    pos = _positions(3000, 5)
    y = c.infer(pos, _directions(3000, 6))
    p = (pos.astype(np.float64) - LO) / (HI - LO)
    x = np.clip(p, 0, 1) * float(res[lvl])
    expect = 0.25 + (x[:, 0] + 2 * x[:, 1] + 3 * x[:, 2]) / 64.0
    np.testing.assert_allclose(y[:, 0], expect, rtol=2 ** -8)
    assert np.all(y[:, 1:] == 0)


def test_direction_encoding_closed_form(log_case):
    """MLP routing the l = 1, m = 0 harmonic (0.4886 z) and the constant one to two outputs."""
    log_case("rcache direction encoding closed form")
    c = _cache(7)
    mlp = np.zeros(N_MLP, np.float32)
    W = _layers(mlp)
    W[0][0, 32 + 2] = 1.0      # Y_1^0 = c1 z
    W[0][1, 32 + 0] = 1.0      # Y_0^0 = c0
    for l in range(1, 5):
        W[l][0, 0] = 1.0
        W[l][1, 1] = 1.0
    W[5][0, 0] = 1.0
    W[5][1, 1] = 1.0
    c.set_params(mlp, np.zeros(R.LEVELS * R.TABLE * 2, np.float32))
    d = _directions(2000, 8)
    d[:, 2] = np.abs(d[:, 2])  # z >= 0 so the ReLU passes it
    y = c.infer(_positions(2000, 9), d * 3.0)
    np.testing.assert_allclose(y[:, 0], R.SH_C["c1"] * d[:, 2], rtol=2 ** -8, atol=1e-7)
    np.testing.assert_allclose(y[:, 1], R.SH_C["c0"], rtol=2 ** -8)


def test_gradients_match_autograd(log_case):
    log_case("rcache gradients vs autograd, n=3000")
    c = _cache(5)
    mlp, tables = _set_random(c, 31, table_scale=0.5)
    _, _, res = c.get_params()
    pos, dirs = _positions(3000, 32), _directions(3000, 34)
    rng = np.random.default_rng(33)
    tgt = rng.random((3000, 3)).astype(np.float32)
    gm, gt = c.gradients(pos, tgt, dirs)
    ref = R.CacheReference(mlp, tables, res, LO, HI)
    _, rm, rt = ref.gradients(pos, tgt, dirs)
    off = 0
    for l, (r, cc) in enumerate(R.LAYER_SHAPES):
        a, b = gm[off:off + r * cc], rm[off:off + r * cc]
        if l == 5:
            assert np.all(a[3 * 64:] == 0)
            a, b = a[:3 * 64], b[:3 * 64]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        # Output gradients and pre-activation gradients are carried in bf16 (8-bit significand).
        assert rel <= 3e-2, (l, rel)
        off += r * cc
    # The direction columns of W0 get gradients too.
    g0 = gm[:64 * 48].reshape(64, 48)
    assert np.abs(g0[:, 32:]).max() > 0
    rel_t = np.linalg.norm(gt - rt) / np.linalg.norm(rt)
    assert rel_t <= 3e-2, rel_t
    # No gradient lands on an entry the reference leaves untouched.
    assert np.abs(gt[rt == 0]).max() <= 1e-6 * np.abs(rt).max()
    # No stale gradients: a second call returns the same values.
    gm2, _ = c.gradients(pos, tgt, dirs, tables=False)
    assert np.linalg.norm(gm2 - gm) <= 1e-5 * np.linalg.norm(gm)


def _field(pos, dirs):
    p = (pos - LO) / (HI - LO)
    r = 0.6 + 0.4 * np.sin(6 * p[:, 0]) * np.cos(4 * p[:, 1])
    g = 0.3 + p[:, 2] ** 2 + 0.2 * np.maximum(dirs[:, 1], 0.0)
    b = 0.2 + 0.5 * p[:, 0] * p[:, 1]
    return np.stack([r, g, b], axis=1).astype(np.float32)


def test_training_fits_smooth_field(log_case):
    log_case("rcache training on a smooth field, 400 steps x 16384")
    c = _cache(6)
    c.set_learning_rate(1e-2)
    test_pos, test_dir = _positions(8192, 99), _directions(8192, 98)
    test_tgt = _field(test_pos, test_dir)
    err0 = np.mean((c.infer(test_pos, test_dir) - test_tgt) ** 2)
    rng = np.random.default_rng(7)
    losses = []
    for it in range(400):
        pos = (LO + rng.random((16384, 3)) * (HI - LO)).astype(np.float32)
        d = rng.normal(size=(16384, 3)).astype(np.float32)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        losses.append(c.train(pos, _field(pos, d), d))
    err1 = np.mean((c.infer(test_pos, test_dir) - test_tgt) ** 2)
    assert losses[-1] < 0.05 * losses[0], (losses[0], losses[-1])
    assert err1 < 0.01 * err0 and err1 < 1e-3, (err0, err1)


def _adam(w, g, m, v, step, lr, b1=0.9, b2=0.99, eps=1e-15):
    """One Adam step (Kingma and Ba 2015, algorithm 1) in float64; returns the new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    return w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps), m, v


def test_adam_is_lazy_on_table_entries(log_case):
    """Two training steps on batches that touch different table entries: the MLP takes dense Adam
    with its l2 term, the tables take Adam only where the step's batch lands (an entry touched by the
    first batch and not the second keeps its weight exactly; dense Adam would move it by its decayed
    first moment)."""
    log_case("rcache lazy table Adam, two steps, n=2000 then a sub-box of 1000")
    c = _cache(8)
    lr, l2 = 3e-3, 1e-6
    c.set_learning_rate(lr)
    _set_random(c, 41, table_scale=0.5)
    p0m, p0t, _ = c.get_params()
    rng = np.random.default_rng(42)
    pos_a, dir_a = _positions(2000, 43), _directions(2000, 44)
    tgt_a = rng.random((2000, 3)).astype(np.float32)
    sub = HI - 0.75 * (HI - LO)
    pos_b = (LO + np.random.default_rng(45).random((1000, 3)) * (sub - LO)).astype(np.float32)
    dir_b, tgt_b = _directions(1000, 46), rng.random((1000, 3)).astype(np.float32)

    ga_m, ga_t = c.gradients(pos_a, tgt_a, dir_a)
    c.train(pos_a, tgt_a, dir_a)
    p1m, p1t, _ = c.get_params()
    gb_m, gb_t = c.gradients(pos_b, tgt_b, dir_b)
    c.train(pos_b, tgt_b, dir_b)
    p2m, p2t, _ = c.get_params()

    w, g = p0t.astype(np.float64), ga_t.astype(np.float64)
    ta, tb = ga_t != 0, gb_t != 0
    e1, m1, v1 = _adam(w, g, 0.0, 0.0, 1, lr)
    assert np.array_equal(p1t[~ta], p0t[~ta])
    assert np.allclose(p1t[ta], e1[ta], rtol=0, atol=2e-6)
    w, g = p1t.astype(np.float64), gb_t.astype(np.float64)
    m1, v1 = np.where(ta, m1, 0.0), np.where(ta, v1, 0.0)
    e2, _, _ = _adam(w, g, m1, v1, 2, lr)
    only_a = ta & ~tb
    assert only_a.sum() > 1000 and tb.sum() > 1000
    assert np.array_equal(p2t[~tb], p1t[~tb])
    assert np.allclose(p2t[tb], e2[tb], rtol=0, atol=2e-6)
    dense, _, _ = _adam(w, g, m1, v1, 2, lr)
    assert np.abs(dense[only_a] - p1t[only_a]).max() > 1e-4  # what dense Adam would have done

    # MLP: dense, gradient plus l2 * w; its weight gradients are summed with atomics, so a few
    # near-zero components may round to the other sign between the two passes.
    live = np.ones(N_MLP, bool)
    live[-(16 - 3) * 64:] = False
    e1m, mm, vm = _adam(p0m.astype(np.float64), ga_m + l2 * p0m.astype(np.float64), 0.0, 0.0, 1, lr)
    e2m, _, _ = _adam(p1m.astype(np.float64), gb_m + l2 * p1m.astype(np.float64), mm, vm, 2, lr)
    for got, exp in ((p1m, e1m), (p2m, e2m)):
        assert np.mean(np.abs(got[live] - exp[live]) > 1e-3 * lr) < 1e-3
