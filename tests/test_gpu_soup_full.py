"""The 1M-triangle soup with 2^24 rays (the ray-cast benchmark workload) on the GPU.

At this size the oracle only checks a seeded subset; every ray is checked through properties
that hold independently of size: any-hit agrees with closest-hit's hit flag, both visiting
orders agree, and the returned (t, u, v) is exactly what Moller-Trumbore gives for the returned
triangle (recomputed in numpy float32 in Rendu's operation order).
"""
import numpy as np
import pytest

from conftest import log_case
from oracle import oracle as orc
from rendu_amd import workloads
from rendu_amd.raycaster import Raycaster

pytestmark = pytest.mark.gpu
F32 = np.float32


def _dot(a, b):
    t = a * b
    return (t[:, 0] + t[:, 1]) + t[:, 2]


def _cross(x, y):
    return np.stack([x[:, 1] * y[:, 2] - y[:, 1] * x[:, 2], x[:, 2] * y[:, 0] - y[:, 2] * x[:, 0],
                     x[:, 0] * y[:, 1] - y[:, 0] * x[:, 1]], -1)


@pytest.fixture(scope="module")
def soup():
    log_case("full soup: 1M triangles, 2^24 rays")
    mesh = workloads.triangle_soup(1_000_000, seed=1)
    rays = workloads.soup_rays(1 << 23, 1 << 23, seed=2)
    rc = Raycaster()
    rc.addMesh(mesh)
    rc.updateHierarchy()
    return mesh, rays, rc


def test_full_size_properties(soup):
    mesh, rays, rc = soup
    log_case("full soup closest near-first + any-hit")
    hits = rc.intersect_batch(rays)
    occ = rc.intersect_any_batch(rays)
    assert np.array_equal(occ, hits["hit"] == 1)
    h = hits["hit"] == 1
    assert 0.5 < h.mean() <= 1.0
    assert np.all(hits["mesh_id"] == 0) and np.all(hits["local_id"] % 3 == 0)
    assert np.all((hits["t"][h] > F32(1e-4)) & (hits["t"][h] < F32(1e8)))
    assert np.all(hits["t"][~h] == np.finfo(F32).max)
    # Recompute Moller-Trumbore on the returned triangles (Raycaster.cpp:232-262).
    idx = np.nonzero(h)[0]
    lid = hits["local_id"][idx].astype(np.int64)
    p = mesh.positions
    v0, v1, v2 = p[lid], p[lid + 1], p[lid + 2]
    o = np.stack([rays["ox"][idx], rays["oy"][idx], rays["oz"][idx]], -1)
    d = np.stack([rays["dx"][idx], rays["dy"][idx], rays["dz"][idx]], -1)
    n2 = _dot(d, d)
    d = d * (F32(1.0) / np.sqrt(n2))[:, None]
    e1, e2 = v1 - v0, v2 - v0
    pp = _cross(d, e2)
    inv = F32(1.0) / _dot(e1, pp)
    q = o - v0
    u = inv * _dot(q, pp)
    r = _cross(q, e1)
    v = inv * _dot(d, r)
    t = inv * _dot(e2, r)
    assert np.array_equal(t.view(np.uint32), hits["t"][idx].view(np.uint32))
    assert np.array_equal(u.view(np.uint32), hits["u"][idx].view(np.uint32))
    assert np.array_equal(v.view(np.uint32), hits["v"][idx].view(np.uint32))


def test_orders_agree_and_subset_matches_oracle(soup):
    mesh, rays, rc = soup
    log_case("full soup reference order vs near-first; oracle subset")
    ref_order = rc.intersect_batch(rays, reference_order=True)
    near = rc.intersect_batch(rays)
    same = (ref_order["local_id"] == near["local_id"]) & (ref_order["hit"] == near["hit"])
    tie = np.abs(ref_order["t"] - near["t"]) <= 1e-5 * np.abs(ref_order["t"])
    assert np.all(same | tie) and (~same).mean() < 1e-4
    ora = orc.OracleRaycaster()
    ora.add_mesh(mesh.positions, mesh.indices)
    ora.update()
    rng = np.random.default_rng(0)
    sub = np.sort(rng.choice(rays.shape[0], 20000, replace=False))
    exp = ora.intersect(rays[sub], "closest")
    for k in ("t", "u", "v", "hit", "mesh_id", "local_id"):
        assert np.array_equal(ref_order[k][sub].view(np.uint32), exp[k].view(np.uint32)), k
