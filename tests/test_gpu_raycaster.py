"""GPU parity of the batch traversal kernels (trace.hip) against the oracle, through the C ABI.

Bar: reference order is bit-exact on every field (t, u, v, hit, mesh_id, local_id) against the
oracle's restatement of Raycaster::intersects. Near-first order returns the same hit except on
near-ties, where the returned t must still match within 1e-5 relative. Any-hit flags are exact.
"""
import os

import numpy as np
import pytest

from conftest import log_case
from oracle import oracle as orc
from rendu_amd import _lib, workloads
from rendu_amd.raycaster import Raycaster, make_rays

pytestmark = pytest.mark.gpu
F32 = np.float32


def build(meshes):
    ours, ref = Raycaster(), orc.OracleRaycaster()
    for mesh, model in meshes:
        ours.addMesh(mesh, model)
        ref.add_mesh(mesh.positions, mesh.indices, model)
    ours.updateHierarchy()
    ref.update()
    return ours, ref


def mixed_scene():
    out = [(workloads.triangle_soup(20000, seed=21, log_edge=-3.5), None)]
    for k in range(4):
        s = workloads.icosphere(3, radius=0.12, center=(0.25 * k, 0.5, 0.5))
        m = np.eye(4, dtype=F32)
        m[:3, :3] = np.array([[1.2, 0.1, 0.0], [-0.1, 0.8, 0.2], [0.0, -0.2, 1.0]], F32)
        m[:3, 3] = [0.1, 0.05 * k, -0.1]
        out.append((s, m))
    for q in workloads.box_quads((-0.1, -0.1, -0.1), (1.1, 1.1, 1.1)):
        out.append((q, None))
    return out


def special_rays():
    """Axis-parallel rays, origins on box planes, -0 components, zero and tiny directions, short tmax."""
    o, d, tmax = [], [], []
    for a in range(3):
        for s in (-1.0, -0.0, 0.0, 1.0):
            dd = np.zeros(3, F32)
            dd[a] = s if s != 0 else 0.0
            dd[(a + 1) % 3] = s
            o.append([0.5, 0.5, 0.5]); d.append(dd); tmax.append(1e8)
    for a in range(3):
        dd = np.zeros(3, F32)
        dd[a] = 1.0
        for plane in (-0.1, 0.0, 1.0, 1.1):
            oo = np.full(3, 0.5, F32)
            oo[(a + 1) % 3] = plane
            o.append(oo); d.append(dd); tmax.append(1e8)
    o.append([0.5, 0.5, -1.0]); d.append([0.0, 0.0, 0.0]); tmax.append(1e8)          # NaN direction
    o.append([0.5, 0.5, -1.0]); d.append([1e-30, 1e-30, 1.0]); tmax.append(1e8)     # huge reciprocals
    o.append([0.5, 0.5, -1.0]); d.append([0.0, 0.0, 1.0]); tmax.append(0.5)         # stops short
    o.append([0.5, 0.5, 0.5]); d.append([0.3, 0.4, 0.5]); tmax.append(1e-4)         # empty interval
    return make_rays(np.array(o, F32), np.array(d, F32), 1e-4, np.array(tmax, F32))


def compare_exact(got, ref):
    for k in ("t", "u", "v"):
        g = got[k].view(np.uint32)
        r = ref[k].view(np.uint32)
        bad = np.nonzero(g != r)[0]
        assert bad.size == 0, f"{k} differs at {bad[:8]}"
    for k in ("hit", "mesh_id", "local_id"):
        bad = np.nonzero(got[k] != ref[k])[0]
        assert bad.size == 0, f"{k} differs at {bad[:8]}"


def compare_near_first(got, ref):
    assert np.array_equal(got["hit"], ref["hit"])
    same = (got["mesh_id"] == ref["mesh_id"]) & (got["local_id"] == ref["local_id"])
    tie = np.abs(got["t"] - ref["t"]) <= 1e-5 * np.abs(ref["t"])
    assert np.all(same | tie)
    assert np.all(got["t"][same].view(np.uint32) == ref["t"][same].view(np.uint32))
    assert (~same).mean() < 1e-3


@pytest.fixture(scope="module")
def mixed():
    return build(mixed_scene())


@pytest.fixture(scope="module")
def mixed_rays():
    r = workloads.soup_rays(32768, 32768, seed=5, width=256)
    return np.concatenate([r, special_rays()])


def test_reference_order_bit_exact(mixed, mixed_rays):
    ours, ref = mixed
    log_case("closest reference-order mixed scene 65k rays")
    got = ours.intersect_batch(mixed_rays, reference_order=True)
    exp = ref.intersect(mixed_rays, "closest")
    compare_exact(got, exp)
    assert 0.3 < exp["hit"].mean() < 1.0


def test_near_first_closest(mixed, mixed_rays):
    ours, ref = mixed
    log_case("closest near-first mixed scene")
    got = ours.intersect_batch(mixed_rays)
    compare_near_first(got, ref.intersect(mixed_rays, "closest"))


@pytest.mark.parametrize("layout", ["wide", "wide-simple", "binary"])
@pytest.mark.parametrize("incoherent", [False, True])
def test_near_first_layouts_and_ordering(mixed, mixed_rays, monkeypatch, layout, incoherent):
    ours, ref = mixed
    if layout == "binary":
        monkeypatch.setenv("RDU_TRACE_LAYOUT", "binary")
    if layout == "wide-simple":
        monkeypatch.setenv("RDU_TRACE_KERNEL", "simple")
    log_case(f"near-first layout={layout} incoherent={incoherent}")
    exp = ref.intersect(mixed_rays, "closest")
    compare_near_first(ours.intersect_batch(mixed_rays, incoherent=incoherent), exp)
    occ = ours.intersect_any_batch(mixed_rays, incoherent=incoherent)
    assert np.array_equal(occ, exp["hit"].astype(bool))


@pytest.mark.parametrize("reference_order", [False, True])
def test_any_hit_exact(mixed, mixed_rays, reference_order):
    ours, ref = mixed
    log_case(f"any-hit mixed scene ref={reference_order}")
    got = ours.intersect_any_batch(mixed_rays, reference_order=reference_order)
    exp = ref.intersect(mixed_rays, "any")["hit"].astype(bool)
    assert np.array_equal(got, exp)


@pytest.fixture(params=[False, True], ids=["host-walk", "gpu-launch"])
def per_ray_mode(request, mixed):
    """Per-ray calls on the host walk of the uploaded tree (default) or as one-ray GPU launches."""
    ours, _ = mixed
    ours.set_per_ray_device(request.param)
    yield request.param
    ours.set_per_ray_device(False)


def test_per_ray_api(mixed, per_ray_mode):
    ours, ref = mixed
    log_case(f"per-ray intersects/intersectsAny/visible (GPU launch: {per_ray_mode})")
    rays = workloads.soup_rays(0, 64, seed=9)
    exp = ref.intersect(rays, "closest")
    for i in range(64):
        o = [rays["ox"][i], rays["oy"][i], rays["oz"][i]]
        d = [rays["dx"][i], rays["dy"][i], rays["dz"][i]]
        h = ours.intersects(o, d)
        assert h.hit == bool(exp["hit"][i])
        if h.hit:
            assert (np.float32(h.dist), np.float32(h.u), np.float32(h.v)) == (exp["t"][i], exp["u"][i], exp["v"][i])
            assert (h.meshId, h.localId) == (exp["mesh_id"][i], exp["local_id"][i])
            assert np.float32(h.w) == exp["w"][i]
        assert ours.intersectsAny(o, d) == bool(exp["hit"][i])
        p1 = np.asarray(o, F32) + F32(0.05) * np.asarray(d, F32)
        vis = ours.visible(o, p1)
        seg = make_rays(np.asarray(o, F32)[None], (p1 - np.asarray(o, F32))[None], 1e-4,
                        np.sqrt(np.sum((p1 - np.asarray(o, F32)) ** 2, dtype=F32)))
        assert vis == (not bool(ref.intersect(seg, "any")["hit"][0]))


def test_per_ray_api_from_many_threads(mixed, per_ray_mode):
    """Per-ray calls from 8 host threads at once (host walk: no lock; GPU launches: each thread has
    its own stream and pinned record): every result equals the oracle's, as from one thread."""
    import threading
    ours, ref = mixed
    log_case(f"per-ray intersects from 8 threads (GPU launch: {per_ray_mode})")
    rays = workloads.soup_rays(0, 8 * 96, seed=19)
    exp = ref.intersect(rays, "closest")
    got = [None] * rays.shape[0]
    occ = [None] * rays.shape[0]

    def work(k):
        for i in range(k, rays.shape[0], 8):
            o = [rays["ox"][i], rays["oy"][i], rays["oz"][i]]
            d = [rays["dx"][i], rays["dy"][i], rays["dz"][i]]
            got[i] = ours.intersects(o, d)
            occ[i] = ours.intersectsAny(o, d)

    ths = [threading.Thread(target=work, args=(k,)) for k in range(8)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for i, h in enumerate(got):
        assert h.hit == bool(exp["hit"][i]) and occ[i] == bool(exp["hit"][i])
        if h.hit:
            assert (np.float32(h.dist), np.float32(h.u), np.float32(h.v)) == (exp["t"][i], exp["u"][i], exp["v"][i])
            assert (h.meshId, h.localId) == (exp["mesh_id"][i], exp["local_id"][i])


def test_spill_path_bit_exact(mixed, mixed_rays, monkeypatch):
    ours, ref = mixed
    monkeypatch.setenv("RDU_DEBUG_LDS_STACK", "3")
    log_case("closest + any with 3 LDS stack levels (global spill)")
    got = ours.intersect_batch(mixed_rays[:20000], reference_order=True)
    compare_exact(got, ref.intersect(mixed_rays[:20000], "closest"))
    got = ours.intersect_batch(mixed_rays[:20000])
    compare_near_first(got, ref.intersect(mixed_rays[:20000], "closest"))
    occ = ours.intersect_any_batch(mixed_rays[:20000])
    assert np.array_equal(occ, ref.intersect(mixed_rays[:20000], "any")["hit"].astype(bool))
    monkeypatch.setenv("RDU_TRACE_LAYOUT", "binary")
    compare_near_first(ours.intersect_batch(mixed_rays), ref.intersect(mixed_rays, "closest"))


@pytest.mark.parametrize("n", [1, 300, 1100])
def test_spill_small_batch_fresh_raycaster(mixed_rays, monkeypatch, n):
    """Small near-first batches on a raycaster that has not traced anything yet: the persistent grid
    rounds its blocks up to a whole number per XCD (more lanes than ceil(n / 128) blocks), and the
    spill stack is sized for that grid."""
    ours, ref = build(mixed_scene())
    monkeypatch.setenv("RDU_DEBUG_LDS_STACK", "1")
    log_case(f"fresh raycaster, {n} rays, 1 LDS stack level")
    rays = mixed_rays[:n]
    compare_near_first(ours.intersect_batch(rays), ref.intersect(rays, "closest"))
    assert np.array_equal(ours.intersect_any_batch(rays), ref.intersect(rays, "any")["hit"].astype(bool))


def test_many_roots_and_tiny_meshes():
    meshes = [(workloads.triangle_soup(n, seed=100 + k, log_edge=-2.0), None)
              for k, n in enumerate([1, 2, 3, 4, 5, 7, 16, 33] * 25)]
    ours, ref = build(meshes)
    rays = workloads.soup_rays(4096, 4096, seed=3, width=64)
    log_case("200 meshes closest reference-order")
    compare_exact(ours.intersect_batch(rays, reference_order=True), ref.intersect(rays, "closest"))
    compare_near_first(ours.intersect_batch(rays), ref.intersect(rays, "closest"))


def test_autotune_keeps_results(mixed, mixed_rays):
    ours, ref = mixed
    log_case("autotune incoherent closest + any, then compare")
    rep = ours.autotune(mixed_rays, incoherent=True)
    assert rep["refill_min"] in (4, 8, 16, 24, 32, 40, 48, 64) and rep["best_ms"] > 0
    rep = ours.autotune(mixed_rays, any_hit=True, incoherent=True)
    assert rep["refill_min"] > 0
    exp = ref.intersect(mixed_rays, "closest")
    compare_near_first(ours.intersect_batch(mixed_rays, incoherent=True), exp)
    assert np.array_equal(ours.intersect_any_batch(mixed_rays, incoherent=True), exp["hit"].astype(bool))


def test_empty_batch(mixed):
    ours, _ = mixed
    assert ours.intersect_batch(np.zeros(0, _lib.RAY_DTYPE)).shape == (0,)
    assert ours.intersect_any_batch(np.zeros(0, _lib.RAY_DTYPE)).shape == (0,)


def test_device_pointer_entry(mixed, mixed_rays):
    import torch
    ours, ref = mixed
    log_case("device-pointer entry on torch stream")
    rays = np.ascontiguousarray(mixed_rays)
    d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_hits = torch.empty(rays.shape[0] * 24, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    ours.intersect_device(d_rays.data_ptr(), rays.shape[0], d_hits.data_ptr(), stream, reference_order=True)
    torch.cuda.synchronize()
    got = d_hits.cpu().numpy().view(_lib.HIT_DTYPE)
    compare_exact(got, ref.intersect(rays, "closest"))


def test_traversal_stats_counts(mixed, mixed_rays):
    ours, ref = mixed
    r = mixed_rays[:8192]
    log_case("trace stats reference-order")
    nodes, tris = ours.trace_stats(r, reference_order=True)
    _, (onodes, otris) = ref.intersect(r, "closest", stats=True)
    # The reference-order kernel makes exactly the oracle's expansions and triangle tests.
    assert (nodes, tris) == (onodes, otris)
