"""C ABI checks that need no GPU: symbols, argument/state errors, and the host-built tree.

The product builds Rendu's tree on the host (update_hierarchy) before uploading it; without a HIP
device the upload fails loudly (RDU_ERR_HIP) but the host tree is already there, so it can be
compared node for node with the oracle's here.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GPU, ROOT
from oracle import oracle as orc
from rendu_amd import _lib, workloads
from rendu_amd.raycaster import Raycaster

F32 = np.float32


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(rdu_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = _declared_functions()
    assert "rdu_raycaster_intersect" in names and len(names) >= 14
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string():
    assert b"gfx950" in _lib.lib.rdu_version()
    assert _lib.lib.rdu_last_error() is not None


def _update(rc):
    try:
        rc.updateHierarchy()
    except _lib.RduError as e:
        assert not GPU and e.code == _lib.RDU_ERR_HIP, e
        assert "GPU" in e.message or "device" in e.message


def _scene_pair(meshes):
    ours, ref = Raycaster(), orc.OracleRaycaster()
    for mesh, model in meshes:
        ours.addMesh(mesh, model)
        ref.add_mesh(mesh.positions, mesh.indices, model)
    _update(ours)
    ref.update()
    return ours, ref


def _meshes_mixed():
    out = [(workloads.triangle_soup(5000, seed=11), None)]
    for k in range(3):
        s = workloads.icosphere(2 + (k % 2), radius=0.2, center=(0.2 * k, 0.1, -0.3))
        m = np.eye(4, dtype=F32)
        m[:3, :3] = np.array([[0.9, 0.1, 0.0], [-0.1, 0.9, 0.2], [0.0, -0.2, 1.1]], F32)
        m[:3, 3] = [0.5, 0.25 * k, 0.125]
        out.append((s, m))
    for q in workloads.box_quads((0, 0, 0), (1, 1, 1)):
        out.append((q, None))
    return out


@pytest.mark.parametrize("case", ["soup", "mixed", "tiny", "big", "big-mixed"])
def test_host_tree_equals_oracle_tree(case):
    """The host build equals the oracle's node for node. The big cases pass the size above which
    the build splits subtrees over threads (32768 triangles), within one mesh and across meshes."""
    if case == "soup":
        meshes = [(workloads.triangle_soup(20000, seed=3), None)]
    elif case == "mixed":
        meshes = _meshes_mixed()
    elif case == "big":
        meshes = [(workloads.triangle_soup(300000, seed=4), None)]
    elif case == "big-mixed":
        meshes = [(workloads.triangle_soup(n, seed=n), None) for n in (70000, 7, 40000, 90000, 3)] + _meshes_mixed()
    else:
        meshes = [(workloads.triangle_soup(n, seed=n), None) for n in (1, 2, 3, 4, 5, 9)]
    ours, ref = _scene_pair(meshes)
    a, b = ours.export_nodes(), ref.nodes()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    mesh, local = ours.export_triangles()
    _, rlocal, rmesh, _ = ref.triangles()
    assert np.array_equal(mesh, rmesh) and np.array_equal(local, rlocal.astype(np.uint32))
    info = ours.info()
    assert info["nodes"] == len(b[0]) and info["max_depth"] == ref.max_depth()
    assert info["meshes"] == len(meshes)


def test_state_and_argument_errors():
    rc = Raycaster()
    with pytest.raises(_lib.RduError) as e:
        rc.intersect_batch(np.zeros(1, _lib.RAY_DTYPE))
    assert e.value.code == _lib.RDU_ERR_STATE
    bad = workloads.MeshData(np.zeros((3, 3), F32), np.array([0, 1, 3], np.uint32))
    with pytest.raises(_lib.RduError) as e:
        rc.addMesh(bad)
    assert e.value.code == _lib.RDU_ERR_INVALID
    with pytest.raises(_lib.RduError) as e:
        rc.addMesh(workloads.MeshData(np.zeros((3, 3), F32), np.array([0, 1], np.uint32)))
    assert e.value.code == _lib.RDU_ERR_INVALID
    rc.addMesh(workloads.triangle_soup(10, seed=1))
    _update(rc)
    with pytest.raises(_lib.RduError) as e:
        rc.addMesh(workloads.triangle_soup(10, seed=2))
    assert e.value.code == _lib.RDU_ERR_STATE
    with pytest.raises(_lib.RduError) as e:
        rc.updateHierarchy()
    assert e.value.code == _lib.RDU_ERR_STATE


def test_tuning_and_graph_argument_errors():
    import ctypes
    lib = _lib.lib
    rc = Raycaster()
    rep = _lib.TuneReport()
    # Tuning queries need a built tree; explicit parameters are range-checked.
    assert lib.rdu_raycaster_tuning(rc._h, 1024, 0, 0, ctypes.byref(rep)) == _lib.RDU_ERR_STATE
    rc.addMesh(workloads.triangle_soup(100, seed=1))
    _update(rc)
    assert lib.rdu_raycaster_tuning(rc._h, 1024, 0, 0, None) == _lib.RDU_ERR_INVALID
    rc.set_tuning(24, 128, 2, incoherent=True)
    t = rc.tuning(1 << 20, incoherent=True)
    assert (t["refill_min"], t["chunk"], t["blocks_per_cu"], t["source"]) == (24, 128, 2, 1)
    assert rc.tuning(1 << 20)["source"] == 0  # another class
    for bad in ((65, 0, 0), (16, 8, 0), (16, 0, 65)):
        with pytest.raises(_lib.RduError) as e:
            rc.set_tuning(*bad)
        assert e.value.code == _lib.RDU_ERR_INVALID
    assert lib.rdu_raycaster_set_tuning(None, 0, 0, ctypes.byref(rep)) == _lib.RDU_ERR_INVALID
    # Scene-level entry points on a null handle.
    assert lib.rdu_scene_set_graph(None, 1) == _lib.RDU_ERR_INVALID
    assert lib.rdu_scene_autotune(None, None, 16, 5, 64, 64, None, None) == _lib.RDU_ERR_INVALID
    assert lib.rdu_tune_cache_size(None) == _lib.RDU_ERR_INVALID
    assert lib.rdu_tune_cache_save(None) == _lib.RDU_ERR_INVALID
    assert lib.rdu_tune_cache_load(None) == _lib.RDU_ERR_INVALID


def test_zero_triangle_mesh_rejected():
    rc = Raycaster()
    rc.addMesh(workloads.MeshData(np.zeros((3, 3), F32), np.zeros(0, np.uint32)))
    with pytest.raises(_lib.RduError) as e:
        rc.updateHierarchy()
    assert e.value.code == _lib.RDU_ERR_INVALID


@pytest.mark.skipif(GPU, reason="checks the no-device path")
def test_queries_fail_loudly_without_a_gpu():
    rc = Raycaster()
    rc.addMesh(workloads.triangle_soup(100, seed=1))
    _update(rc)
    with pytest.raises(_lib.RduError) as e:
        rc.intersect_batch(workloads.soup_rays(0, 16))
    assert e.value.code == _lib.RDU_ERR_HIP
    with pytest.raises(_lib.RduError):
        rc.intersects([0.5, 0.5, -1], [0, 0, 1])


def test_extension_entry_points_reject_bad_arguments():
    """Argument and state errors of the extension entry points are reported without a GPU."""
    L = _lib.lib
    c_void = ctypes.c_void_p
    # Radiance cache: null arguments, bad learning rate.
    lo = np.zeros(3, F32)
    assert L.rdu_rcache_create(None, _lib.fptr(lo), _lib.fptr(lo), 0) == _lib.RDU_ERR_INVALID
    assert L.rdu_rcache_set_learning_rate(None, ctypes.c_float(1e-3)) == _lib.RDU_ERR_INVALID
    assert L.rdu_rcache_infer(None, None, 1, None, 0, None) == _lib.RDU_ERR_INVALID
    assert L.rdu_rcache_train(None, None, None, 0, 1, 0, None, None) == _lib.RDU_ERR_INVALID
    assert L.rdu_rcache_get_params(None, None, None, None) == _lib.RDU_ERR_INVALID
    # Output step: components out of range, empty image is fine.
    assert L.rdu_image_to_ldr(None, 4, 4, 5, None, 0, None) == _lib.RDU_ERR_INVALID
    assert L.rdu_image_to_ldr(None, 0, 4, 3, None, 0, None) == _lib.RDU_OK
    # Scene: GPU-build and instancing switches and cached render on a null / uncommitted scene.
    assert L.rdu_scene_set_gpu_build(None, 1) == _lib.RDU_ERR_INVALID
    assert L.rdu_scene_set_instancing(None, 1) == _lib.RDU_ERR_INVALID
    h = c_void()
    assert L.rdu_scene_create(ctypes.byref(h)) == _lib.RDU_OK
    try:
        assert L.rdu_scene_set_gpu_build(h, 1) == _lib.RDU_OK
        assert L.rdu_scene_set_instancing(h, 1) == _lib.RDU_OK and L.rdu_scene_set_instancing(h, 0) == _lib.RDU_OK
        assert L.rdu_pathtracer_render_cached(h, None, None, 1, 1, None, 1, 1, 3, None, None) == _lib.RDU_ERR_INVALID
        stats = (ctypes.c_uint64 * 5)()
        assert L.rdu_scene_last_cache_stats(h, ctypes.byref(stats)) == _lib.RDU_OK
        n = ctypes.c_size_t(7)
        assert L.rdu_scene_last_cache_pairs(h, None, None, 0, ctypes.byref(n)) == _lib.RDU_OK and n.value == 0
    finally:
        L.rdu_scene_destroy(h)
    # Raycaster: GPU build twice / with no triangles, timing on a null handle.
    rc = c_void()
    assert L.rdu_raycaster_create(ctypes.byref(rc)) == _lib.RDU_OK
    try:
        assert L.rdu_raycaster_update_hierarchy_gpu(rc, None) == _lib.RDU_ERR_INVALID  # no triangles
    finally:
        L.rdu_raycaster_destroy(rc)
    assert L.rdu_raycaster_update_hierarchy_gpu(None, None) == _lib.RDU_ERR_INVALID
    assert L.rdu_raycaster_set_kernel_timing(None, 1) == _lib.RDU_ERR_INVALID
    ms, k = ctypes.c_double(), ctypes.c_uint64()
    assert L.rdu_raycaster_kernel_time(None, ctypes.byref(ms), ctypes.byref(k)) == _lib.RDU_ERR_INVALID


def test_batch_render_rejects_bad_jobs_without_a_gpu():
    """rdu_pathtracer_render_batch validates its job list before touching a device: an empty or
    null list, a job without a committed scene, a null camera or image, fewer than three
    components, an empty image or zero samples are argument or state errors."""
    from rendu_amd import scene as S
    L = _lib.lib
    assert L.rdu_pathtracer_render_batch(None, 1, 0, None, None) == _lib.RDU_ERR_INVALID
    jobs = (S.BatchJob * 1)()
    assert L.rdu_pathtracer_render_batch(ctypes.cast(jobs, ctypes.c_void_p), 0, 0, None, None) == _lib.RDU_ERR_INVALID
    h = ctypes.c_void_p()
    assert L.rdu_scene_create(ctypes.byref(h)) == _lib.RDU_OK
    try:
        cam = S.CameraDesc()
        img = np.zeros((4, 4, 3), F32)
        jobs[0] = S.BatchJob(h.value, ctypes.pointer(cam), None, 4, 2, _lib.fptr(img), 4, 4, 3, S.RenderOpts(1, 0, 3, 0))
        # The scene is not committed.
        assert L.rdu_pathtracer_render_batch(ctypes.cast(jobs, ctypes.c_void_p), 1, 0, None, None) == _lib.RDU_ERR_STATE
        assert b"committed" in L.rdu_last_error()
    finally:
        L.rdu_scene_destroy(h)
