"""Per-ray calls from a compiled Rendu-style caller (tests/cpp/perray_threads.cpp): 7 host threads
over contiguous spans of rays, one call per ray, as PathTracer::render's forParallel workers call
Raycaster::intersects (PathTracer.cpp:186, 204; System.hpp:72-104).

Every closest hit is bit-identical to the GPU's RDU_REFERENCE_ORDER batch on the same rays (hit flag,
t, u, v, mesh and triangle ids), every any-hit flag equals the batch any hit, in both per-ray modes
(the host walk of the uploaded tree and one-ray GPU launches). On the 1M-triangle soup with
incoherent rays the host walk answers at least Rendu's own per-ray rate on 7 threads
(0.336 M rays/s for Raycaster::intersects, measured with the reference build on an 8-core host).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, log_case
from rendu_amd import _lib, workloads
from rendu_amd.raycaster import Raycaster

pytestmark = pytest.mark.gpu

RENDU_PER_RAY_MRAYS_7_THREADS = 0.336


def build_harness(tmp_path):
    exe = os.path.join(str(tmp_path), "perray_threads")
    subprocess.run(["g++", "-O2", "-std=c++14", "-pthread", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "perray_threads.cpp"), "-L", os.path.join(ROOT, "rendu_amd"),
                    "-lrdu", f"-Wl,-rpath,{os.path.join(ROOT, 'rendu_amd')}", "-o", exe], check=True)
    return exe


def write_scene(path, mesh, rays):
    with open(path, "wb") as f:
        pos = np.ascontiguousarray(mesh.positions, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(mesh.indices, np.uint32)
        np.array([pos.shape[0], idx.shape[0], rays.shape[0]], np.uint32).tofile(f)
        pos.tofile(f)
        idx.tofile(f)
        np.ascontiguousarray(rays).tofile(f)


def run_harness(exe, scene, hits, threads, device=False):
    r = subprocess.run([exe, scene, hits, str(threads), "1" if device else "0"], capture_output=True, text=True,
                       timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    fields = dict(kv.split("=") for kv in r.stdout.split())
    n = int(fields["rays"])
    raw = np.fromfile(hits, np.uint8)
    got = raw[: n * 24].view(_lib.HIT_DTYPE)
    anyhit = raw[n * 24:].astype(bool)
    return got, anyhit, float(fields["closest_mrays"]), float(fields["any_mrays"])


def check_against_batch(rc, rays, got, anyhit):
    exp = rc.intersect_batch(rays, reference_order=True)
    for k in ("t", "u", "v", "hit", "mesh_id", "local_id"):
        assert np.array_equal(got[k].view(np.uint32), exp[k].view(np.uint32)), k
    assert np.array_equal(anyhit, rc.intersect_any_batch(rays, reference_order=True))


@pytest.mark.parametrize("device", [False, True], ids=["host-walk", "gpu-launch"])
def test_per_ray_threads_match_the_reference_order_batch(tmp_path, device):
    mesh = workloads.triangle_soup(50000, seed=5, log_edge=-3.8)
    rays = workloads.soup_rays(2048, 2048, seed=6, width=64)
    exe = build_harness(tmp_path)
    scene, hits = str(tmp_path / "scene.bin"), str(tmp_path / "hits.bin")
    write_scene(scene, mesh, rays)
    log_case(f"per-ray harness, 50k soup, 4096 rays, 7 threads, GPU launch: {device}")
    got, anyhit, _, _ = run_harness(exe, scene, hits, 7, device)
    rc = Raycaster()
    rc.addMesh(mesh)
    rc.updateHierarchy()
    assert got["hit"].mean() > 0.3
    check_against_batch(rc, rays, got, anyhit)


def test_per_ray_rate_on_the_1m_soup(tmp_path):
    mesh = workloads.triangle_soup(1_000_000, seed=1)
    rays = workloads.soup_rays(0, 1 << 18, seed=2)
    exe = build_harness(tmp_path)
    scene, hits = str(tmp_path / "scene.bin"), str(tmp_path / "hits.bin")
    write_scene(scene, mesh, rays)
    log_case("per-ray harness, 1M soup, 2^18 incoherent rays, 7 threads, host walk")
    got, anyhit, closest_rate, any_rate = run_harness(exe, scene, hits, 7)
    print(f"per-ray 7 threads: closest {closest_rate:.3f} M rays/s, any {any_rate:.3f} M rays/s", flush=True)
    rc = Raycaster()
    rc.addMesh(mesh)
    rc.updateHierarchy()
    check_against_batch(rc, rays, got, anyhit)
    assert closest_rate >= RENDU_PER_RAY_MRAYS_7_THREADS
