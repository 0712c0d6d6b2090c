"""Batched render jobs: many independent (scene, camera, image, spp, depth) jobs on one GPU.

No Rendu counterpart (Rendu renders one scene per process). A job is the whole drop-in sequence of
the path-tracer app for one scene: build the raycaster over the scene's objects (host BVH build and
upload, rdu_scene_commit) and render it (rdu_pathtracer_render). Jobs run on a pool of host threads;
the library releases Python's lock inside every call and renders each scene on its own HIP stream,
so one job's host-side build overlaps other jobs' GPU work and small renders share the GPU.
Scenes stay alive until the batch ends, so that no device free (which waits for the whole device)
lands inside the batch.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import scene as S
from .pathtracer import PathTracer


@dataclass
class Job:
    scene: S.Scene
    camera: object
    width: int
    height: int
    samples: int
    depth: int
    seed: int = 0
    estimator: int = S.ESTIMATOR_REFERENCE
    rr_start: int = 3
    gpu_build: bool = False  # build the tree on the GPU instead of Rendu's host build
    instancing: bool = False  # an instanced scene (rdu_scene_set_instancing): shared meshes built once
    # Passes of this job's render kept in flight at once (0 = the library's default, 3). One by
    # default: the batch already runs several jobs at once, and more streams per job cost the batch
    # throughput (64 jobs: 53-55 jobs/s with 3, 56-63 with 1) for a lower p99 latency.
    passes_in_flight: int = 1
    max_paths: int = 0  # paths per pass (0 = the library's default)


@dataclass
class JobResult:
    index: int
    image: np.ndarray
    latency_s: float           # build + render, from the job's start
    build_s: float
    render_s: float
    paths: int
    rays: int
    started_s: float = 0.0     # relative to the batch start
    stats: dict = field(default_factory=dict)


def _build(job: Job):
    t0 = time.perf_counter()
    pt = PathTracer(job.scene, gpu_build=job.gpu_build, instancing=job.instancing)
    return pt, t0, time.perf_counter()


def _render(job: Job, pt: PathTracer, t0: float, t1: float) -> JobResult:
    """Render a built job; t0 / t1 are its build's start and end."""
    img = np.zeros((job.height, job.width, 3), np.float32)
    r0 = time.perf_counter()
    st = pt.render(job.camera, job.samples, job.depth, img, seed=job.seed, estimator=job.estimator,
                   rr_start=job.rr_start, max_paths=job.max_paths, passes_in_flight=job.passes_in_flight)
    t2 = time.perf_counter()
    return JobResult(-1, img, t2 - t0, t1 - t0, t2 - r0, st["paths"], st["extension_rays"] + st["shadow_rays"],
                     stats=st)


def run_job(job: Job, keep: Optional[list] = None) -> JobResult:
    pt, t0, t1 = _build(job)
    r = _render(job, pt, t0, t1)
    if keep is not None:
        keep.append(pt)
    return r


def estimated_cost(job: Job) -> float:
    """Rough seconds of one job on its own: the tree build (about 0.35 s per million triangles on
    the host, 20x less on the GPU) plus the render (about 1e9 path vertices per second)."""
    tris = sum(int(np.asarray(job.scene.meshes[o.mesh].indices).size) // 3 for o in job.scene.objects)
    build = tris * (0.35e-6 / (20.0 if job.gpu_build else 1.0))
    spp = 2 ** round(np.log2(max(job.samples, 1)))
    return build + job.width * job.height * spp * job.depth / 1e9


def run_batch(jobs: List[Job], workers: int = 4, on_done: Optional[Callable[[JobResult], None]] = None,
              order: str = "input", build_workers: int = 0):
    """Run every job; returns (results in job order, wall seconds of the batch). order: "input"
    starts jobs in list order; "lpt" starts the longest (estimated_cost) first, which shortens the
    batch when a few large jobs would otherwise start last.

    build_workers = 0: each of the `workers` threads builds and then renders its job. build_workers
    > 0: the builds run in a pool of their own, in the same order, ahead of the renders (each
    render waits for its own build only), so the GPU starts after the first build instead of
    after `workers` concurrent ones, and later builds overlap earlier renders."""
    if order not in ("input", "lpt"):
        raise ValueError("order must be 'input' or 'lpt'")
    if build_workers < 0:
        raise ValueError("build_workers must be >= 0")
    seq = list(range(len(jobs)))
    if order == "lpt":
        seq.sort(key=lambda i: -estimated_cost(jobs[i]))
    keep: list = []
    lock = threading.Lock()
    results: List[Optional[JobResult]] = [None] * len(jobs)
    b0 = time.perf_counter()
    bex = ThreadPoolExecutor(max_workers=build_workers) if build_workers > 0 else None
    builds = {i: bex.submit(_build, jobs[i]) for i in seq} if bex else {}

    def work(i):
        if bex:
            pt, t0, t1 = builds[i].result()
            r = _render(jobs[i], pt, t0, t1)
            s = t0 - b0
        else:
            s = time.perf_counter() - b0
            pt, t0, t1 = _build(jobs[i])
            r = _render(jobs[i], pt, t0, t1)
            with lock:
                keep.append(pt)
        r.index, r.started_s = i, s
        with lock:
            results[i] = r
        if on_done:
            on_done(r)

    try:
        with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
            list(ex.map(work, seq))
    finally:
        if bex:
            bex.shutdown(wait=True)
    wall = time.perf_counter() - b0
    del keep, builds
    return results, wall


def render_batch(items, max_paths: int = 0):
    """Renders several built jobs in one call (rdu_pathtracer_render_batch): every bounce of every
    job in one set of wavefront launches. items: (PathTracer, Job) pairs. Returns (images, each
    job's completion seconds from the call's start, render stats)."""
    import ctypes
    from ._lib import check, lib
    n = len(items)
    arr = (S.BatchJob * n)()
    images, cams = [], []
    for k, (pt, job) in enumerate(items):
        img = np.zeros((job.height, job.width, 3), np.float32)
        images.append(img)
        cams.append(S.camera_desc(job.camera))
        arr[k] = S.BatchJob(pt._h.value, ctypes.pointer(cams[-1]), None, job.samples, job.depth,
                            img.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), job.width, job.height, 3,
                            S.RenderOpts(job.seed, job.estimator, job.rr_start, 0))
    secs = np.zeros(n, np.float64)
    st = S.RenderStats()
    check(lib.rdu_pathtracer_render_batch(ctypes.cast(arr, ctypes.c_void_p), n, int(max_paths),
                                          secs.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
    return images, secs, {k: getattr(st, k) for k, _ in S.RenderStats._fields_}


def merge_groups(jobs: List[Job], groups: int) -> List[List[int]]:
    """Job indices longest first (estimated_cost), cut into at most `groups` consecutive non-empty
    groups of about equal estimated cost."""
    if groups < 1:
        raise ValueError("groups must be >= 1")
    seq = sorted(range(len(jobs)), key=lambda i: -estimated_cost(jobs[i]))
    cost = [estimated_cost(jobs[i]) for i in seq]
    total = sum(cost) or 1.0
    parts: List[List[int]] = [[] for _ in range(groups)]
    acc = 0.0
    for i, c in zip(seq, cost):
        parts[min(groups - 1, int(acc / total * groups))].append(i)
        acc += c
    return [g for g in parts if g]


def run_merged(jobs: List[Job], build_workers: int = 3, groups: int = 4, max_paths: int = 0):
    """The batch as a few renderer calls that each render many jobs at once
    (rdu_pathtracer_render_batch: every bounce of every job of the call in one set of wavefront
    launches). Jobs are taken longest first (estimated_cost) and cut into `groups` consecutive
    groups of about equal estimated cost; the scenes are built in that order on `build_workers`
    host threads, and each group's call starts, on a thread of its own, as soon as its own jobs are
    built, so that builds overlap earlier groups' renders and the groups' calls overlap each other
    on the GPU. Returns (results in job order, wall seconds); a job's latency runs from the batch
    start to its completion inside its group's call."""
    if build_workers < 1:
        raise ValueError("build_workers must be >= 1")
    parts = merge_groups(jobs, groups)
    seq = [i for g in parts for i in g]
    results: List[Optional[JobResult]] = [None] * len(jobs)
    b0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=build_workers) as bex:
        builds = {i: bex.submit(_build, jobs[i]) for i in seq}

        def render_group(members):
            built = [builds[i].result() for i in members]
            c0 = time.perf_counter() - b0
            images, secs, st = render_batch([(pt, jobs[i]) for (pt, _, _), i in zip(built, members)],
                                            max_paths=max_paths)
            for k, i in enumerate(members):
                pt, t0, t1 = built[k]
                spp = 2 ** round(np.log2(max(jobs[i].samples, 1)))
                r = JobResult(i, images[k], c0 + float(secs[k]), t1 - t0, float(secs[k]),
                              jobs[i].width * jobs[i].height * spp, 0, started_s=c0)
                results[i] = r
            results[members[0]].rays = int(st["extension_rays"] + st["shadow_rays"])
            results[members[0]].stats = st
            del built

        with ThreadPoolExecutor(max_workers=len(parts)) as rex:
            list(rex.map(render_group, parts))
        del builds
    wall = time.perf_counter() - b0
    return results, wall


def summary(results: List[JobResult], wall: float) -> dict:
    lat = np.array([r.latency_s for r in results])
    return {
        "jobs": len(results),
        "jobs_per_s": round(len(results) / wall, 3),
        "wall_s": round(wall, 3),
        "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
        "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
        "build_s_total": round(float(sum(r.build_s for r in results)), 3),
        "msamples_per_s": round(sum(r.paths for r in results) / wall / 1e6, 2),
        "mrays_per_s": round(sum(r.rays for r in results) / wall / 1e6, 2),
    }
