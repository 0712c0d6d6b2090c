"""Ray-cast benchmark: incoherent closest-hit MRays/s on one MI355X.

Workload (the ray-cast microbenchmark): a 1,000,000-triangle soup (centroids U[0,1]^3, edge
exp(-5 + 0.5 z), random rotation) with 2^23 incoherent rays (origins U[0,1]^3, directions uniform on
the sphere) and 2^23 coherent pinhole primaries (4096 x 2048), Rendu's default interval (1e-4, 1e8).
One step = one closest-hit pass (rdu_raycaster_intersect, near-first order, RDU_INCOHERENT hint so
the library orders the rays itself inside the timed region) over the 2^23 incoherent rays, rays
and hits resident in HBM. The traversal's launch parameters (refill threshold, ray tile, workgroups
per CU) are autotuned on the same batch during warm-up.

Alongside: coherent closest hit, any hit, Rendu's exact visiting order, and the small path-tracing
workload (Cornell box + 3 icospheres, 256x256, 16 spp, depth 5) in samples/s. The CPU baseline is
the oracle's restatement of Rendu's Raycaster (same tree, same traversal, forParallel's thread
policy) on a bounded sample of the same incoherent rays, on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_TRIS = 1_000_000
N_HALF = 1 << 23
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=1 << 23, help="rays in the CPU baseline sample")
    ap.add_argument("--cpu-passes", type=int, default=2, help="passes over the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--main-only", action="store_true", help="only the timed incoherent closest-hit passes")
    ap.add_argument("--refill", type=int, default=0,
                    help="fixed launch parameters instead of autotuning (profiling runs: one kernel "
                         "configuration): refill threshold, with --chunk and --blocks-per-cu")
    ap.add_argument("--chunk", type=int, default=0, help="fixed ray tile with --refill (0 = default)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="fixed workgroups per CU with --refill (0 = all)")
    ap.add_argument("--tree", choices=("rendu", "gpu"), default="rendu",
                    help="tree of the timed queries: Rendu's host build (updateHierarchy) or the GPU linear BVH")
    ap.add_argument("--no-cache-experiment", action="store_true", help="skip the radiance-cache comparison")
    ap.add_argument("--cache-ref-spp", type=int, default=4096, help="samples of the reference frame")
    ap.add_argument("--profile-frame", action="store_true",
                    help="only the single frame (1920x1080, 64 spp, depth 8), plain then with a fresh radiance "
                         "cache, after one untimed render of each and a 1 s pause: the workload of the "
                         "path-tracer kernel profiles")
    ap.add_argument("--profile-frame-instanced", action="store_true",
                    help="only the single frame, flattened then as an instanced scene (one untimed render of "
                         "each, a 1 s pause, then one timed render of each): the workload of the instanced "
                         "frame's kernel profile")
    ap.add_argument("--jobs-only", action="store_true",
                    help="only the 64 batched jobs: a warm-up batch, a 1.5 s pause, then the timed batch with "
                         "2 build and 6 render threads (the workload of the job-batch kernel trace)")
    ap.add_argument("--jobs-one-call", action="store_true",
                    help="only the 64 jobs, built, then rendered in one rdu_pathtracer_render_batch call: a "
                         "warm-up call, a 1.5 s pause, then the timed call (the workload of its kernel trace)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device three times: the reference reading of the "
                         "FETCH_SIZE / WRITE_SIZE counters that converts them to bytes")
    return ap.parse_args()


def copy_calibration():
    import torch
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copied_bytes_per_dispatch": a.numel() * 4, "dispatches": 3}), flush=True)


def profile_frame():
    from rendu_amd import scene as S
    from rendu_amd import scenes
    from rendu_amd.pathtracer import PathTracer
    from rendu_amd.rcache import RadianceCache
    scn, cam_of = scenes.terrain_spheres()
    pt = PathTracer(scn)
    cam = cam_of(1920, 1080)
    img = np.zeros((1080, 1920, 3), np.float32)
    # Warm-up of both renders (kernel loads, the working sets' first allocation), then a 1 s pause
    # that separates the timed renders in a trace.
    pt.render(cam, 64, 8, img, seed=1, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    pt.render_cached(RadianceCache(*S.scene_bounds(scn), seed=4), cam, 64, 8, img, seed=1,
                     estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    time.sleep(1.0)
    img[:] = 0
    t0 = time.perf_counter()
    st = pt.render(cam, 64, 8, img, seed=2, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    t1 = time.perf_counter()
    img[:] = 0
    cst = pt.render_cached(RadianceCache(*S.scene_bounds(scn), seed=5), cam, 64, 8, img, seed=2,
                           estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    t2 = time.perf_counter()
    print(json.dumps({"plain_s": round(t1 - t0, 4), "cached_s": round(t2 - t1, 4), "paths": st["paths"],
                      "cache_queries": cst["cache_queries"], "train_steps": cst["train_steps"]}), flush=True)


def profile_frame_instanced():
    from rendu_amd import scene as S
    from rendu_amd import scenes
    from rendu_amd.pathtracer import PathTracer
    scn, cam_of = scenes.terrain_spheres()
    cam = cam_of(1920, 1080)
    img = np.zeros((1080, 1920, 3), np.float32)
    t0 = time.perf_counter()
    flat = PathTracer(scn)
    t1 = time.perf_counter()
    inst = PathTracer(scn, instancing=True)
    t2 = time.perf_counter()
    for pt in (flat, inst):
        pt.render(cam, 64, 8, img, seed=1, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    time.sleep(1.0)
    out = {"flattened_setup_s": round(t1 - t0, 3), "instanced_setup_s": round(t2 - t1, 3)}
    for name, pt in (("flattened", flat), ("instanced", inst)):
        img[:] = 0
        w0 = time.perf_counter()
        pt.render(cam, 64, 8, img, seed=2, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
        out[name + "_s"] = round(time.perf_counter() - w0, 4)
        out[name + "_mean"] = round(float(img.mean()), 6)
    print(json.dumps(out), flush=True)


def job_list_64():
    from rendu_amd import jobs as J
    from rendu_amd import scenes
    out = []
    for s in range(64):
        scn_j, cam_j, w_j, h_j, spp_j, depth_j = scenes.random_job(s)
        out.append(J.Job(scn_j, cam_j(w_j, h_j), w_j, h_j, spp_j, depth_j, seed=s))
    return out


def jobs_only():
    from rendu_amd import jobs as J
    job_list = job_list_64()
    J.run_batch(job_list, workers=8, order="lpt")
    time.sleep(1.5)
    res, wall = J.run_batch(job_list, workers=6, order="lpt", build_workers=2)
    print(json.dumps(dict(J.summary(res, wall), workers=6, build_workers=2, order="longest first")), flush=True)


def jobs_one_call():
    from rendu_amd import jobs as J
    from rendu_amd.pathtracer import PathTracer
    job_list = job_list_64()
    items = [(PathTracer(j.scene), j) for j in job_list]
    J.render_batch(items)
    time.sleep(1.5)
    t0 = time.perf_counter()
    _, secs, st = J.render_batch(items)
    wall = time.perf_counter() - t0
    print(json.dumps({"jobs": len(items), "render_wall_s": round(wall, 4), "jobs_per_s_render_only":
                      round(len(items) / wall, 2), "passes": st["passes"], "paths": st["paths"],
                      "latency_p50_s": round(float(np.percentile(secs, 50)), 4)}), flush=True)


def cache_experiment(pt, scn, cam, ref_spp):
    """Radiance cache on the single frame (1920 x 1080, depth 8, RR from bounce 3): RMSE against a
    ref_spp plain render, plain at 64 spp against cached renders (fresh cache each, online training
    included in the wall time) at 16 / 32 / 64 spp, with the trust test and with every vertex from
    the termination bounce (2, 1, or the primary vertex) ending in the cache; also a cache warmed
    by one earlier 64-spp frame.
    Reports the wall-time ratio of the fastest cached render whose RMSE reaches plain 64 spp, and
    what the tail past bounce 2 holds (its share of the mean and of the rays, and the time of the
    render truncated there, the floor for a cache that terminates paths at bounce 2)."""
    from rendu_amd import scene as S
    from rendu_amd.rcache import RadianceCache
    ext = S.ESTIMATOR_EXTENDED
    h, w = 1080, 1920
    lo, hi = S.scene_bounds(scn)
    ref = np.zeros((h, w, 3), np.float32)
    t0 = time.perf_counter()
    pt.render(cam, ref_spp, 8, ref, seed=1000, estimator=ext, rr_start=3)
    ref_s = time.perf_counter() - t0
    rmse = lambda a: float(np.sqrt(np.mean((a.astype(np.float64) - ref) ** 2)))

    def plain(spp, depth=8):
        img = np.zeros((h, w, 3), np.float32)
        t = time.perf_counter()
        pt.render(cam, spp, depth, img, seed=7, estimator=ext, rr_start=3)
        st = pt.last_stats()
        return {"spp": spp, "s": round(time.perf_counter() - t, 4), "rmse": round(rmse(img), 6),
                "mean": round(float(img.mean()), 5),
                "rays_per_sample": round((st["extension_rays"] + st["shadow_rays"]) / (w * h * spp), 4)}

    def cached(spp, cache=None, **kw):
        cache = cache or RadianceCache(lo, hi, seed=5)
        img = np.zeros((h, w, 3), np.float32)
        t = time.perf_counter()
        st = pt.render_cached(cache, cam, spp, 8, img, seed=7, estimator=ext, rr_start=3, **kw)
        out = {"spp": spp, "s": round(time.perf_counter() - t, 4), "rmse": round(rmse(img), 6),
               "mean": round(float(img.mean()), 5), "cache_queries": st["cache_queries"],
               "train_steps": st["train_steps"], "trusted_cells": st["trusted_cells"]}
        out.update({k: v for k, v in kw.items()})
        return out

    pt.render(cam, 1, 8, np.zeros((h, w, 3), np.float32), seed=1, estimator=ext, rr_start=3)  # warm-up
    p64 = plain(64)
    runs = {"plain": [plain(16), plain(32), p64]}
    # The tracer without the network: paths end after bounce 2 with nothing added (depth 3). Its time
    # bounds from below what any cache terminating at bounce 2 can reach at this spp.
    trunc = plain(64, depth=3)
    runs["plain_truncated_after_bounce2"] = [trunc]
    runs["cached"] = [cached(s) for s in (16, 32, 64)]
    runs["cached_always"] = [cached(s, always_trust=True) for s in (16, 32, 64)]
    runs["cached_always_b1"] = [cached(s, always_trust=True, terminate_bounce=1) for s in (16, 32, 64)]
    runs["cached_always_primary"] = [cached(s, always_trust=True, from_primary=True) for s in (16, 32)]
    warm = RadianceCache(lo, hi, seed=5)
    pt.render_cached(warm, cam, 64, 8, np.zeros((h, w, 3), np.float32), seed=3, estimator=ext, rr_start=3)
    runs["warm_cache"] = [cached(s, cache=warm) for s in (16, 32)]
    best = None
    for key in ("cached", "cached_always", "cached_always_b1", "cached_always_primary", "warm_cache"):
        for r in runs[key]:
            if r["rmse"] <= p64["rmse"] and (best is None or r["s"] < best[1]["s"]):
                best = (key, r)
    return {"ref_spp": ref_spp, "ref_s": round(ref_s, 3), "ref_mean": round(float(ref.mean()), 5), "runs": runs,
            "plain64_rmse": p64["rmse"], "plain64_s": p64["s"],
            "tail_after_bounce2": {"share_of_mean": round(1.0 - trunc["mean"] / p64["mean"], 4),
                                   "share_of_rays": round(1.0 - trunc["rays_per_sample"] / p64["rays_per_sample"], 4),
                                   "time_ratio_truncated_over_plain64": round(trunc["s"] / p64["s"], 3)},
            "fastest_reaching_plain64_rmse": None if best is None else
            {"mode": best[0], **best[1], "time_ratio": round(best[1]["s"] / p64["s"], 3)}}


def cache_kernels():
    """Radiance-cache kernels alone (HIP events on the launch stream): inference over 2^20 queries and
    training steps of 16,384 pairs (forward, backward, sorted table-gradient reduction, weight
    gradients, Adam), against the dense bf16 MFMA peak. Issued MLP work: 40,960 FLOP per query
    forward (48x64 + 4x64x64 + 16x64 multiply-adds), about 3x that per training pair."""
    import torch
    from rendu_amd.rcache import RadianceCache, make_queries
    rng = np.random.default_rng(4)
    cache = RadianceCache([0, 0, 0], [1, 1, 1], seed=3)
    nq = 1 << 20
    q = torch.from_numpy(make_queries(rng.random((nq, 3)).astype(np.float32))).cuda()
    out = torch.empty((nq, 4), dtype=torch.float32, device="cuda")
    nt = 16384
    cent = rng.random((64, 3))
    pos = (cent[rng.integers(0, 64, nt)] + 0.02 * rng.standard_normal((nt, 3))).clip(0, 1).astype(np.float32)
    qt = torch.from_numpy(make_queries(pos)).cuda()
    tt = torch.from_numpy(rng.random((nt, 4)).astype(np.float32)).cuda()
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def timed(fn, reps):
        fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    inf_ms = timed(lambda: cache.infer_device(q.data_ptr(), nq, out.data_ptr(), stream=sh), 20)
    tr_ms = timed(lambda: cache.train_device(qt.data_ptr(), tt.data_ptr(), nt, steps=1, stream=sh), 20)
    flop_q = 2 * (48 * 64 + 4 * 64 * 64 + 16 * 64)
    inf_tf = nq * flop_q / (inf_ms * 1e-3) / 1e12
    tr_tf = nt * 3 * flop_q / (tr_ms * 1e-3) / 1e12
    return {"infer_2e20_ms": round(inf_ms, 4), "infer_gqueries_per_s": round(nq / (inf_ms * 1e-3) / 1e9, 3),
            "infer_tflops_issued": round(inf_tf, 2), "train_step_16k_ms": round(tr_ms, 4),
            "train_mpairs_per_s": round(nt / (tr_ms * 1e-3) / 1e6, 2), "train_tflops_issued": round(tr_tf, 2),
            "bf16_dense_peak_tflops": 2500.0, "infer_frac_of_peak": round(inf_tf / 2500.0, 4),
            "table_gather_bytes_per_query": 512}


def cache_experiment_cornell():
    """Radiance cache on the closed Cornell box (256 x 256, depth 8, RR from bounce 3; most light is
    indirect): RMSE against an 8192-spp plain render, plain at 64 / 128 spp against cached renders
    at 64 spp (fresh cache each; 2 spp per pass so that the cache trains between passes)."""
    from rendu_amd import scene as S
    from rendu_amd import scenes
    from rendu_amd.pathtracer import PathTracer
    from rendu_amd.rcache import RadianceCache
    ext = S.ESTIMATOR_EXTENDED
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=4)
    lo, hi = S.scene_bounds(scn)
    pt = PathTracer(scn)
    w = h = 256
    cam = cam_of(w, h)
    ref = np.zeros((h, w, 3), np.float32)
    pt.render(cam, 8192, 8, ref, seed=1000, estimator=ext, rr_start=3)
    rmse = lambda a: round(float(np.sqrt(np.mean((a.astype(np.float64) - ref) ** 2))), 6)
    out = {"ref_spp": 8192}
    for spp in (64, 128):
        img = np.zeros((h, w, 3), np.float32)
        t = time.perf_counter()
        pt.render(cam, spp, 8, img, seed=7, estimator=ext, rr_start=3)
        out[f"plain{spp}"] = {"s": round(time.perf_counter() - t, 4), "rmse": rmse(img)}
    for name, kw in (("cached64_trust_test_b2", {}), ("cached64_always_b1", {"always_trust": True, "terminate_bounce": 1})):
        img = np.zeros((h, w, 3), np.float32)
        t = time.perf_counter()
        st = pt.render_cached(RadianceCache(lo, hi, seed=5), cam, 64, 8, img, seed=7, estimator=ext, rr_start=3,
                              max_paths=w * h * 2, **kw)
        out[name] = {"s": round(time.perf_counter() - t, 4), "rmse": rmse(img), "cache_queries": st["cache_queries"],
                     "train_steps": st["train_steps"]}
    out["variance_ratio_plain64_over_cached64_always_b1"] = round(
        (out["plain64"]["rmse"] / out["cached64_always_b1"]["rmse"]) ** 2, 3)
    return out


def host_cpu():
    """Logical CPUs and the model name of the host (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count() or 1, "model": model}


def cpu_jobs(orc, threads):
    """The 64 batched jobs on the host cores, one after another as Rendu would run them (its
    PathTracer renders one scene per call, over forParallel's threads): each job's scene is built
    in full (the oracle's Raycaster::updateHierarchy, single-threaded as in Rendu) and rendered at a
    quarter of its width and height with 4 samples per pixel; the render time is scaled by the
    job's pixels x spp over the rendered ones (per-sample cost is the same scene, camera and depth).
    Reports the serial batch's jobs/s and the per-job time distribution."""
    from rendu_amd import scene as S
    from rendu_amd import scenes
    per_job, builds, renders = [], 0.0, 0.0
    for s in range(64):
        scn, cam, w, h, spp, depth = scenes.random_job(s)
        t0 = time.perf_counter()
        opt = orc.OraclePathTracer(scn)
        b = time.perf_counter() - t0
        ws, hs = max(1, w // 4), max(1, h // 4)
        img = np.zeros((hs, ws, 3), np.float32)
        t0 = time.perf_counter()
        opt.render(S.camera_desc(cam(ws, hs)), 4, depth, img, S.RenderOpts(s, 0, 3, 0), threads=threads)
        r = (time.perf_counter() - t0) * (w * h * 2 ** round(np.log2(spp))) / (ws * hs * 4)
        per_job.append(b + r)
        builds += b
        renders += r
        del opt
    tot = sum(per_job)
    return {"jobs": 64, "jobs_per_s": round(64 / tot, 4), "serial_batch_s": round(tot, 2),
            "job_s_p50": round(float(np.percentile(per_job, 50)), 3),
            "job_s_p99": round(float(np.percentile(per_job, 99)), 3),
            "build_s_total": round(builds, 2), "render_s_total_scaled": round(renders, 2), "threads": threads,
            "sample": "every job built in full; rendered at 1/4 x 1/4 resolution, 4 spp, scaled to its "
                      "pixels x spp (power of two, as Rendu rounds)"}


def load_pmc():
    """The committed rocprofv3 --pmc summary of the dominant kernel (None when absent)."""
    path = os.path.join(ROOT, "profiles", "trace_closest_soup1m_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def load_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3 --pmc summary."""
    pmc = load_pmc()
    return pmc.get("hbm_bytes_per_launch") if pmc else None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; how the path "
              "would shard is a design note of the repository, not a measurement", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.profile_frame:
        profile_frame()
        return
    if args.profile_frame_instanced:
        profile_frame_instanced()
        return
    if args.jobs_one_call:
        jobs_one_call()
        return
    if args.jobs_only:
        jobs_only()
        return

    import torch
    from rendu_amd import workloads
    from rendu_amd.raycaster import Raycaster

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.init()
    t0 = time.perf_counter()
    mesh = workloads.triangle_soup(N_TRIS, seed=1)
    rays = workloads.soup_rays(N_HALF, N_HALF, seed=2)
    gen_s = time.perf_counter() - t0
    rc = Raycaster()
    rc.addMesh(mesh)
    t0 = time.perf_counter()
    if args.tree == "gpu":
        rc.updateHierarchyGpu()
        torch.cuda.synchronize()
    else:
        rc.updateHierarchy()
    build_s = time.perf_counter() - t0
    info = rc.info()

    coherent = np.ascontiguousarray(rays[:N_HALF])
    incoherent = np.ascontiguousarray(rays[N_HALF:])
    d_inc = torch.from_numpy(incoherent.view(np.uint8)).cuda()
    d_coh = torch.from_numpy(coherent.view(np.uint8)).cuda()
    d_hits = torch.empty(N_HALF * 24, dtype=torch.uint8, device="cuda")
    d_occ = torch.empty(N_HALF, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def closest(d, incoh=True, ref=False):
        rc.intersect_device(d.data_ptr(), N_HALF, d_hits.data_ptr(), sh, reference_order=ref, incoherent=incoh)

    def anyhit(d, incoh=True):
        rc.intersect_device(d.data_ptr(), N_HALF, d_occ.data_ptr(), sh, any_hit=True, incoherent=incoh)

    def timed(fn, steps, warmup, after_warmup=None):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        if after_warmup:
            after_warmup()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        w0 = time.perf_counter()
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        return wall, float(np.mean([a.elapsed_time(b) for a, b in ev]))

    if args.refill:
        rc.set_tuning(args.refill, args.chunk, args.blocks_per_cu, incoherent=True)
        tune = {"refill_min": args.refill, "chunk": args.chunk, "blocks_per_cu": args.blocks_per_cu, "fixed": True}
    else:
        tune = rc.autotune(rays_ptr=d_inc.data_ptr(), n=N_HALF, incoherent=True, stream=sh)
    # The traversal kernel alone is timed with HIP events the library records on the launch stream.
    rc.set_kernel_timing(True)
    wall, k_ms = timed(lambda: closest(d_inc), args.steps, args.warmup, after_warmup=rc.kernel_time)
    trav_total, trav_n = rc.kernel_time()
    rc.set_kernel_timing(False)
    trav_ms = trav_total / max(1, trav_n)
    value = args.steps * N_HALF / wall / 1e6
    if args.main_only:  # profiling runs: the timed incoherent closest-hit passes only
        print(json.dumps({"value": round(value, 2), "step_ms": round(k_ms, 4), "traversal_kernel_ms": round(trav_ms, 4),
                          "tune": tune}), flush=True)
        return
    hits = d_hits.cpu().numpy().view(np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("hit", "<u4"),
                                                ("mesh_id", "<u4"), ("local_id", "<u4")]))
    hit_rate = float((hits["hit"] == 1).mean())

    steps2 = max(3, args.steps // 2)
    rate = lambda w, k: round(k * N_HALF / w / 1e6, 2)
    tune_c = rc.autotune(rays_ptr=d_coh.data_ptr(), n=N_HALF, incoherent=False, stream=sh)
    extra = {
        "coherent_closest_mrays": rate(timed(lambda: closest(d_coh, incoh=False), steps2, 1)[0], steps2),
        "incoherent_closest_input_order_mrays": rate(timed(lambda: closest(d_inc, incoh=False), steps2, 1)[0], steps2),
        "incoherent_any_mrays": rate(timed(lambda: anyhit(d_inc), steps2, 1)[0], steps2),
        "coherent_any_mrays": rate(timed(lambda: anyhit(d_coh, incoh=False), steps2, 1)[0], steps2),
    }
    # Host buffers in and out (PCIe-inclusive: 32 B per ray up, 24 B per hit down); not the metric.
    rc.intersect_batch(incoherent[: 1 << 20], incoherent=True)
    t0 = time.perf_counter()
    for _ in range(2):
        rc.intersect_batch(incoherent, incoherent=True)
    extra["incoherent_closest_host_buffers_mrays"] = round(2 * N_HALF / (time.perf_counter() - t0) / 1e6, 2)
    if args.tree == "rendu":  # Rendu's visiting order exists only on Rendu's tree
        extra["incoherent_closest_reference_order_mrays"] = rate(timed(lambda: closest(d_inc, ref=True), 2, 1)[0], 2)
    sub = np.ascontiguousarray(incoherent[: 1 << 20])
    nv, tt = rc.trace_stats(sub)
    extra["near_first_wide_nodes_per_ray"] = round(nv / sub.shape[0], 2)
    extra["near_first_triangle_tests_per_ray"] = round(tt / sub.shape[0], 2)
    # Demand bytes of this traversal order (112 B per 4-wide node fetched, 48 B per triangle
    # record), against the fabric traffic and L2 hit rate of the committed rocprofv3 counters.
    demand = (nv * 112 + tt * 48) / sub.shape[0]
    extra["demand_bytes_per_ray"] = round(demand, 1)
    extra["demand_gbs"] = round(demand * N_HALF / (trav_ms * 1e-3) / 1e9, 1)
    pmc = load_pmc()
    if pmc:
        extra["fabric_gbs_from_counters"] = round(pmc["hbm_bytes_per_launch"] / (trav_ms * 1e-3) / 1e9, 1)
        extra["l2_hit_rate_from_counters"] = pmc.get("tcc_hit_rate")

    # The other tree over the same soup: its build time and the same query.
    rg = Raycaster()
    rg.addMesh(mesh)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if args.tree == "rendu":
        rg.updateHierarchyGpu()
    else:
        rg.updateHierarchy()
    torch.cuda.synchronize()
    other_build = time.perf_counter() - t0
    rg.autotune(rays_ptr=d_inc.data_ptr(), n=N_HALF, incoherent=True, stream=sh)
    wall_g = timed(lambda: rg.intersect_device(d_inc.data_ptr(), N_HALF, d_hits.data_ptr(), sh, incoherent=True),
                   steps2, 1)[0]
    nv2, tt2 = rg.trace_stats(sub)
    extra["other_tree"] = {"tree": "gpu linear BVH" if args.tree == "rendu" else "rendu (host build)",
                           "build_s": round(other_build, 4), "incoherent_closest_mrays": rate(wall_g, steps2),
                           "wide_nodes": rg.info()["wide_nodes"], "wide_nodes_per_ray": round(nv2 / sub.shape[0], 2),
                           "triangle_tests_per_ray": round(tt2 / sub.shape[0], 2)}
    del rg

    # Small path-tracing workload on the GPU (samples/s).
    from rendu_amd import scenes
    from rendu_amd.pathtracer import PathTracer
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=4)
    pt = PathTracer(scn)
    img = np.zeros((256, 256, 3), np.float32)
    pt.render(cam_of(256, 256), 16, 5, img, seed=1)
    best = None
    for rep in range(3):
        img[:] = 0
        st = pt.render(cam_of(256, 256), 16, 5, img, seed=1 + rep)
        best = st if best is None or st["seconds"] < best["seconds"] else best
    extra["cornell_256_16spp_depth5_msamples_per_s"] = round(best["paths"] / best["seconds"] / 1e6, 2)
    extra["cornell_256_16spp_depth5_ms"] = round(best["seconds"] * 1e3, 3)
    extra["cornell_rays_per_sample"] = round((best["extension_rays"] + best["shadow_rays"]) / best["paths"], 3)
    # Image error of the 16-spp render: per-pixel RMSE against a 4096-spp render of the same scene
    # (another seed, so the two are independent estimates).
    ref = np.zeros((256, 256, 3), np.float32)
    pt.render(cam_of(256, 256), 4096, 5, ref, seed=4096)
    img[:] = 0
    pt.render(cam_of(256, 256), 16, 5, img, seed=1)
    extra["cornell_256_16spp_depth5_rmse_vs_4096spp"] = round(float(np.sqrt(np.mean((img - ref) ** 2))), 5)
    extra["cornell_256_mean"] = round(float(ref.mean()), 5)
    del pt, ref
    # End to end (PathTracer construction = raycaster build + upload, then the render).
    t0 = time.perf_counter()
    pt = PathTracer(scn)
    img[:] = 0
    pt.render(cam_of(256, 256), 16, 5, img, seed=9)
    extra["cornell_256_16spp_depth5_end_to_end_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    del pt
    # HIP graph replay of render passes, eager against replayed: the Cornell workload in 16 passes
    # of one sample each, and a progressive loop of 16 one-sample renders with new seeds (the way an
    # interactive viewer calls render).
    graph = {}
    img1 = np.zeros((256, 256, 3), np.float32)
    cam = cam_of(256, 256)
    for mode in ("eager", "graph"):
        ptg = PathTracer(scn)
        ptg.set_graph(mode == "graph")
        for _ in range(2):  # warm-up (the graph is captured here)
            img[:] = 0
            ptg.render(cam, 16, 5, img, seed=1, max_paths=256 * 256)
        times = []
        for rep in range(3):
            img[:] = 0
            st = ptg.render(cam, 16, 5, img, seed=1, max_paths=256 * 256)
            times.append(st["seconds"])
        graph[f"cornell_16_passes_{mode}_ms"] = round(min(times) * 1e3, 3)
        graph[f"cornell_16_passes_{mode}_replayed"] = st["graph_passes"]
        img1[:] = 0
        t0 = time.perf_counter()
        for k in range(16):
            ptg.render(cam, 1, 5, img1, seed=100 + k)
        graph[f"progressive_16x1spp_{mode}_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        del ptg
    extra["graph_replay"] = graph

    # Single frame: fBm terrain + 2000 icospheres (3.08M triangles), lat-long sky, 1920x1080,
    # 64 spp, depth 8, Russian roulette from bounce 3.
    scn3, cam3 = scenes.terrain_spheres()
    t0 = time.perf_counter()
    pt3 = PathTracer(scn3)
    pt3.raycaster()
    setup3 = time.perf_counter() - t0
    img3 = np.zeros((1080, 1920, 3), np.float32)
    from rendu_amd.scene import ESTIMATOR_EXTENDED
    pt3.render(cam3(1920, 1080), 1, 8, img3, seed=1, estimator=ESTIMATOR_EXTENDED, rr_start=3)  # warm-up
    img3[:] = 0
    w0 = time.perf_counter()
    st3 = pt3.render(cam3(1920, 1080), 64, 8, img3, seed=2, estimator=ESTIMATOR_EXTENDED, rr_start=3)
    frame_s = time.perf_counter() - w0
    rays3 = st3["extension_rays"] + st3["shadow_rays"]
    extra["frame_1080p_64spp_depth8_s"] = round(frame_s, 4)
    extra["frame_1080p_msamples_per_s"] = round(st3["paths"] / frame_s / 1e6, 2)
    extra["frame_1080p_mrays_per_s"] = round(rays3 / frame_s / 1e6, 2)
    extra["frame_1080p_rays_per_sample"] = round(rays3 / st3["paths"], 3)
    extra["frame_scene_setup_s"] = round(setup3, 3)
    # The pass autotuner on this frame (one 2-spp pass timed per candidate), then the frame with the
    # chosen parameters, and with graph replay of its passes.
    t0 = time.perf_counter()
    tune3 = pt3.autotune(cam3(1920, 1080), 64, 8, 1920, 1080, seed=2, estimator=ESTIMATOR_EXTENDED, rr_start=3)
    tune3_s = time.perf_counter() - t0
    img3[:] = 0
    w0 = time.perf_counter()
    pt3.render(cam3(1920, 1080), 64, 8, img3, seed=2, estimator=ESTIMATOR_EXTENDED, rr_start=3)
    tuned_s = time.perf_counter() - w0
    pt3.set_graph(True)
    img3[:] = 0
    pt3.render(cam3(1920, 1080), 64, 8, img3, seed=3, estimator=ESTIMATOR_EXTENDED, rr_start=3)  # captures
    img3[:] = 0
    w0 = time.perf_counter()
    stg = pt3.render(cam3(1920, 1080), 64, 8, img3, seed=2, estimator=ESTIMATOR_EXTENDED, rr_start=3)
    graph_s = time.perf_counter() - w0
    pt3.set_graph(False)
    extra["frame_autotune"] = {k: tune3[k] for k in ("refill_min", "chunk", "blocks_per_cu", "candidates")}
    extra["frame_autotune"].update({"pass_default_ms": round(tune3["default_ms"], 3),
                                    "pass_best_ms": round(tune3["best_ms"], 3), "tuning_s": round(tune3_s, 3),
                                    "frame_tuned_s": round(tuned_s, 4), "frame_tuned_graph_s": round(graph_s, 4),
                                    "graph_passes": stg["graph_passes"], "passes": stg["passes"]})
    if not args.no_cache_experiment:
        extra["cache_1080p"] = cache_experiment(pt3, scn3, cam3(1920, 1080), args.cache_ref_spp)
        extra["cache_cornell_256"] = cache_experiment_cornell()
    extra["cache_kernels"] = cache_kernels()
    # The same frame as an instanced scene (one tree per mesh, a top level over the 2001 objects;
    # an extension whose hits differ from Rendu's in the last bits): set-up, render time, and its
    # image against the flattened render of the same seed.
    ref3 = np.zeros_like(img3)
    pt3.render(cam3(1920, 1080), 64, 8, ref3, seed=2, estimator=ESTIMATOR_EXTENDED, rr_start=3)
    del pt3
    t0 = time.perf_counter()
    pti = PathTracer(scn3, instancing=True)
    setup_i = time.perf_counter() - t0
    pti.render(cam3(1920, 1080), 1, 8, img3, seed=1, estimator=ESTIMATOR_EXTENDED, rr_start=3)  # warm-up
    img3[:] = 0
    w0 = time.perf_counter()
    pti.render(cam3(1920, 1080), 64, 8, img3, seed=2, estimator=ESTIMATOR_EXTENDED, rr_start=3)
    inst_s = time.perf_counter() - w0
    same = np.abs(img3 - ref3) <= 1e-4 * np.abs(ref3) + 1e-5
    extra["frame_instanced"] = {"setup_s": round(setup_i, 3), "frame_s": round(inst_s, 4),
                                "channels_within_1e-4_of_flattened": round(float(same.mean()), 4),
                                "mean": round(float(img3.mean()), 6), "flattened_mean": round(float(ref3.mean()), 6)}
    del pti, ref3

    # Batched jobs: 64 seeded scenes (1e4..5e5 triangles, 128^2..1024^2 pixels, 4..256 spp, depth 6),
    # each built (host BVH + upload) and rendered; one HIP stream per scene.
    from rendu_amd import jobs as J
    job_list = job_list_64()
    # Throughput configuration: longest estimated job first, host builds in a pool of 2 threads
    # running ahead of 6 render threads. Reported beside it: builds inside each job's own thread
    # (8 threads; lower p50 latency), and 4 threads in list order (lower p99 latency).
    # Warm-up: one whole batch (kernel loads, and the library's memory pools grow to the batch's
    # peak once), so that every schedule below is timed in the steady state of a job service.
    # The throughput configuration is timed three times; the line reports the median batch (the
    # batch-to-batch spread of a concurrent schedule is several jobs/s) with every wall listed.
    J.run_batch(job_list, workers=8, order="lpt")
    runs = []
    for _ in range(3):
        res_j, wall_j = J.run_batch(job_list, workers=6, order="lpt", build_workers=2)
        runs.append((wall_j, J.summary(res_j, wall_j)))
        del res_j
    runs.sort(key=lambda r: r[0])
    extra["jobs64"] = dict(runs[1][1], workers=6, build_workers=2, order="longest first",
                           walls_s=[round(r[0], 3) for r in runs], batch="median of 3")
    res_j, wall_j = J.run_batch(job_list, workers=8, order="lpt")
    extra["jobs64_build_in_job"] = dict(J.summary(res_j, wall_j), workers=8, order="longest first")
    del res_j
    res_j, wall_j = J.run_batch(job_list, workers=4)
    extra["jobs64_4_threads_list_order"] = J.summary(res_j, wall_j)
    del res_j
    # The same jobs as 4 renderer calls of 16 jobs each (rdu_pathtracer_render_batch: one set of
    # wavefront launches per bounce for all jobs of a call), started as their builds (3 threads) end.
    res_j, wall_j = J.run_merged(job_list, build_workers=3, groups=4)
    extra["jobs64_merged_calls"] = dict(J.summary(res_j, wall_j), calls=4, build_workers=3)
    del res_j
    # Render time alone of all 64 built jobs in one rdu_pathtracer_render_batch call (warm-up call
    # first), beside the concurrent single renders of the same built jobs on 6 threads.
    from rendu_amd.pathtracer import PathTracer as _PT
    built = [(_PT(j.scene), j) for j in job_list]
    J.render_batch(built)
    t0 = time.perf_counter()
    _, _, st_b = J.render_batch(built)
    one_call = time.perf_counter() - t0

    def _solo(item):
        p, j = item
        im = np.zeros((j.height, j.width, 3), np.float32)
        p.render(j.camera, j.samples, j.depth, im, seed=j.seed, passes_in_flight=1)

    order = sorted(built, key=lambda it: -J.estimated_cost(it[1]))
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=6) as ex:
        list(ex.map(_solo, order))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=6) as ex:
        list(ex.map(_solo, order))
    solo = time.perf_counter() - t0
    extra["jobs64_render_only"] = {"one_call_s": round(one_call, 4), "one_call_passes": st_b["passes"],
                                   "concurrent_single_renders_6_threads_s": round(solo, 4)}
    del built, order
    for j in job_list:
        j.gpu_build = True
    J.run_batch(job_list, workers=8, order="lpt")  # warm-up of the GPU build's own temporaries
    res_j, wall_j = J.run_batch(job_list, workers=6, order="lpt", build_workers=2)
    extra["jobs64_gpu_build"] = dict(J.summary(res_j, wall_j), workers=6, build_workers=2, order="longest first")
    # The same jobs as instanced scenes (each job's spheres share one tree; host builds).
    for j in job_list:
        j.gpu_build = False
        j.instancing = True
    J.run_batch(job_list, workers=8, order="lpt")  # warm-up
    res_i, wall_i = J.run_batch(job_list, workers=6, order="lpt", build_workers=2)
    extra["jobs64_instanced"] = dict(J.summary(res_i, wall_i), workers=6, build_workers=2, order="longest first")
    for j in job_list:
        j.instancing = False
    del res_j

    # Algorithmic (compulsory) bytes per traversal launch, the per-ray figure of the workload's
    # definition: each ray read once (32 B), each hit written once (16 B: t, u, v, triangle id), and
    # the scene once per launch (Rendu's nodes at 32 B and triangles at 48 B: the box-and-children /
    # vertex records a traversal must touch), here 48 + (nodes * 32 + triangles * 48) / 2^23 B per ray.
    # The bytes this layout actually moves per ray (order index 4 B + ray 32 B + 24-B hit + the
    # 4-wide node and triangle arrays once) are reported beside it.
    alg_per_ray = 32 + 16 + (info["nodes"] * 32 + info["triangles"] * 48) / N_HALF
    alg_bytes = N_HALF * alg_per_ray
    achieved = alg_bytes / (trav_ms * 1e-3) / 1e9
    layout_per_ray = 4 + 32 + 24 + info["wide_bytes"] / N_HALF
    extra["algorithmic_bytes_per_ray"] = round(alg_per_ray, 2)
    extra["layout_bytes_per_ray"] = round(layout_per_ray, 2)
    extra["layout_gbs"] = round(N_HALF * layout_per_ray / (trav_ms * 1e-3) / 1e9, 2)
    out = {
        "metric": "incoherent MRays/s at 1 GPU (closest hit, 1M-triangle soup)",
        "value": round(value, 2),
        "unit": "MRays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded soup and rays, generated in-process; seeds: soup 1, rays 2, "
                "Cornell scene 3, single frame 11, jobs 0..63)",
        "config": {
            "workload": "ray-cast microbench: 1M-triangle soup, 2^23 incoherent closest-hit rays per step",
            "triangles": N_TRIS,
            "rays_per_step": N_HALF,
            "bvh_nodes": info["nodes"],
            "bvh_depth": info["max_depth"],
            "wide_nodes": info["wide_nodes"],
            "visiting_order": "near-first, 4-wide, rays Morton-ordered in the timed region",
            "tree": "rendu (host build, updateHierarchy)" if args.tree == "rendu" else "gpu linear BVH",
            "refill_min": tune["refill_min"],
            "chunk": tune.get("chunk", 0),
            "blocks_per_cu": tune.get("blocks_per_cu", 0),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": load_traffic(),
        },
        "extra": dict(extra, step_ms_events=round(k_ms, 4), traversal_kernel_ms=round(trav_ms, 4),
                      traversal_kernel_launches_timed=trav_n, hit_rate_incoherent=round(hit_rate, 4),
                      autotune_incoherent=tune, autotune_coherent=tune_c, bvh_build_s=round(build_s, 3),
                      input_generation_s=round(gen_s, 3)),
    }

    if not args.no_cpu:
        from oracle import oracle as orc
        ncpu = os.cpu_count() or 1
        threads = max(1, min(ncpu - 1, 15))
        host = host_cpu()
        ora = orc.OracleRaycaster()
        ora.add_mesh(mesh.positions, mesh.indices)
        ora.update()
        sample = incoherent[: args.cpu_sample]
        c0 = time.perf_counter()
        for _ in range(args.cpu_passes):
            exp = ora.intersect(sample, "closest", threads=threads)
        cpu_s = time.perf_counter() - c0
        # Hit-ID parity of the timed GPU results against the oracle on the same rays.
        got = hits[: sample.shape[0]]
        same = (got["hit"] == exp["hit"]) & (got["mesh_id"] == exp["mesh_id"]) & (got["local_id"] == exp["local_id"])
        exact = same & (got["t"] == exp["t"]) & (got["u"] == exp["u"]) & (got["v"] == exp["v"])
        t_ok = (got["hit"] == exp["hit"]) & ((exp["hit"] == 0) | (np.abs(got["t"] - exp["t"]) <= 1e-5 * np.abs(exp["t"])))
        out["extra"]["parity_vs_oracle"] = {"rays": int(sample.shape[0]), "same_hit_id": round(float(same.mean()), 7),
                                            "bit_exact_hit": round(float(exact.mean()), 7),
                                            "t_within_1e-5": round(float(t_ok.mean()), 7)}
        # The path tracer on the host cores (oracle), on bounded samples scaled explicitly: the
        # Cornell workload at full size, and a 240 x 135 crop-equivalent of the single frame
        # (same scene and camera, 8 spp) scaled to 1920 x 1080 x 64 spp.
        from rendu_amd import scene as S
        from rendu_amd import scenes
        scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=4)
        opt = orc.OraclePathTracer(scn)
        im = np.zeros((256, 256, 3), np.float32)
        c0 = time.perf_counter()
        opt.render(S.camera_desc(cam_of(256, 256)), 16, 5, im, S.RenderOpts(1, 0, 3, 0), threads=threads)
        cs = time.perf_counter() - c0
        scn3, cam3 = scenes.terrain_spheres()
        opt3 = orc.OraclePathTracer(scn3)
        im3 = np.zeros((135, 240, 3), np.float32)
        c0 = time.perf_counter()
        opt3.render(S.camera_desc(cam3(240, 135)), 8, 8, im3, S.RenderOpts(2, S.ESTIMATOR_EXTENDED, 3, 0), threads=threads)
        fs = time.perf_counter() - c0
        rate3 = 240 * 135 * 8 / fs
        out["extra"]["cpu_path_tracer"] = {
            "threads": threads,
            "cornell_256_16spp_depth5_s": round(cs, 3),
            "cornell_msamples_per_s": round(256 * 256 * 16 / cs / 1e6, 4),
            "frame_sample": "240x135, 8 spp, depth 8, RR 3 (same scene and camera)",
            "frame_sample_s": round(fs, 3),
            "frame_msamples_per_s": round(rate3 / 1e6, 4),
            "frame_1080p_64spp_s_scaled": round(1920 * 1080 * 64 / rate3, 1),
        }
        out["extra"]["cpu_jobs64"] = cpu_jobs(orc, threads)
        out["cpu_baseline"] = {
            "value": round(args.cpu_passes * sample.shape[0] / cpu_s / 1e6, 4),
            "unit": "MRays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_passes} passes over the first {sample.shape[0]} of the same incoherent rays, "
                      f"closest hit, {cpu_s:.1f} s, {threads} threads on a host with {host['nproc']} logical CPUs "
                      f"({host['model']})",
            "host_nproc": host["nproc"],
            "host_cpu_model": host["model"],
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
