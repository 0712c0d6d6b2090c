"""Batched render jobs in one renderer call (rdu_pathtracer_render_batch): every job's image equals,
bit for bit, the image of the same scene rendered alone with the same seed and options, whatever
the pass split (jobs cut across passes, several jobs per pass). The batch mixes terrain + sphere jobs
(reference estimator), a Cornell box with a point light (shadow rays), the small single-frame scene
with the extended estimator (environment samples, Russian roulette), a scene with alpha cut-outs,
spot and directional lights and a cube map (the cut-out walk, which then serves every job's shadow
rays), and a job on a GPU-built tree; depths differ between jobs."""
import numpy as np
import pytest

from conftest import log_case
from rendu_amd import jobs as J
from rendu_amd import scene as S
from rendu_amd import scenes
from rendu_amd.pathtracer import PathTracer
from test_gpu_pathtracer import feature_scene

pytestmark = pytest.mark.gpu


def _jobs():
    out = []
    for s in range(4):
        scn, cam, w, h, spp, depth = scenes.random_job(400 + s)
        side = 24 + 8 * s
        out.append(J.Job(scn, cam(side, side), side, side, min(spp, 8), depth, seed=s))
    scn, cam_of = scenes.cornell_spheres(seed=4, subdivisions=2, point_light=True)
    out.append(J.Job(scn, cam_of(40, 30), 40, 30, 8, 5, seed=11))
    scn, cam_of = scenes.terrain_spheres(seed=5, grid=64, spheres=60, subdivisions=1, sky_size=(128, 64))
    out.append(J.Job(scn, cam_of(48, 27), 48, 27, 4, 8, seed=12, estimator=S.ESTIMATOR_EXTENDED, rr_start=3))
    scn, cam_of = feature_scene(S.BG_SKYBOX)
    out.append(J.Job(scn, cam_of(32, 24), 32, 24, 4, 4, seed=13))
    scn, cam, w, h, spp, depth = scenes.random_job(450)
    out.append(J.Job(scn, cam(36, 36), 36, 36, 4, depth, seed=14, gpu_build=True))
    return out


def _alone(job, pt):
    img = np.zeros((job.height, job.width, 3), np.float32)
    pt.render(job.camera, job.samples, job.depth, img, seed=job.seed, estimator=job.estimator, rr_start=job.rr_start)
    return img


@pytest.mark.parametrize("max_paths", [0, 3000], ids=["default-passes", "small-passes"])
def test_batch_images_equal_alone(max_paths):
    js = _jobs()
    log_case(f"batch of {len(js)} jobs in one call, max_paths {max_paths}, vs alone")
    pts = [PathTracer(j.scene, gpu_build=j.gpu_build) for j in js]
    alone = [_alone(j, pt) for j, pt in zip(js, pts)]
    images, secs, st = J.render_batch(list(zip(pts, js)), max_paths=max_paths)
    for k, (a, b) in enumerate(zip(alone, images)):
        assert a.mean() > 0.0, k
        assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))
    assert st["passes"] >= (1 if max_paths == 0 else 4)
    assert st["extension_rays"] >= st["paths"] > 0 and st["shadow_rays"] > 0
    assert np.all(secs > 0) and np.all(secs <= st["seconds"] + 1e-6)


@pytest.mark.parametrize("groups", [1, 2])
def test_merged_runner_matches_alone(groups):
    js = _jobs()[:5]
    log_case(f"jobs.run_merged (threaded builds, {groups} batch call(s)) vs alone")
    res, wall = J.run_merged(js, build_workers=4, groups=groups)
    for j, r in zip(js, res):
        np.testing.assert_array_equal(r.image, J.run_job(j).image)
    summ = J.summary(res, wall)
    assert summ["jobs"] == 5 and summ["jobs_per_s"] > 0
    assert all(0 < r.render_s <= r.latency_s for r in res)
