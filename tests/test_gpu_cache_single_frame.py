"""The radiance cache on the single-frame workload: the full 2M-triangle scene (512 x 512 fBm
terrain, 2000 icospheres, lat-long sky; scenes.terrain_spheres(11)) at 480 x 270, depth 8, Russian
roulette from bounce 3, the extended estimator, cache trained online during the render.

The method without the network is the same tracer truncated at the termination bounce (depth 3,
i.e. paths end after bounce 2 with nothing added). Against a 2048-spp plain render, at 64 spp:
* the cached render (every bounce-2 vertex ending in the cache) removes most of the truncation's
  bias (measured -1.14% without the cache, +0.31% with it) and has less error than the truncated
  render, at the same rays per sample;
* with the trust test (the cache answers where its held-out error is small) the error stays within
  2% of the full plain render and the mean within 1%.
What this scene leaves to a cache terminating at bounce 2 is small: the tail past bounce 2 holds
about 1% of the radiance and 4% of the rays (2.206 of 2.298 rays per sample), so the cache cannot
cut this scene's time to a given error by much; the measured numbers are printed.
"""
import numpy as np
import pytest

from conftest import log_case
from rendu_amd import scene as S
from rendu_amd import scenes
from rendu_amd.pathtracer import PathTracer
from rendu_amd.rcache import RadianceCache

pytestmark = pytest.mark.gpu
EXT = S.ESTIMATOR_EXTENDED


def test_cache_on_the_single_frame_scene():
    log_case("radiance cache, single-frame scene (2M triangles) at 480x270, 64 spp vs 2048-spp plain")
    scn, cam_of = scenes.terrain_spheres(seed=11)
    lo, hi = S.scene_bounds(scn)
    pt = PathTracer(scn)
    w, h = 480, 270
    cam = cam_of(w, h)
    ref = np.zeros((h, w, 3), np.float32)
    pt.render(cam, 2048, 8, ref, seed=1000, estimator=EXT, rr_start=3)
    rmse = lambda a: float(np.sqrt(np.mean((a.astype(np.float64) - ref) ** 2)))
    bias = lambda a: float(a.mean() / ref.mean() - 1.0)
    rays = lambda st: (st["extension_rays"] + st["shadow_rays"]) / (w * h * 64)

    def plain(depth):
        img = np.zeros_like(ref)
        pt.render(cam, 64, depth, img, seed=7, estimator=EXT, rr_start=3)
        return img, pt.last_stats()

    full, st_full = plain(8)
    trunc, st_trunc = plain(3)
    cached = np.zeros_like(ref)
    st_c = pt.render_cached(RadianceCache(lo, hi, seed=5), cam, 64, 8, cached, seed=7, estimator=EXT, rr_start=3,
                            always_trust=True)
    cache = RadianceCache(lo, hi, seed=5)
    pt.render_cached(cache, cam, 64, 8, np.zeros_like(ref), seed=50, estimator=EXT, rr_start=3)
    trusted = np.zeros_like(ref)
    st_t = pt.render_cached(cache, cam, 64, 8, trusted, seed=7, estimator=EXT, rr_start=3)
    for name, img, st in (("plain", full, st_full), ("truncated at bounce 2", trunc, st_trunc),
                          ("cached, every bounce-2 vertex", cached, st_c), ("cached, trust test", trusted, st_t)):
        print(f"[single frame] {name}: rmse {rmse(img):.5f} bias {bias(img):+.4f} rays/sample {rays(st):.3f}",
              flush=True)
    assert st_c["train_pairs"] > 0 and st_c["cache_queries"] > 0 and st_t["cache_queries"] > 0
    assert abs(bias(full)) < 0.002
    assert bias(trunc) < -0.008  # the tail the cache stands in for
    assert abs(bias(cached)) < 0.5 * abs(bias(trunc)), (bias(cached), bias(trunc))
    assert rmse(cached) < rmse(trunc), (rmse(cached), rmse(trunc))
    assert rays(st_c) <= rays(st_trunc) * 1.01
    assert abs(bias(trusted)) < 0.01 and rmse(trusted) < 1.02 * rmse(full), (bias(trusted), rmse(trusted), rmse(full))
