"""The extended estimator's light sampling on the GPU against the oracle, sample for sample (bars of
test_gpu_pathtracer.py): lat-long environment sampling, emissive-triangle sampling (with a textured
emitter), both together with an analytic light, and with alpha cut-outs in the visibility walk.
The estimator itself (its expectation equals Rendu's) is pinned in tests/test_oracle_extension.py."""
import os

import numpy as np
import pytest

from conftest import log_case
from rendu_amd import ingest
from rendu_amd import scene as S
from rendu_amd import scenes
from test_gpu_pathtracer import assert_close, feature_scene, render_both

pytestmark = pytest.mark.gpu
EXT = S.ESTIMATOR_EXTENDED
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("case", ["env", "env+tri+point", "tri", "cutouts-env", "emissive-grid"])
def test_extended_light_sampling_matches_oracle(case):
    if case == "env":
        scn, cam_of = scenes.sky_spheres(sun=True)
    elif case == "env+tri+point":
        scn, cam_of = scenes.sky_spheres(sun=True, point_light=True, emissive=2.0)
    elif case == "tri":
        scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=2)
    elif case == "cutouts-env":
        scn, cam_of = feature_scene(S.BG_LATLONG)
    else:
        scn, cam_of = ingest.load_scene(os.path.join(GOLD, "pbrdemo", "cornellbox-emissive.scene"),
                                        [os.path.join(GOLD, d) for d in ("cornellbox", "common", "pbrdemo")])
    w, h = 48, 36
    log_case(f"extended estimator, {case}, {w}x{h} 16 spp depth 5, RR from 3")
    got, exp, st = render_both(scn, cam_of(w, h), w, h, 16, 5, seed=9, estimator=EXT, rr_start=3)
    assert exp.mean() > 1e-3 and st["shadow_rays"] > 0
    assert_close(got, exp)
