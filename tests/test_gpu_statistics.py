"""Statistical checks of the GPU path tracer, independent of sample-for-sample agreement.

The sample-for-sample tests (test_gpu_pathtracer.py) reuse the oracle's random stream; these draw
independent streams on each side, so they would also catch an estimator that matches the oracle
only because both share a mistake in how samples are drawn:
- bias: K independent GPU renders against K independent oracle renders of the same scene, per-pixel
  z-scores of the difference of the means;
- convergence: the RMSE against a 4096-spp render falls as spp^-1/2 (Monte Carlo rate).
"""
import numpy as np
import pytest

from conftest import log_case
from oracle import oracle as orc
from rendu_amd import scene as S
from rendu_amd import scenes
from rendu_amd.pathtracer import PathTracer

pytestmark = pytest.mark.gpu
F32 = np.float32


def test_bias_z_scores_against_the_oracle():
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=2)
    w = h = 24
    k, spp, depth = 12, 4, 5
    log_case(f"bias: {k} GPU vs {k} oracle renders, {w}x{h}, {spp} spp, independent seeds")
    cam = cam_of(w, h)
    pt = PathTracer(scn)
    ora = orc.OraclePathTracer(scn)
    gpu = np.zeros((k, h, w, 3), np.float64)
    cpu = np.zeros((k, h, w, 3), np.float64)
    for i in range(k):
        img = np.zeros((h, w, 3), F32)
        pt.render(cam, spp, depth, img, seed=1 + i)
        gpu[i] = img
        img = np.zeros((h, w, 3), F32)
        ora.render(S.camera_desc(cam), spp, depth, img, S.RenderOpts(1001 + i, 0, 3, 0))
        cpu[i] = img
    mg, mc = gpu.mean(0), cpu.mean(0)
    vg, vc = gpu.var(0, ddof=1), cpu.var(0, ddof=1)
    se = np.sqrt(vg / k + vc / k)
    live = se > 0
    # Where neither side varies (pixels that see only the black background), the means agree exactly.
    np.testing.assert_array_equal(mg[~live], mc[~live])
    z = (mg[live] - mc[live]) / se[live]
    n = z.size
    assert n > 0.4 * mg.size
    # Unbiased: z is about N(0, 1) (a t law with 2k - 2 degrees of freedom, heavier tails).
    assert abs(z.mean()) < 4.0 / np.sqrt(n), f"mean z {z.mean():.4f} over {n} channels"
    assert np.mean(np.abs(z) > 3.5) < 0.01, f"{np.mean(np.abs(z) > 3.5):.4f} of channels beyond 3.5 sigma"
    assert 0.7 < z.std() < 1.3


def test_convergence_rate_against_4096_spp():
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=3)
    w = h = 64
    log_case(f"convergence: {w}x{h}, 16/64/256 spp vs 4096 spp")
    cam = cam_of(w, h)
    pt = PathTracer(scn)
    ref = np.zeros((h, w, 3), F32)
    pt.render(cam, 4096, 5, ref, seed=4242)
    spps = [16, 64, 256]
    rmse = []
    for s in spps:
        img = np.zeros((h, w, 3), F32)
        pt.render(cam, s, 5, img, seed=7)
        rmse.append(float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2))))
    slope = np.polyfit(np.log(spps), np.log(rmse), 1)[0]
    # The 4096-spp reference carries noise of its own (1/16 of the variance at 256 spp), which
    # flattens the fitted slope slightly.
    assert -0.6 < slope < -0.4, f"RMSE {rmse} gives slope {slope:.3f}"
    assert rmse[0] > rmse[1] > rmse[2]
