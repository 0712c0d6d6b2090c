"""Rendu's own scenes (resources/pbrdemo, copied as data under tests/golden) rendered by the GPU path
tracer against the oracle, sample for sample (bars of test_gpu_pathtracer.py):

- furnace-test.scene: 22 white spheres, dielectric and metal, roughness 0 ... 1, in a constant 0.5
  cubemap. Beyond parity, each sphere's mean over the central part of its disk equals 0.5 times the
  material's directional albedo averaged over the view angles there (a convex sphere in a uniform
  environment reflects L E(mu) in expectation; E from quadrature of the GGX evaluation, whose
  sampling estimator and textbook forms tests/test_oracle_pins.py pins), within 3.5%: the
  neighbouring spheres (radius 1, 3 apart) hide part of the environment from wide lobes (rough
  metal 1.0 lands 2-3% low at 1024 spp). Roughness below 0.2 is left out of this check: there
  alpha^2 < 1e-4 and D's denominator floor (MaterialGGX.cpp:11) makes the evaluated pdf differ
  from the sampled GGX density, so Rendu's estimator does not converge to the integral of its own
  BRDF (a smooth metal renders about 0.5 while the floored BRDF integrates to a few percent).
- cornellbox-emissive.scene: emissive walls and an emissive box textured with an EXR grid.
- cornellbox-cutout.scene: alpha-masked leaves (the shadow walk through cut-outs) under two spot
  lights (the cos(cos(outer)) cone), a vase with a JPEG normal map. Its light paths are chaotic:
  perturbing one cosine of the oracle's direction sampling by a single ulp changes 4% of its own
  pixel channels at depth 2 and 10% at depth 5 (the device's cos/sin differ from glibc's by an ulp
  on some inputs), so it is compared sample for sample at depth 1 (every channel: emission, cut-outs,
  normal maps, both spot lights and their shadow walks), and at depth 5 by independent-stream
  renders, per-pixel z-scores of the difference of the means (as test_gpu_statistics.py).
- cornellbox-parallax.scene (parallax materials shaded as Rendu's path tracer shades them, JPEG
  textures, a spot light) and cornellbox-textured.scene (JPEG maps on the floor and boxes, a
  skybox of EXR faces): the same two comparisons.
"""
import math
import os

import numpy as np
import pytest

from conftest import log_case
from oracle import oracle as orc
from rendu_amd import ingest
from rendu_amd import scene as S
from rendu_amd.pathtracer import PathTracer
from test_gpu_pathtracer import assert_close, render_both

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOTS = [os.path.join(GOLD, "cornellbox"), os.path.join(GOLD, "common"), os.path.join(GOLD, "pbrdemo")]


def load(name):
    return ingest.load_scene(os.path.join(GOLD, "pbrdemo", name), ROOTS)


def directional_albedo(mu, rough, metal, n=96):
    """E(mu) = integral of the GGX evaluation over the hemisphere (white base), midpoint rule."""
    wo = np.array([math.sqrt(max(0.0, 1 - mu * mu)), 0.0, mu])
    zg = (np.arange(n) + 0.5) / n
    pg = (np.arange(n) + 0.5) / n * 2 * math.pi
    acc = 0.0
    for z in zg:
        r = math.sqrt(1 - z * z)
        for p in pg:
            acc += orc.ggx_eval(wo, [1.0, 1.0, 1.0], rough, metal, [r * math.cos(p), r * math.sin(p), z])[0]
    return acc / n * (2 * math.pi / n)


def test_furnace_scene_gpu_vs_oracle_and_energy():
    scn, cam_of = load("furnace-test.scene")
    w, h, spp = 192, 144, 64
    cam = cam_of(w, h)
    log_case("Rendu furnace-test.scene 192x144 64spp depth 5")
    got, exp, _ = render_both(scn, cam, w, h, spp, 5, seed=1)
    assert_close(got, exp)
    assert np.all(got[0, 0] == 0.5)  # the background seen directly
    corner, dx, dy = (np.asarray(v, np.float64) for v in cam.pixelShifts())
    eye = np.asarray(cam.position, np.float64)
    yy, xx = np.mgrid[0:h, 0:w] + 0.5
    # Disk radius in pixels of a unit sphere at distance d: the half-width of the image plane at
    # distance d is d * tan(fov / 2) * ratio over w / 2 pixels.
    half_h = math.tan(0.5 * cam.fov)
    gl = np.polynomial.legendre.leggauss(6)
    for o in scn.objects:
        mat = scn.materials[o.material]
        rough = float(scn.textures[mat.textures[2]][0, 0, 0])
        metal = float(scn.textures[mat.textures[2]][0, 0, 1])
        c = np.asarray(o.model, np.float64)[:3, 3]
        u, v, _ = np.linalg.solve(np.stack([dx, dy, -(c - eye)], 1), eye - corner)
        d = np.linalg.norm(c - eye)
        r_px = (1.0 / d) / half_h * (h / 2)
        k = 0.7  # central part of the disk: view cosines from sqrt(1 - k^2) to 1
        mask = (xx - u * w) ** 2 + (yy - v * h) ** 2 < (k * r_px) ** 2
        mean = float(got[mask].mean())
        if rough < 0.2:
            assert 0.45 < mean < 0.55, (rough, metal, mean)
            continue
        mu0 = math.sqrt(1 - k * k)
        mus = 0.5 * (gl[0] + 1) * (1 - mu0) + mu0
        ws = 0.5 * gl[1] * (1 - mu0)
        e_avg = sum(wt * directional_albedo(m, rough, metal) * 2 * m for m, wt in zip(mus, ws)) / (1 - mu0 * mu0)
        assert mean == pytest.approx(0.5 * e_avg, rel=0.035), (rough, metal, mean, 0.5 * e_avg)


def test_emissive_cornellbox_gpu_vs_oracle():
    scn, cam_of = load("cornellbox-emissive.scene")
    log_case("Rendu cornellbox-emissive.scene 64x64 16spp depth 5")
    got, exp, _ = render_both(scn, cam_of(64, 64), 64, 64, 16, 5, seed=2)
    assert exp.mean() > 0.05
    assert_close(got, exp)


def test_cutout_cornellbox_gpu_vs_oracle():
    scn, cam_of = load("cornellbox-cutout.scene")
    log_case("Rendu cornellbox-cutout.scene 64x64 16spp depth 1")
    got, exp, st = render_both(scn, cam_of(64, 64), 64, 64, 16, 1, seed=3)
    assert exp.mean() > 0.01 and st["shadow_rays"] > 0
    assert_close(got, exp, frac=0.9999)
    independent_streams_agree(scn, cam_of, "cornellbox-cutout.scene")


def independent_streams_agree(scn, cam_of, name, w=32, h=32, k=8, spp=8, depth=5):
    """k GPU renders and k oracle renders with independent seeds: per-pixel z-scores of the
    difference of the two means are standard normal when both estimate the same image."""
    log_case(f"Rendu {name} depth {depth}: {k} GPU vs {k} oracle renders, independent seeds")
    cam = cam_of(w, h)
    pt = PathTracer(scn)
    ora = orc.OraclePathTracer(scn)
    gpu = np.zeros((k, h, w, 3))
    cpu = np.zeros((k, h, w, 3))
    for i in range(k):
        img = np.zeros((h, w, 3), np.float32)
        pt.render(cam, spp, depth, img, seed=1 + i)
        gpu[i] = img
        img = np.zeros((h, w, 3), np.float32)
        ora.render(S.camera_desc(cam), spp, depth, img, S.RenderOpts(1001 + i, 0, 3, 0))
        cpu[i] = img
    se = np.sqrt(gpu.var(0, ddof=1) / k + cpu.var(0, ddof=1) / k)
    live = se > 0
    z = (gpu.mean(0)[live] - cpu.mean(0)[live]) / se[live]
    assert z.size >= 500  # channels that vary between renders (a directly seen background does not)
    assert abs(z.mean()) < 4.0 / np.sqrt(z.size), z.mean()
    assert np.mean(np.abs(z) > 3.5) < 0.01
    assert 0.7 < z.std() < 1.3


def test_parallax_cornellbox_gpu_vs_oracle():
    """Parallax materials (Rendu's path tracer shades every non-emissive material with the GGX
    model of its first three textures, PathTracer.cpp:224-236, so the fourth, the depth map, is
    loaded and unused), JPEG textures, a spot light and normal maps."""
    scn, cam_of = load("cornellbox-parallax.scene")
    assert [m.type for m in scn.materials].count(ingest.MATERIAL_TYPES["Parallax"]) == 2
    log_case("Rendu cornellbox-parallax.scene 64x64 16spp depth 1")
    got, exp, st = render_both(scn, cam_of(64, 64), 64, 64, 16, 1, seed=4)
    assert exp.mean() > 0.005 and st["shadow_rays"] > 0
    assert_close(got, exp, frac=0.9999)
    independent_streams_agree(scn, cam_of, "cornellbox-parallax.scene")


def test_textured_cornellbox_gpu_vs_oracle():
    """JPEG colour, normal and roughness/metalness maps on the floor and both boxes, an emissive
    light and a skybox of uncompressed half-float EXR faces."""
    scn, cam_of = load("cornellbox-textured.scene")
    assert scn.background.mode == S.BG_SKYBOX
    log_case("Rendu cornellbox-textured.scene 64x64 16spp depth 1")
    got, exp, _ = render_both(scn, cam_of(64, 64), 64, 64, 16, 1, seed=5)
    assert exp.mean() > 0.05
    assert_close(got, exp, frac=0.9999)
    independent_streams_agree(scn, cam_of, "cornellbox-textured.scene")
