"""GPU path tracer (pt.hip) against the oracle's restatement of PathTracer::render, sample for sample.

Both draw the same Philox stream per (seed, pixel, sample, bounce), trace the same tree and follow
the same estimator. They differ only in the last ulps of transcendentals (device pow/cos/sin/exp vs
glibc), which very rarely flip a discrete choice and send one path elsewhere. Bar: at least 98% of
pixel channels within 1e-4 relative (+1e-6 absolute), mean absolute difference below 1e-3 of the
mean, and no systematic bias (mean signed difference within 3 standard errors).
"""
import numpy as np
import pytest

from conftest import log_case
from oracle import oracle as orc
from rendu_amd import scene as S
from rendu_amd import scenes
from rendu_amd.pathtracer import PathTracer
from rendu_amd.workloads import icosphere, quad

pytestmark = pytest.mark.gpu
F32 = np.float32


def render_both(scn, cam, w, h, spp, depth, seed=5, estimator=0, rr_start=3, max_paths=0):
    desc = S.camera_desc(cam)
    exp = np.zeros((h, w, 3), F32)
    orc.OraclePathTracer(scn).render(desc, spp, depth, exp, S.RenderOpts(seed, estimator, rr_start, 0))
    got = np.zeros((h, w, 3), F32)
    stats = PathTracer(scn).render(cam, spp, depth, got, seed=seed, estimator=estimator, rr_start=rr_start,
                                   max_paths=max_paths)
    return got, exp, stats


def assert_close(got, exp, frac=0.98):
    d = np.abs(got.astype(np.float64) - exp)
    ok = d <= 1e-4 * np.abs(exp) + 1e-6
    assert ok.mean() >= frac, f"only {ok.mean():.4f} of channels match"
    assert d.mean() <= 1e-3 * max(np.abs(exp).mean(), 1e-6) + 1e-7
    sd = (got.astype(np.float64) - exp).ravel()
    assert abs(sd.mean()) <= 3 * sd.std() / np.sqrt(sd.size) + 1e-7


def test_cornell_reference_estimator():
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=3)
    log_case("cornell 48x48 16spp depth5 reference")
    got, exp, st = render_both(scn, cam_of(48, 48), 48, 48, 16, 5)
    assert exp.mean() > 0.01
    assert_close(got, exp)
    assert st["samples"] == 16 and st["paths"] == 48 * 48 * 16
    assert st["extension_rays"] >= st["paths"]


def test_cornell_point_light_nee_and_passes():
    scn, cam_of = scenes.cornell_spheres(seed=4, subdivisions=2, point_light=True)
    log_case("cornell + point light, NEE, 3 passes")
    got, exp, st = render_both(scn, cam_of(40, 32), 40, 32, 8, 4, max_paths=40 * 32 * 3)
    assert_close(got, exp)
    assert st["shadow_rays"] > 0


def test_pass_split_does_not_change_the_image():
    """Samples accumulate per pixel in sample order whatever the pass split, so one pass of all
    samples (the default pass size), one sample per pass and uneven passes give the same bits
    (extended estimator, point light and Russian roulette: every kernel of a pass)."""
    from rendu_amd.pathtracer import PathTracer
    scn, cam_of = scenes.cornell_spheres(seed=4, subdivisions=2, point_light=True)
    log_case("pass split: 1, 16 and 6 passes of a 64x48 16-spp render, 1-3 passes in flight")
    pt = PathTracer(scn)
    w, h = 64, 48
    imgs, passes, rays = [], [], []
    # (max_paths, passes in flight): one pass; 16 one-sample passes one at a time, 2 and 3 at once;
    # 6 uneven passes 3 at once (each pass's accumulation waits for the previous one's).
    for mp, lanes in ((0, 0), (w * h, 1), (w * h, 2), (w * h, 3), (w * h * 3, 3), (w * h * 3, 0)):
        img = np.zeros((h, w, 3), np.float32)
        st = pt.render(cam_of(w, h), 16, 5, img, seed=9, estimator=S.ESTIMATOR_EXTENDED, rr_start=2, max_paths=mp,
                       passes_in_flight=lanes)
        imgs.append(img)
        passes.append(st["passes"])
        rays.append((st["extension_rays"], st["shadow_rays"]))
    assert passes == [1, 16, 16, 16, 6, 6]
    assert imgs[0].mean() > 0.01
    for img in imgs[1:]:
        assert np.array_equal(imgs[0], img)
    assert len(set(rays)) == 1 and rays[0][1] > 0


def feature_scene(bg_mode):
    """Masked cut-outs with UVs, normal maps, a transformed object, spot + directional + point lights."""
    rng = np.random.default_rng(7)
    scn = S.Scene()
    checker = np.zeros((16, 16, 4), F32)
    yy, xx = np.mgrid[0:16, 0:16]
    checker[..., 0] = 0.8
    checker[..., 1] = 0.3 + 0.4 * ((xx // 4) % 2)
    checker[..., 2] = 0.2
    checker[..., 3] = ((xx // 4 + yy // 4) % 2).astype(F32)  # alpha cut-out pattern
    t_alpha = scn.add_texture(checker)
    nm = np.zeros((16, 16, 4), F32)
    nm[..., 0] = 0.5 + 0.3 * np.sin(xx / 2.0)
    nm[..., 1] = 0.5 + 0.3 * np.cos(yy / 3.0)
    nm[..., 2] = 0.9
    nm[..., 3] = 1
    t_nm = scn.add_texture(nm)
    t_rm = scn.add_texture(S.constant_texture([0.4, 0.2, 1.0, 1.0]))
    t_rm2 = scn.add_texture(S.constant_texture([0.15, 1.0, 1.0, 1.0]))
    t_white = scn.add_texture(S.constant_texture([0.8, 0.8, 0.8, 1.0]))
    t_flat = scn.add_texture(S.constant_texture([0.5, 0.5, 1.0, 1.0]))
    m_mask = scn.add_material(S.Material(S.MATERIAL_REGULAR, (t_alpha, t_nm, t_rm), masked=True))
    m_metal = scn.add_material(S.Material(S.MATERIAL_REGULAR, (t_white, t_nm, t_rm2)))
    m_floor = scn.add_material(S.Material(S.MATERIAL_REGULAR, (t_white, t_flat, t_rm)))
    floor = scn.add_mesh(quad((-2, 0, 2), (2, 0, 2), (2, 0, -2), (-2, 0, -2)))
    scn.objects.append(S.Object(floor, m_floor))
    cut = scn.add_mesh(quad((-1, 0.2, 0.3), (1, 0.2, 0.3), (1, 1.6, -0.2), (-1, 1.6, -0.2)))
    scn.objects.append(S.Object(cut, m_mask, use_texcoords=True))
    sph = scn.add_mesh(icosphere(3, radius=0.4, center=(0, 0, 0)))
    model = np.eye(4, dtype=F32)
    model[:3, :3] = np.array([[1.3, 0.2, 0.0], [0.0, 0.9, 0.1], [0.1, 0.0, 1.1]], F32)
    model[:3, 3] = [0.3, 0.45, -0.6]
    scn.objects.append(S.Object(sph, m_metal, model=model, use_texcoords=True))
    scn.lights.append(S.Light(S.LIGHT_POINT, (6.0, 5.0, 4.0), (0.2, 2.2, 0.8), radius=5.0))
    scn.lights.append(S.Light(S.LIGHT_SPOT, (8.0, 8.0, 9.0), (-1.0, 2.0, 1.0), direction=(0.4, -0.8, -0.45),
                              radius=6.0, angles=(0.3, 0.6)))
    scn.lights.append(S.Light(S.LIGHT_DIRECTIONAL, (0.6, 0.6, 0.7), direction=(-0.3, -0.9, -0.2)))
    if bg_mode == S.BG_SKYBOX:
        faces = [scn.add_texture(np.full((8, 8, 4), 0.2 + 0.1 * k, F32)) for k in range(6)]
        scn.background = S.Background(S.BG_SKYBOX, cube=tuple(faces))
    elif bg_mode == S.BG_IMAGE:
        img = rng.random((12, 20, 4)).astype(F32)
        scn.background = S.Background(S.BG_IMAGE, texture=scn.add_texture(img))
    elif bg_mode == S.BG_LATLONG:
        env = np.zeros((32, 64, 4), F32)
        env[..., 0] = np.linspace(0.2, 1.5, 64)[None, :]
        env[..., 1] = np.linspace(1.0, 0.1, 32)[:, None]
        env[..., 2] = 0.5
        scn.background = S.Background(S.BG_LATLONG, texture=scn.add_texture(env))
    elif bg_mode == S.BG_ATMOSPHERE:
        table = (0.5 + 0.5 * rng.random((512, 512, 4))).astype(F32)
        scn.background = S.Background(S.BG_ATMOSPHERE, texture=scn.add_texture(table),
                                      sun_direction=(0.3, 0.8, 0.52))
    else:
        scn.background = S.Background(S.BG_COLOR, (0.1, 0.2, 0.3))

    def cam_of(w, h):
        from rendu_amd.camera import Camera
        c = Camera()
        c.pose([0.3, 1.2, 3.0], [0.0, 0.6, 0.0], [0.0, 1.0, 0.0])
        c.projection(ratio=w / h, fov=1.0)
        return c

    return scn, cam_of


@pytest.mark.parametrize("bg", [S.BG_COLOR, S.BG_SKYBOX, S.BG_IMAGE, S.BG_LATLONG, S.BG_ATMOSPHERE])
def test_feature_scene_backgrounds(bg):
    scn, cam_of = feature_scene(bg)
    log_case(f"feature scene (masked walk, normal maps, 3 lights) background {bg}")
    got, exp, st = render_both(scn, cam_of(40, 30), 40, 30, 8, 5, seed=11)
    assert exp.mean() > 0.01
    assert_close(got, exp, frac=0.97)
    assert st["shadow_rays"] > 0


def test_extended_estimator_russian_roulette():
    scn, cam_of = scenes.cornell_spheres(seed=5, subdivisions=2, point_light=True)
    log_case("cornell extended estimator, RR from bounce 2")
    got, exp, _ = render_both(scn, cam_of(32, 32), 32, 32, 16, 8, seed=3, estimator=S.ESTIMATOR_EXTENDED, rr_start=2)
    assert_close(got, exp)


def test_white_furnace_gpu():
    scn, cam_of = scenes.furnace(subdivisions=3)
    log_case("white furnace")
    img = np.zeros((32, 32, 3), F32)
    PathTracer(scn).render(cam_of(32, 32), 64, 4, img, seed=2)
    assert np.allclose(img[0, 0], 0.5, atol=1e-5)
    assert abs(img[12:20, 12:20].mean() - 0.5) < 0.03


def test_cornell_full_size_config():
    """The small path-tracing workload at its full size: 256x256, 16 spp, depth 5."""
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=4)
    log_case("cornell 256x256 16spp depth5 (full size)")
    got, exp, st = render_both(scn, cam_of(256, 256), 256, 256, 16, 5, seed=1)
    assert_close(got, exp)
    assert st["paths"] == 256 * 256 * 16


def test_raycaster_of_the_scene():
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=2)
    pt = PathTracer(scn)
    rc = pt.raycaster()
    h = rc.intersects([0.0, 1.0, 2.5], [0.0, 0.0, -1.0])
    assert h.hit and h.meshId < len(scn.objects)
    assert rc.info()["meshes"] == len(scn.objects)


def test_terrain_spheres_small():
    """The single-frame workload's construction at a reduced size: fBm terrain, instanced
    icospheres with random GGX materials, lat-long sky, RR from bounce 3."""
    scn, cam_of = scenes.terrain_spheres(seed=5, grid=64, spheres=60, subdivisions=1, sky_size=(128, 64))
    log_case("terrain+spheres small 64x36 8spp depth8 extended")
    got, exp, _ = render_both(scn, cam_of(64, 36), 64, 36, 8, 8, seed=9, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    assert exp.mean() > 0.05
    assert_close(got, exp)


def test_terrain_spheres_full_geometry():
    """Full single-frame scene (3,082,242 triangles over 2001 meshes, 1024x512 sky) at a reduced
    resolution and sample count, sample for sample against the oracle."""
    scn, cam_of = scenes.terrain_spheres()
    assert sum(scn.meshes[o.mesh].triangle_count for o in scn.objects) == 3_082_242
    log_case("terrain+spheres full geometry 96x54 4spp depth8 extended")
    got, exp, st = render_both(scn, cam_of(96, 54), 96, 54, 4, 8, seed=2, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    assert_close(got, exp)
    assert st["paths"] == 96 * 54 * 4


def test_single_frame_at_full_resolution():
    """The single frame at its configured size, 1920x1080 over the full scene: 2 spp sample for
    sample against the oracle (about 4M paths on the host cores), then the configured 64 spp on
    the GPU, where size-independent properties hold: the same bits from the default (8 passes of
    2^24 paths, 3 in flight), from 4 passes of 2^25 one at a time and from 32 passes of 2^22, every
    pixel finite and non-negative, and a mean within 5 standard errors of the mean of the 2-spp
    oracle render (the two are independent estimates of one image)."""
    scn, cam_of = scenes.terrain_spheres()
    w, h = 1920, 1080
    cam = cam_of(w, h)
    log_case("single frame 1920x1080 2spp depth8 extended vs oracle")
    got, exp, st = render_both(scn, cam, w, h, 2, 8, seed=4, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    assert_close(got, exp)
    assert st["paths"] == w * h * 2
    log_case("single frame 1920x1080 64spp depth8 extended: 8, 4 and 32 passes")
    pt = PathTracer(scn)
    imgs = []
    for mp, lanes, passes in ((0, 0, 8), (1 << 25, 1, 4), (1 << 22, 0, 32)):
        img = np.zeros((h, w, 3), F32)
        st = pt.render(cam, 64, 8, img, seed=11, estimator=S.ESTIMATOR_EXTENDED, rr_start=3, max_paths=mp,
                       passes_in_flight=lanes)
        imgs.append(img)
        assert st["passes"] == passes
    assert np.array_equal(imgs[0], imgs[1]) and np.array_equal(imgs[0], imgs[2])
    full = imgs[0].astype(np.float64)
    assert np.all(np.isfinite(full)) and full.min() >= 0.0
    # Standard error of the 2-spp image mean, from its per-pixel spread (pixels are independent).
    se = exp.astype(np.float64).std() / np.sqrt(exp.size)
    assert abs(full.mean() - exp.mean()) < 5 * se + 1e-4, (full.mean(), exp.mean(), se)
