"""Instanced scenes (rdu_scene_set_instancing, an extension): one tree per mesh in object space,
shared by its objects, under a top level over the objects' boxes; rays enter an object's tree
transformed by its inverse model. Hits differ from the flattened scene's (Rendu's world-space
triangles) only in the transform's rounding, so renders are compared with the flattened scene's
renders of the same seeds: the same paths almost everywhere (a path changes only where an ulp of
its hit point flips a later decision), and equal means."""
import math

import numpy as np
import pytest

from conftest import log_case
from rendu_amd import jobs as J
from rendu_amd import scene as S
from rendu_amd import scenes
from rendu_amd.pathtracer import PathTracer
from rendu_amd.workloads import icosphere, quad

pytestmark = pytest.mark.gpu
F32 = np.float32


def shared_mesh_scene(seed=5, n=40, emissive=False, light=True):
    """A ground quad and n objects of one icosphere mesh under random rotations, non-uniform scales
    and translations; a point light (any-hit shadow rays); optionally emissive instances and a
    lat-long sky (the extended estimator's area-light sampling)."""
    rng = np.random.default_rng(seed)
    scn = S.Scene()
    t_flat = scn.add_texture(S.constant_texture([0.5, 0.5, 1.0, 1.0]))
    t_rm = scn.add_texture(S.constant_texture([0.5, 0.0, 1.0, 1.0]))
    g = scn.add_material(S.Material(S.MATERIAL_REGULAR, (scn.add_texture(S.constant_texture([0.6, 0.6, 0.55, 1])), t_flat, t_rm)))
    ground = scn.add_mesh(quad((-6, 0, 6), (6, 0, 6), (6, 0, -6), (-6, 0, -6)))
    scn.objects.append(S.Object(ground, g))
    sph = scn.add_mesh(icosphere(2, radius=1.0))
    for k in range(n):
        a, b = rng.uniform(0, 2 * math.pi, 2)
        ca, sa, cb, sb = math.cos(a), math.sin(a), math.cos(b), math.sin(b)
        rot = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, cb, -sb], [0, sb, cb]])
        scale = np.diag(rng.uniform(0.2, 0.6, 3))
        model = np.eye(4, dtype=F32)
        model[:3, :3] = (rot @ scale).astype(F32)
        model[:3, 3] = (rng.uniform(-4, 4), rng.uniform(0.3, 1.5), rng.uniform(-4, 2))
        col = rng.uniform(0.2, 0.9, 3)
        if emissive and k % 8 == 0:
            mat = scn.add_material(S.Material(S.MATERIAL_EMISSIVE, (scn.add_texture(S.constant_texture([4.0, 3.0, 2.0, 1])), t_flat, t_rm)))
        else:
            mat = scn.add_material(S.Material(S.MATERIAL_REGULAR, (scn.add_texture(S.constant_texture([*col, 1])), t_flat,
                                                                   scn.add_texture(S.constant_texture([rng.uniform(0.1, 0.9), float(k % 2), 1, 1])))))
        scn.objects.append(S.Object(sph, mat, model))
    if light:
        scn.lights.append(S.Light(S.LIGHT_POINT, (30.0, 28.0, 25.0), (2.0, 6.0, 3.0), radius=20.0))
    if emissive:
        env = np.zeros((32, 64, 4), F32)
        env[..., 0] = np.linspace(0.2, 1.2, 64)[None, :]
        env[..., 1] = np.linspace(0.9, 0.1, 32)[:, None]
        env[..., 2] = 0.6
        env[..., 3] = 1
        scn.background = S.Background(S.BG_LATLONG, texture=scn.add_texture(env))
    else:
        scn.background = S.Background(S.BG_COLOR, (0.2, 0.25, 0.3))

    def cam_of(w, h):
        from rendu_amd.camera import Camera
        c = Camera()
        c.pose([0.0, 3.0, 9.0], [0.0, 0.5, 0.0], [0.0, 1.0, 0.0])
        c.projection(ratio=w / h, fov=0.9)
        return c

    return scn, cam_of


def both(scn, cam, w, h, spp, depth, seed=3, **kw):
    imgs = []
    for inst in (False, True):
        img = np.zeros((h, w, 3), F32)
        st = PathTracer(scn, instancing=inst).render(cam, spp, depth, img, seed=seed, **kw)
        imgs.append((img, st))
    return imgs


@pytest.mark.parametrize("depth", [2, 5])
def test_instanced_render_matches_flattened(depth):
    scn, cam_of = shared_mesh_scene()
    w, h = 96, 64
    log_case(f"instanced vs flattened, {len(scn.objects)} objects of 2 meshes, 96x64 16spp depth {depth}")
    (a, sa), (b, sb) = both(scn, cam_of(w, h), w, h, 16, depth)
    assert a.mean() > 0.01 and sa["shadow_rays"] > 0 and sb["shadow_rays"] > 0
    close = np.abs(a - b) <= 1e-4 * np.abs(a) + 1e-5
    # Depth 2: camera rays, the light's shadow rays and one bounce; deeper paths may branch where
    # an ulp of a hit point flips a later choice.
    assert close.mean() >= (0.99 if depth == 2 else 0.97), close.mean()
    assert abs(float(a.mean()) - float(b.mean())) <= 2e-3 * float(a.mean())
    assert abs(sa["extension_rays"] - sb["extension_rays"]) <= 2e-3 * sa["extension_rays"]


def test_instanced_extended_estimator_with_emissive_instances():
    """Emissive instances (their world triangles sampled by the extended estimator), a lat-long
    sky, one and three passes in flight."""
    scn, cam_of = shared_mesh_scene(seed=9, emissive=True, light=False)
    w, h = 80, 60
    log_case("instanced vs flattened, emissive instances, extended estimator, 80x60 32spp depth 4")
    (a, _), (b, _) = both(scn, cam_of(w, h), w, h, 32, 4, estimator=S.ESTIMATOR_EXTENDED, rr_start=2,
                           max_paths=w * h * 8)
    close = np.abs(a - b) <= 1e-4 * np.abs(a) + 1e-5
    assert close.mean() >= 0.95, close.mean()
    assert abs(float(a.mean()) - float(b.mean())) <= 5e-3 * float(a.mean())
    c = np.zeros_like(b)
    PathTracer(scn, instancing=True).render(cam_of(w, h), 32, 4, c, seed=3, estimator=S.ESTIMATOR_EXTENDED, rr_start=2,
                                            max_paths=w * h * 8, passes_in_flight=1)
    np.testing.assert_array_equal(b, c)


def test_instanced_frame_scene_matches_flattened():
    """The single frame's scene (terrain + 2000 instances of one sphere mesh) at a small size."""
    scn, cam_of = scenes.terrain_spheres()
    w, h = 96, 54
    log_case("instanced vs flattened, terrain + 2000 sphere instances, 96x54 4spp depth 8 extended")
    (a, _), (b, _) = both(scn, cam_of(w, h), w, h, 4, 8, seed=4, estimator=S.ESTIMATOR_EXTENDED, rr_start=3)
    close = np.abs(a - b) <= 1e-4 * np.abs(a) + 1e-5
    assert close.mean() >= 0.95, close.mean()
    assert abs(float(a.mean()) - float(b.mean())) <= 1e-2 * float(a.mean())


def test_instanced_scene_without_shared_meshes_renders_as_flattened():
    """Meshes used by one object each are baked into world space as Rendu bakes them, and their
    objects enter the top level untransformed: the render equals the flattened scene's bit for bit."""
    scn, cam_of = shared_mesh_scene(seed=7, n=12)
    # Every object its own copy of the sphere mesh: nothing is shared.
    sph = scn.objects[1].mesh
    for o in scn.objects[1:]:
        o.mesh = scn.add_mesh(scn.meshes[sph])
    w, h = 80, 60
    log_case("instanced vs flattened, 13 objects without shared meshes, 80x60 16spp depth 4")
    (a, sa), (b, sb) = both(scn, cam_of(w, h), w, h, 16, 4)
    assert a.mean() > 0.01 and sa["shadow_rays"] > 0
    np.testing.assert_array_equal(a, b)
    assert sa["extension_rays"] == sb["extension_rays"] and sa["shadow_rays"] == sb["shadow_rays"]


def test_instanced_identity_object_of_a_shared_mesh_renders_as_flattened():
    """A shared mesh whose visible object has an identity model (entered untransformed through its
    record, then left at the marker) next to world-space meshes (the ground, hung in the top level
    as it is): hits keep the right object on every path, bit for bit as in the flattened scene. The
    mesh's other object is a tiny transformed copy far above, which no ray reaches."""
    scn, cam_of = shared_mesh_scene(seed=8, n=6)
    sph = scn.objects[1].mesh
    mat = scn.objects[1].material
    far = np.diag([1e-3, 1e-3, 1e-3, 1.0]).astype(F32)
    far[:3, 3] = (0.0, 1000.0, 0.0)
    scn.objects[1] = S.Object(sph, mat, np.eye(4, dtype=F32))
    for o in scn.objects[2:]:
        o.mesh = scn.add_mesh(scn.meshes[sph])
    scn.objects.append(S.Object(sph, mat, far))
    w, h = 80, 60
    log_case("instanced vs flattened, identity-model object of a shared mesh + world-space meshes, 80x60 16spp depth 4")
    (a, sa), (b, sb) = both(scn, cam_of(w, h), w, h, 16, 4)
    assert a.mean() > 0.01
    np.testing.assert_array_equal(a, b)
    assert sa["extension_rays"] == sb["extension_rays"] and sa["shadow_rays"] == sb["shadow_rays"]


def test_instanced_cutout_layers_shadow_walk():
    """Alpha cut-out layers as objects of one shared quad mesh (translated copies): the masked
    shadow walk steps through each layer's instance, and the render equals the flattened scene's
    (translations round differently in object space, so the comparison is to 1e-4)."""
    scn, cam_of = scenes.masked_layers(4)
    layer = quad((-200, 0, 200), (200, 0, 200), (200, 0, -200), (-200, 0, -200))
    shared = scn.add_mesh(layer)
    for o in scn.objects[1:5]:
        y = float(scn.meshes[o.mesh].positions[0][1])
        o.mesh = shared
        o.model = np.eye(4, dtype=F32)
        o.model[1, 3] = y
    w, h = 64, 48
    log_case("instanced vs flattened, 4 cut-out layers of one shared quad, 64x48 8spp depth 2")
    (a, sa), (b, sb) = both(scn, cam_of(w, h), w, h, 8, 2)
    assert a.mean() > 0.01 and sa["shadow_rays"] > 0 and sb["shadow_rays"] > 0
    close = np.abs(a - b) <= 1e-4 * np.abs(a) + 1e-5
    assert close.mean() >= 0.99, close.mean()
    assert abs(float(a.mean()) - float(b.mean())) <= 1e-3 * float(a.mean())


def test_instanced_scene_autotune_graph_replay_and_cache():
    """The rest of the render API on an instanced scene: tuned launch parameters and graph-replayed
    passes change no bit of the image; a cached render (radiance cache from bounce 2) stays close
    to the plain one."""
    from rendu_amd.rcache import RadianceCache
    scn, cam_of = shared_mesh_scene(seed=11, n=24)
    w, h = 64, 48
    cam = cam_of(w, h)
    log_case("instanced scene: autotune, graph replay (4 passes), cached render, 64x48 16spp depth 4")
    pt = PathTracer(scn, instancing=True)
    kw = dict(seed=6, estimator=S.ESTIMATOR_EXTENDED, rr_start=2, max_paths=w * h * 4, passes_in_flight=1)
    ref = np.zeros((h, w, 3), F32)
    pt.render(cam, 16, 4, ref, **kw)
    assert ref.mean() > 0.01
    rep = pt.autotune(cam, 16, 4, w, h, seed=6, estimator=S.ESTIMATOR_EXTENDED, rr_start=2, max_paths=w * h * 4)
    assert rep["candidates"] > 0
    tuned = np.zeros_like(ref)
    pt.render(cam, 16, 4, tuned, **kw)
    np.testing.assert_array_equal(ref, tuned)
    pt.set_graph(True)
    for _ in range(2):
        rep_img = np.zeros_like(ref)
        st = pt.render(cam, 16, 4, rep_img, **kw)
        np.testing.assert_array_equal(ref, rep_img)
    assert st["graph_passes"] > 0
    pt.set_graph(False)
    cached = np.zeros_like(ref)
    cst = pt.render_cached(RadianceCache(*S.scene_bounds(scn), seed=2), cam, 16, 4, cached, seed=6,
                           estimator=S.ESTIMATOR_EXTENDED, rr_start=2)
    assert np.isfinite(cached).all() and cst["cache_queries"] >= 0
    assert abs(float(cached.mean()) - float(ref.mean())) <= 0.1 * float(ref.mean())


@pytest.mark.parametrize("max_paths", [0, 2500], ids=["default-passes", "small-passes"])
def test_instanced_jobs_in_one_batch_call(max_paths):
    """Instanced scenes in rdu_pathtracer_render_batch beside flattened ones (instance records,
    per-mesh trees and top levels rebased into the call's arena): every image equals the job
    rendered alone, bit for bit."""
    a, cam_a = shared_mesh_scene(seed=12, n=16)
    b, cam_b = shared_mesh_scene(seed=13, n=10, emissive=True, light=False)
    c, cam_c, w_c, h_c, spp_c, depth_c = scenes.random_job(7)
    # Cut-out layers sharing one quad: the masked shadow walk runs on the arena with instances.
    d, cam_d = scenes.masked_layers(3)
    shared = d.add_mesh(quad((-200, 0, 200), (200, 0, 200), (200, 0, -200), (-200, 0, -200)))
    for o in d.objects[1:4]:
        y = float(d.meshes[o.mesh].positions[0][1])
        o.mesh = shared
        o.model = np.eye(4, dtype=F32)
        o.model[1, 3] = y
    js = [J.Job(a, cam_a(40, 30), 40, 30, 8, 4, seed=21),
          J.Job(c, cam_c(32, 32), 32, 32, 4, depth_c, seed=22),
          J.Job(b, cam_b(36, 24), 36, 24, 8, 4, seed=23, estimator=S.ESTIMATOR_EXTENDED, rr_start=2),
          J.Job(d, cam_d(24, 20), 24, 20, 4, 2, seed=24)]
    pts = [PathTracer(a, instancing=True), PathTracer(c), PathTracer(b, instancing=True), PathTracer(d, instancing=True)]
    log_case(f"batch of 3 instanced (one with cut-outs) + 1 flattened jobs, max_paths {max_paths}, vs alone")
    alone = []
    for j, pt in zip(js, pts):
        img = np.zeros((j.height, j.width, 3), F32)
        pt.render(j.camera, j.samples, j.depth, img, seed=j.seed, estimator=j.estimator, rr_start=j.rr_start)
        alone.append(img)
    images, secs, st = J.render_batch(list(zip(pts, js)), max_paths=max_paths)
    for k, (x, y) in enumerate(zip(alone, images)):
        assert x.mean() > 0.0, k
        assert np.array_equal(x, y), (k, float(np.abs(x - y).max()))
    assert st["passes"] >= (1 if max_paths == 0 else 3)


def test_instanced_scene_limits():
    scn, cam_of = shared_mesh_scene(n=4)
    pt = PathTracer(scn, instancing=True)
    with pytest.raises(RuntimeError):
        pt.raycaster()
    from rendu_amd._lib import RduError
    with pytest.raises(RduError):
        PathTracer(scn, gpu_build=True, instancing=True)
