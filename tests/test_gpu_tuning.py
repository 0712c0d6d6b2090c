"""Autotuner, tuning cache and graph replay on the GPU.

The launch parameters (refill threshold, ray tile, workgroups per CU) only schedule rays onto
wavefronts: each ray's traversal is the same whatever they are, so tuned queries and renders must
equal untuned ones bit for bit, and a pass replayed from a captured HIP graph must equal the same
pass run eagerly bit for bit.
"""
import numpy as np
import pytest

from conftest import log_case
from rendu_amd import scenes, tuning, workloads
from rendu_amd.pathtracer import PathTracer
from rendu_amd.raycaster import Raycaster

pytestmark = pytest.mark.gpu
F32 = np.float32

INCOHERENT, ANY, GPU_TREE, PATH_PASS = 1, 2, 4, 8


@pytest.fixture(autouse=True)
def _empty_cache():
    tuning.clear()
    yield
    tuning.clear()


def floor_log2(v):
    return max(int(v), 1).bit_length() - 1


def key_line(rc, batch, cls, params):
    inf = rc.info()
    return " ".join(str(v) for v in (floor_log2(inf["triangles"]), inf["wide_depth"], floor_log2(inf["meshes"]),
                                     floor_log2(batch), cls) + tuple(params))


@pytest.fixture(scope="module")
def soup():
    rc = Raycaster()
    rc.addMesh(workloads.triangle_soup(200_000, seed=11))
    rc.updateHierarchy()
    return rc


def same_hits(a, b):
    for f in ("t", "u", "v", "hit", "mesh_id", "local_id"):
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("params", [(4, 32, 1), (64, 512, 2), (0, 16, 0), (24, 0, 3)])
@pytest.mark.parametrize("n,incoherent", [(1 << 17, True), (5000, True), (70000, False)])
def test_loaded_parameters_keep_results(soup, tmp_path, params, n, incoherent):
    """Every launch shape the tuner can pick, loaded through the tuning cache, gives the default hits."""
    rays = workloads.soup_rays(0 if incoherent else n, n if incoherent else 0, seed=7)
    log_case(f"tuning file params={params} n={n} incoherent={incoherent}")
    base_hits = soup.intersect_batch(rays, incoherent=incoherent)
    base_any = soup.intersect_any_batch(rays, incoherent=incoherent)
    assert soup.tuning(n, incoherent=incoherent)["source"] == 0
    cls = INCOHERENT if incoherent else 0
    f = tmp_path / "tune.txt"
    f.write_text("rdu-tune 2\n" + key_line(soup, n, cls, params) + "\n" + key_line(soup, n, cls | ANY, params) + "\n")
    tuning.load(f)
    got = soup.tuning(n, incoherent=incoherent)
    assert got["source"] == 2
    assert (got["refill_min"], got["chunk"], got["blocks_per_cu"]) == params
    assert soup.tuning(n, any_hit=True, incoherent=incoherent)["source"] == 2
    same_hits(soup.intersect_batch(rays, incoherent=incoherent), base_hits)
    assert np.array_equal(soup.intersect_any_batch(rays, incoherent=incoherent), base_any)


def test_autotune_then_cache_serves_a_new_raycaster(tmp_path):
    mesh = workloads.triangle_soup(100_000, seed=12)
    rays = workloads.soup_rays(0, 1 << 17, seed=8)
    a = Raycaster()
    a.addMesh(mesh)
    a.updateHierarchy()
    base = a.intersect_batch(rays, incoherent=True)
    log_case("autotune closest incoherent 2^17 on 1e5 soup")
    rep = a.autotune(rays, incoherent=True)
    assert rep["candidates"] >= 10 and rep["best_ms"] > 0 and rep["default_ms"] > 0
    assert rep["best_ms"] <= rep["default_ms"] * 1.0001
    assert a.tuning(1 << 17, incoherent=True)["source"] == 1
    same_hits(a.intersect_batch(rays, incoherent=True), base)
    assert tuning.size() == 1
    # Another raycaster over a scene with the same statistics finds the entry.
    b = Raycaster()
    b.addMesh(workloads.triangle_soup(100_000, seed=13))
    b.updateHierarchy()
    if b.info()["wide_depth"] == a.info()["wide_depth"]:
        t = b.tuning((1 << 17) + 5, incoherent=True)
        assert t["source"] == 2
        assert (t["refill_min"], t["chunk"], t["blocks_per_cu"]) == (rep["refill_min"], rep["chunk"], rep["blocks_per_cu"])
    # Other classes and batch sizes are not served by it.
    assert b.tuning(1 << 17, any_hit=True, incoherent=True)["source"] == 0
    assert b.tuning(1 << 12, incoherent=True)["source"] == 0
    # Through a file into an empty cache.
    f = tmp_path / "t.txt"
    tuning.save(f)
    tuning.clear()
    assert a.tuning(1 << 17, incoherent=True)["source"] == 1  # the raycaster keeps its own
    tuning.load(f)
    assert tuning.size() == 1


def cornell(w, h):
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=2, point_light=True)
    return scn, cam_of(w, h)


def test_scene_autotune_keeps_the_image():
    scn, cam = cornell(64, 48)
    pt = PathTracer(scn)
    ref = np.zeros((48, 64, 3), F32)
    pt.render(cam, 8, 5, ref, seed=9)
    log_case("scene autotune cornell 64x48 8spp")
    rep = pt.autotune(cam, 8, 5, 64, 48, seed=9)
    assert rep["candidates"] >= 10 and rep["best_ms"] > 0
    assert tuning.size() == 1
    got = np.zeros((48, 64, 3), F32)
    pt.render(cam, 8, 5, got, seed=9)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("spp,per_pass", [(16, 4), (16, 3), (8, 8)])
def test_graph_replay_equals_eager(spp, per_pass):
    """Full passes replay from a captured graph (a short last pass runs eagerly); every render equals
    the eager render bit for bit, across renders, with eager renders in between."""
    w, h = 48, 40
    scn, cam = cornell(w, h)
    eager = PathTracer(scn)
    ref = np.zeros((h, w, 3), F32)
    st = eager.render(cam, spp, 5, ref, seed=4, max_paths=w * h * per_pass)
    passes = -(-spp // per_pass)
    assert st["passes"] == passes and st["graph_passes"] == 0
    graphed = PathTracer(scn)
    graphed.set_graph(True)
    full = spp // per_pass
    for k in range(4):
        log_case(f"graph replay spp={spp} per_pass={per_pass} render {k}")
        img = np.zeros((h, w, 3), F32)
        st = graphed.render(cam, spp, 5, img, seed=4, max_paths=w * h * per_pass)
        assert np.array_equal(img, ref), f"render {k}"
        assert st["passes"] == passes
        # First render: the first full pass runs eagerly, the second is captured; later renders
        # replay every full pass.
        want = max(full - 1, 0) if k == 0 else full
        assert st["graph_passes"] == want, st
        assert st["graph_captures"] == (1 if (k == 0 and full >= 2) or (k == 1 and full == 1) else 0), st
        again = np.zeros((h, w, 3), F32)
        eager.render(cam, spp, 5, again, seed=4, max_paths=w * h * per_pass)
        assert np.array_equal(again, ref)


def test_graph_replays_across_seeds_and_recaptures_on_a_new_frame():
    """The seed lives in device memory: progressive renders with new seeds replay the same graph.
    Another image size is another configuration: it is captured afresh, and so is the first again."""
    scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=2, point_light=True)
    sizes = {(32, 32): cam_of(32, 32), (24, 40): cam_of(24, 40)}
    refs = {}
    for (w, h), cam in sizes.items():
        for seed in (1, 2):
            ref = np.zeros((h, w, 3), F32)
            PathTracer(scn).render(cam, 8, 4, ref, seed=seed, max_paths=w * h * 2)
            refs[(w, h, seed)] = ref
    pt = PathTracer(scn)
    pt.set_graph(True)
    plan = [((32, 32), 1, 1, 3), ((32, 32), 2, 0, 4), ((32, 32), 1, 0, 4), ((24, 40), 2, 1, 3), ((24, 40), 1, 0, 4),
            ((32, 32), 2, 1, 3)]
    for (w, h), seed, captures, replays in plan:
        log_case(f"graph {w}x{h} seed {seed}")
        img = np.zeros((h, w, 3), F32)
        st = pt.render(sizes[(w, h)], 8, 4, img, seed=seed, max_paths=w * h * 2)
        assert np.array_equal(img, refs[(w, h, seed)])
        assert (st["graph_captures"], st["graph_passes"]) == (captures, replays), st
    pt.set_graph(False)
    img = np.zeros((32, 32, 3), F32)
    st = pt.render(sizes[(32, 32)], 8, 4, img, seed=1, max_paths=32 * 32 * 2)
    assert np.array_equal(img, refs[(32, 32, 1)]) and st["graph_passes"] == 0


def test_graph_replay_lights_latlong_and_eager_cut_outs():
    """A scene with analytic lights (any-hit shadow rays) and a lat-long background replays bit for
    bit. A scene with alpha cut-outs renders eagerly even with replay on (its shadow walk reads a
    count on the host between rounds of steps, so it has no fixed launch sequence): same image,
    nothing replayed."""
    from test_gpu_pathtracer import feature_scene
    from rendu_amd import scene as S
    scn, cam_of = scenes.background_probe(S.BG_LATLONG, level=0.6)
    scn.lights.append(S.Light(S.LIGHT_POINT, (3.0, 3.0, 3.0), (0.5, 1.5, 0.5), radius=8.0))
    scn.lights.append(S.Light(S.LIGHT_SPOT, (2.0, 2.0, 2.0), (0.0, 2.0, 1.0), (0.0, -1.0, 0.0), radius=6.0,
                              angles=(0.3, 0.6)))
    w, h = 40, 30
    cam = cam_of(w, h)
    ref = np.zeros((h, w, 3), F32)
    PathTracer(scn).render(cam, 8, 5, ref, seed=3, max_paths=w * h * 2)
    pt = PathTracer(scn)
    pt.set_graph(True)
    for k in range(3):
        log_case(f"graph replay lights + lat-long render {k}")
        img = np.zeros((h, w, 3), F32)
        st = pt.render(cam, 8, 5, img, seed=3, max_paths=w * h * 2)
        assert np.array_equal(img, ref)
        assert st["graph_passes"] == (3 if k == 0 else 4) and st["shadow_rays"] > 0
    scn, cam_of = feature_scene(S.BG_LATLONG)
    cam = cam_of(w, h)
    ref = np.zeros((h, w, 3), F32)
    PathTracer(scn).render(cam, 8, 5, ref, seed=3, max_paths=w * h * 2)
    pt = PathTracer(scn)
    pt.set_graph(True)
    for k in range(2):
        log_case(f"cut-out scene with replay enabled, render {k}")
        img = np.zeros((h, w, 3), F32)
        st = pt.render(cam, 8, 5, img, seed=3, max_paths=w * h * 2)
        assert np.array_equal(img, ref)
        assert st["graph_passes"] == 0 and st["graph_captures"] == 0 and st["shadow_rays"] > 0
