"""Scene ingestion (rendu_amd/ingest.py): Rendu's .scene tokens, OBJ Indexed loading, transforms,
normals, PNG textures, and Rendu's own Cornell box scene (tests/golden/cornellbox: the reference's
scene, meshes and textures, copied as data) rendered on the GPU against the oracle."""
import math
import os

import numpy as np
import pytest

from rendu_amd import ingest
from rendu_amd import scene as S

GOLD = os.path.join(os.path.dirname(__file__), "golden", "cornellbox")


def test_decode_tokens_nesting_and_arrays():
    text = """# comment
* object: # trailing comment
\tmesh: box
\ttranslation: 1.0, 2.0 3.0
* material:
\ttextures:
\t\t- srgb: red
\t\t- rgb: 0.5,0.5,1.0
\tname: m
* background:
\tcube: rgbcube: grey_cube
no colon here
"""
    toks = ingest.decode(text)
    assert [t.key for t in toks] == ["object", "material", "background"]
    assert toks[0].elements[1].values == ["1.0", "2.0", "3.0"]
    tex = toks[1].elements[0]
    assert tex.key == "textures" and [e.key for e in tex.elements] == ["srgb", "rgb"]
    assert tex.elements[1].values == ["0.5", "0.5", "1.0"]
    cube = toks[2].elements[0]
    assert cube.key == "cube" and cube.elements[0].key == "rgbcube" and cube.elements[0].values == ["grey_cube"]


def test_obj_indexed_sharing_and_uv_flip():
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvt 1 0.25\nvn 0 0 1\n" \
           "f 1/1/1 2/2/1 3/1/1\nf 2/2/1 4/1/1 3/1/1 1/1/1\n"
    m = ingest.load_obj(text)
    # Distinct 'v/vt/vn' strings: 1/1/1 2/2/1 3/1/1 4/1/1; the fourth vertex of a face is ignored.
    assert m.positions.tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]
    assert m.indices.tolist() == [0, 1, 2, 1, 3, 2]
    np.testing.assert_array_equal(m.texcoords, [[0, 1], [1, 0.75], [0, 1], [0, 1]])
    assert np.all(m.normals == [0, 0, 1])


def test_obj_uv_index_substring_quirk():
    """'v/vt' faces: the uv index is read from foundL characters after the first slash."""
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\n" + "".join(f"vt {k} 0\n" for k in range(12)) + "f 1/12 2/2 3/3\n"
    m = ingest.load_obj(text)
    assert m.texcoords[0, 0] == 0.0  # '12' truncated to '1'
    assert m.texcoords[1, 0] == 1.0 and m.texcoords[2, 0] == 2.0


def test_compute_normals_and_tangent_frame():
    m = ingest.load_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n")
    assert m.normals is None
    ingest.compute_normals(m)
    np.testing.assert_allclose(m.normals, [[0, 0, 1]] * 3, atol=1e-7)
    ingest.compute_tangents(m)
    np.testing.assert_allclose(m.tangents, [[1, 0, 0]] * 3, atol=1e-6)
    # v is flipped on load (v' = 1 - v): the texture area is negative, MikkTSpace's sign -1, and
    # Rendu's bitangent -sign * cross(n, t) (Mesh.cpp:465) is +y.
    np.testing.assert_allclose(m.bitangents, [[0, 1, 0]] * 3, atol=1e-6)


def test_transformation_trs():
    toks = ingest.decode("* object:\n\torientation: 0 2 0 1.5707963\n\ttranslation: 1 2 3\n\tscaling: 2\n")
    M = ingest.transformation(toks[0].elements)
    p = M @ np.array([1.0, 0.0, 0.0, 1.0])
    # Rotation by +90 degrees about +y sends +x to -z; then scale 2 and translate.
    np.testing.assert_allclose(p[:3], [1.0, 2.0, 1.0], atol=1e-6)
    np.testing.assert_allclose(M[:3, :3] @ M[:3, :3].T, 4 * np.eye(3), atol=1e-5)


def test_cornellbox_scene_loads():
    scn, cam_of = ingest.load_scene(os.path.join(GOLD, "cornellbox.scene"), [GOLD])
    assert len(scn.objects) == 8 and len(scn.materials) == 5 and len(scn.lights) == 1
    l = scn.lights[0]
    assert l.type == S.LIGHT_POINT and np.allclose(l.intensity, [9.0, 8.4, 6.0]) and l.radius == 3.0
    assert scn.background.mode == S.BG_COLOR
    tris = sum(scn.meshes[o.mesh].triangle_count for o in scn.objects)
    assert tris == 32
    c = cam_of(256, 256)
    assert np.allclose(c.position, [0, 1, 2.5]) and abs(c.fov - 1.3) < 1e-7
    # PNG texels are bytes / 255, 4 channels.
    grey = ingest.texture(ingest.Resources([GOLD]), "cornellbox_grey")
    assert grey.shape == (8, 8, 4) and np.all(grey[..., 3] == 1.0)
    assert np.all(np.abs(grey[..., :3] * 255 - np.round(grey[..., :3] * 255)) < 1e-4)


COMMON = os.path.join(os.path.dirname(__file__), "golden", "common")
SUN = "-0.287348 0.000000 0.957826"  # the sun of Rendu's pbrdemo/cornellbox-reflections.scene


def atmosphere_scene(tmp_path):
    """Rendu's Cornell box with the background of its cornellbox-reflections scene (an atmosphere
    with that scene's sun), so the sky is seen through the open front."""
    text = open(os.path.join(GOLD, "cornellbox.scene")).read()
    text = text.replace("\tcolor: 0.000000 0.000000 0.000000 \n", f"\tsun: {SUN} \n", 1)
    assert f"sun: {SUN}" in text
    p = tmp_path / "cornellbox-sky.scene"
    p.write_text(text)
    return ingest.load_scene(str(p), [GOLD, COMMON])


def test_rendu_scattering_table_reads():
    """Rendu's precomputed scattering table (resources/common, copied as data): an uncompressed
    512 x 512 half-float EXR stored A, B, G, R; Image::loadHDR picks channels by name."""
    from rendu_amd.output import read_exr
    path = os.path.join(COMMON, "scattering-precomputed.exr")
    chans = read_exr(open(path, "rb").read())
    assert sorted(chans) == ["A", "B", "G", "R"]
    img = ingest.texture(ingest.Resources([COMMON]), "scattering-precomputed")
    assert img.shape == (512, 512, 4) and img.dtype == np.float32 and np.isfinite(img).all()
    for k, name in enumerate("RGBA"):
        np.testing.assert_array_equal(img[..., k], chans[name])
    # A transmittance-like table: every entry in [0, 1], alpha 1, and not constant.
    assert img[..., :3].min() >= 0.0 and img[..., :3].max() <= 1.0 and np.all(img[..., 3] == 1.0)
    assert img[..., :3].std() > 0.05
    # half values exactly (the file holds halves)
    assert np.array_equal(img, img.astype(np.float16).astype(np.float32))


def test_atmosphere_background_ingested(tmp_path):
    scn, cam_of = atmosphere_scene(tmp_path)
    b = scn.background
    assert b.mode == S.BG_ATMOSPHERE
    sun = np.array([float(v) for v in SUN.split()], np.float32)
    np.testing.assert_allclose(b.sun_direction, sun / np.linalg.norm(sun), rtol=1e-6)
    assert scn.textures[b.texture].shape == (512, 512, 4)
    bad = tmp_path / "null-sun.scene"
    bad.write_text(open(os.path.join(GOLD, "cornellbox.scene")).read().replace(
        "\tcolor: 0.000000 0.000000 0.000000 \n", "\tsun: 0 0 0 \n", 1))
    with pytest.raises(ValueError):
        ingest.load_scene(str(bad), [GOLD, COMMON])


@pytest.mark.gpu
def test_atmosphere_scene_with_rendus_table_gpu_vs_oracle(tmp_path):
    from conftest import log_case
    from oracle import oracle as orc
    from rendu_amd.pathtracer import PathTracer
    scn, cam_of = atmosphere_scene(tmp_path)
    log_case("Rendu cornellbox.scene with the reflections scene's atmosphere, 48x48 16spp depth5")
    cam = cam_of(48, 48)
    got = np.zeros((48, 48, 3), np.float32)
    PathTracer(scn).render(cam, 16, 5, got, seed=2)
    exp = np.zeros_like(got)
    orc.OraclePathTracer(scn).render(S.camera_desc(cam), 16, 5, exp, S.RenderOpts(2, 0, 3, 0))
    dark = np.zeros_like(got)
    scn.background = S.Background(S.BG_COLOR, (0.0, 0.0, 0.0))
    PathTracer(scn).render(cam, 16, 5, dark, seed=2)
    assert exp.mean() > dark.mean() + 1e-3  # the sky lights the box
    d = np.abs(got.astype(np.float64) - exp)
    assert (d <= 1e-4 * np.abs(exp) + 1e-6).mean() >= 0.97
    assert d.mean() <= 1e-3 * np.abs(exp).mean()


@pytest.mark.gpu
def test_cornellbox_scene_gpu_vs_oracle():
    from conftest import log_case
    from oracle import oracle as orc
    from rendu_amd.pathtracer import PathTracer
    scn, cam_of = ingest.load_scene(os.path.join(GOLD, "cornellbox.scene"), [GOLD])
    log_case("Rendu cornellbox.scene 64x64 16spp depth5")
    cam = cam_of(64, 64)
    got = np.zeros((64, 64, 3), np.float32)
    PathTracer(scn).render(cam, 16, 5, got, seed=1)
    exp = np.zeros_like(got)
    orc.OraclePathTracer(scn).render(S.camera_desc(cam), 16, 5, exp, S.RenderOpts(1, 0, 3, 0))
    assert exp.mean() > 0.01
    d = np.abs(got.astype(np.float64) - exp)
    assert (d <= 1e-4 * np.abs(exp) + 1e-6).mean() >= 0.98
    assert d.mean() <= 1e-3 * np.abs(exp).mean()


PBR = os.path.join(os.path.dirname(__file__), "golden", "pbrdemo")
ROOTS = [GOLD, COMMON, PBR]


def test_furnace_scene_loads():
    """resources/pbrdemo/furnace-test.scene: 22 spheres (radius 1) in rows of 5 and 6, dielectrics
    then metals, roughness 0.0 ... 1.0 from left to right, white albedo, in Rendu's grey cubemap (the
    first level of grey_cube_0_*.exr, 0.5 everywhere)."""
    scn, cam_of = ingest.load_scene(os.path.join(PBR, "furnace-test.scene"), ROOTS)
    assert len(scn.objects) == 22 and len(scn.materials) == 22 and not scn.lights
    assert scn.background.mode == S.BG_SKYBOX
    for f in scn.background.cube:
        assert np.all(scn.textures[f][..., :3] == 0.5)
    rough = [float(scn.textures[m.textures[2]][0, 0, 0]) for m in scn.materials]
    metal = [float(scn.textures[m.textures[2]][0, 0, 1]) for m in scn.materials]
    assert rough == pytest.approx([k / 10 for k in range(11)] * 2)
    assert metal == [0.0] * 11 + [1.0] * 11
    assert all(np.all(scn.textures[m.textures[0]][..., :3] == 1.0) for m in scn.materials)
    ys = sorted({round(float(np.asarray(o.model)[1, 3]), 3) for o in scn.objects})
    assert ys == [-4.5, -1.5, 1.5, 4.5]
    assert not any(o.use_texcoords for o in scn.objects)
    cam = cam_of(64, 48)
    assert np.allclose(cam.position, [23.2, 0, 0]) and cam.fov == pytest.approx(0.57)


def test_emissive_cornellbox_loads():
    """resources/pbrdemo/cornellbox-emissive.scene: three emissive walls (radiance 1.25 in one
    channel), an emissive box textured with light_grid.exr (uses its UVs), no analytic light."""
    scn, _ = ingest.load_scene(os.path.join(PBR, "cornellbox-emissive.scene"), ROOTS)
    assert len(scn.objects) == 7 and not scn.lights
    emissive = [m for m in scn.materials if m.type == S.MATERIAL_EMISSIVE]
    assert len(emissive) == 4
    cols = sorted(tuple(scn.textures[m.textures[0]][0, 0, :3].round(3)) for m in emissive)
    assert (0.0, 0.0, 1.25) in cols and (0.0, 1.25, 0.0) in cols and (1.25, 0.0, 0.0) in cols
    grid = [m for m in emissive if scn.textures[m.textures[0]].shape[:2] == (32, 32)]
    assert len(grid) == 1 and scn.textures[grid[0].textures[0]][..., 0].max() > 3.0
    # Objects are sorted by material type: the emissive ones come last, the grid box uses UVs.
    types = [scn.materials[o.material].type for o in scn.objects]
    assert types == sorted(types) and sum(o.use_texcoords for o in scn.objects) == 1


def test_cutout_cornellbox_loads():
    """resources/pbrdemo/cornellbox-cutout.scene: an alpha-masked leaves material (palette PNG
    with transparency, JPEG normal and roughness maps), a vase, two spot lights with their cones."""
    scn, _ = ingest.load_scene(os.path.join(PBR, "cornellbox-cutout.scene"), ROOTS)
    masked = [m for m in scn.materials if m.masked]
    assert len(masked) == 1
    alpha = scn.textures[masked[0].textures[0]][..., 3]
    assert alpha.min() == 0.0 and alpha.max() == 1.0
    assert [l.type for l in scn.lights] == [S.LIGHT_SPOT, S.LIGHT_SPOT]
    assert scn.lights[0].angles == pytest.approx((0.5, 0.6)) and scn.lights[1].angles == pytest.approx((0.35, 0.4))
    assert scn.lights[0].radius == 5.0 and scn.lights[0].intensity == pytest.approx((6.0, 5.4, 3.0))


def test_parallax_cornellbox_loads():
    """resources/pbrdemo/cornellbox-parallax.scene: two Parallax materials of four JPEG textures
    (colour, normal, roughness/metalness/AO, depth; the path tracer reads the first three), a
    rotated and scaled sphere, one spot light, a colour background."""
    scn, cam_of = ingest.load_scene(os.path.join(PBR, "cornellbox-parallax.scene"), ROOTS)
    assert len(scn.objects) == 3 and [l.type for l in scn.lights] == [S.LIGHT_SPOT]
    par = [m for m in scn.materials if m.type == ingest.MATERIAL_TYPES["Parallax"]]
    assert len(par) == 2
    for m in par:
        col = scn.textures[m.textures[0]]
        assert col.shape == (512, 512, 4) and 0.0 <= col.min() and col.max() <= 1.0 and col[..., :3].std() > 0.01
    assert scn.background.mode == S.BG_COLOR and scn.background.color == pytest.approx((0.039,) * 3)
    sphere = [o for o in scn.objects if np.asarray(o.model)[1, 3] == pytest.approx(0.75)]
    assert len(sphere) == 1 and sphere[0].use_texcoords


def test_textured_cornellbox_loads():
    """resources/pbrdemo/cornellbox-textured.scene: wood, tiles and brushed metal JPEG maps, painted
    walls, an emissive light, and the kiara_dawn skybox (first level: 512^2 half-float EXR faces)."""
    scn, _ = ingest.load_scene(os.path.join(PBR, "cornellbox-textured.scene"), ROOTS)
    assert len(scn.objects) == 8 and len(scn.materials) == 7
    assert scn.background.mode == S.BG_SKYBOX and len(scn.background.cube) == 6
    faces = [scn.textures[f] for f in scn.background.cube]
    assert all(f.shape[:2] == (512, 512) and f[..., :3].std() > 0.01 for f in faces)
    jpeg = [scn.textures[m.textures[0]] for m in scn.materials if scn.textures[m.textures[0]].shape[0] > 8]
    assert len(jpeg) == 3
