"""Seeded path-tracing scenes.

* cornell_spheres: a Cornell-style box in Rendu's Cornell frame (floor y = 0, ceiling y = 2, walls
  at x = +-1 and z = -1; camera from resources/pbrdemo/cornellbox.scene:67-72) built from 12 quads —
  5 walls, 1 emissive ceiling quad of radiance 15 and the 6 faces of one block — plus 3 icospheres
  at subdivision 4 at seeded positions; one Regular material per object with a constant albedo drawn
  from U[0.2, 0.8] (stored as albedo^(1/2.2), since the tracer applies pow(., 2.2)), roughness 1,
  metalness 0; black background; no analytic light.
* furnace: the white-furnace check of resources/pbrdemo/furnace-test.scene — albedo-1 spheres
  inside a constant 0.5 environment (cubemap).
"""
from __future__ import annotations

import numpy as np

from . import scene as S
from .camera import Camera
from .workloads import MeshData, box_quads, icosphere, quad

F32 = np.float32


def _regular(scn: S.Scene, albedo, roughness=1.0, metalness=0.0, masked=False) -> int:
    a = np.asarray(albedo, F32)
    col = np.concatenate([np.power(a, F32(1.0 / 2.2)), [1.0]]).astype(F32)
    t0 = scn.add_texture(S.constant_texture(col))
    t1 = scn.add_texture(S.constant_texture([0.5, 0.5, 1.0, 1.0]))  # flat normal map
    t2 = scn.add_texture(S.constant_texture([roughness, metalness, 1.0, 1.0]))
    return scn.add_material(S.Material(S.MATERIAL_REGULAR, (t0, t1, t2), masked))


def _emissive(scn: S.Scene, radiance) -> int:
    t0 = scn.add_texture(S.constant_texture(list(radiance) + [1.0]))
    t1 = scn.add_texture(S.constant_texture([0.5, 0.5, 1.0, 1.0]))
    t2 = scn.add_texture(S.constant_texture([1.0, 0.0, 1.0, 1.0]))
    return scn.add_material(S.Material(S.MATERIAL_EMISSIVE, (t0, t1, t2), False))


def cornell_camera(width: int, height: int) -> Camera:
    cam = Camera()
    cam.pose([0.0, 1.0, 2.5], [0.0, 1.0, 1.0], [0.0, 1.0, 0.0])
    cam.projection(ratio=width / height, fov=1.3)
    return cam


def cornell_spheres(seed: int = 3, subdivisions: int = 4, radiance: float = 15.0, point_light: bool = False):
    """(scene, camera factory). 15 meshes, 24 + 3 * 20 * 4^subdivisions triangles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scn = S.Scene()
    quads = [
        quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)),      # floor
        quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),      # ceiling
        quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)),    # back wall
        quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1)),    # left wall
        quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)),        # right wall
    ]
    light = quad((-0.3, 1.98, -0.3), (0.3, 1.98, -0.3), (0.3, 1.98, 0.3), (-0.3, 1.98, 0.3))
    block = box_quads((0.15, 0.0, -0.65), (0.65, 0.6, -0.15))
    for q in quads:
        m = scn.add_mesh(q)
        scn.objects.append(S.Object(m, _regular(scn, rng.uniform(0.2, 0.8, 3))))
    m = scn.add_mesh(light)
    scn.objects.append(S.Object(m, _emissive(scn, [radiance] * 3)))
    block_albedo = rng.uniform(0.2, 0.8, 3)
    for q in block:
        m = scn.add_mesh(q)
        scn.objects.append(S.Object(m, _regular(scn, block_albedo)))
    for k in range(3):
        r = float(rng.uniform(0.18, 0.3))
        c = (float(rng.uniform(-0.7, 0.1)), r + float(rng.uniform(0.0, 0.6)), float(rng.uniform(-0.6, 0.5)))
        m = scn.add_mesh(icosphere(subdivisions, radius=r, center=c))
        scn.objects.append(S.Object(m, _regular(scn, rng.uniform(0.2, 0.8, 3))))
    if point_light:
        scn.lights.append(S.Light(S.LIGHT_POINT, (9.0, 8.4, 6.0), (-0.03, 1.75, 0.0), radius=3.0))
    scn.background = S.Background(S.BG_COLOR, (0.0, 0.0, 0.0))
    return scn, cornell_camera


def fbm_heights(n: int, seed: int, octaves: int = 7) -> np.ndarray:
    """(n, n) fractal Brownian motion: bilinearly upsampled value noise, octave k at 2^(k+2) cells
    per side with amplitude 0.5^k, normalised to [0, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = np.zeros((n, n), np.float64)
    t = np.linspace(0.0, 1.0, n)
    for k in range(octaves):
        cells = 2 ** (k + 2)
        g = rng.random((cells + 1, cells + 1))
        x = t * cells
        i = np.minimum(np.floor(x).astype(int), cells - 1)
        f = x - i
        f = f * f * (3.0 - 2.0 * f)  # smoothstep between lattice values
        rows = g[i] * (1 - f)[:, None] + g[i + 1] * f[:, None]          # (n, cells+1)
        h += (rows[:, i] * (1 - f)[None, :] + rows[:, i + 1] * f[None, :]) * 0.5 ** k
    h -= h.min()
    return h / h.max()


def heightfield_mesh(heights: np.ndarray, size: float, amplitude: float) -> MeshData:
    """Grid mesh over [-size/2, size/2]^2 in xz, y = amplitude * height; 2 (n-1)^2 triangles."""
    n = heights.shape[0]
    t = np.linspace(-0.5 * size, 0.5 * size, n)
    xx, zz = np.meshgrid(t, t)
    y = heights * amplitude
    pos = np.stack([xx, y, zz], axis=-1).reshape(-1, 3).astype(F32)
    d = size / (n - 1)
    gx = np.gradient(y, d, axis=1)
    gz = np.gradient(y, d, axis=0)
    nrm = np.stack([-gx, np.ones_like(y), -gz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    tan = np.stack([np.ones_like(y), gx, np.zeros_like(y)], axis=-1)
    tan /= np.linalg.norm(tan, axis=-1, keepdims=True)
    bit = np.cross(nrm, tan)
    uv = np.stack([(xx / size + 0.5), (zz / size + 0.5)], axis=-1).reshape(-1, 2).astype(F32)
    ii, jj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (ii * n + jj).ravel()
    b, c, e = a + 1, a + n, a + n + 1
    idx = np.stack([a, c, b, b, c, e], axis=1).reshape(-1).astype(np.uint32)  # counter-clockwise seen from +y
    return MeshData(pos, idx, nrm.reshape(-1, 3).astype(F32), tan.reshape(-1, 3).astype(F32),
                    bit.reshape(-1, 3).astype(F32), uv)


def latlong_sky(width: int = 1024, height: int = 512, sun=(0.35, 0.55, -0.76), sun_radiance: float = 25.0,
                sun_angle: float = 0.05) -> np.ndarray:
    """Procedural (height, width, 4) fp32 lat-long sky: row 0 is the zenith, u follows the azimuth
    from -z towards +x (the tracer's lat-long lookup). Horizon-to-zenith gradient, dark ground below
    the horizon and a sun disk of angular radius `sun_angle`."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v[:, None] * np.pi
    phi = (u[None, :] - 0.5) * 2.0 * np.pi
    d = np.stack([np.sin(theta) * np.sin(phi), np.cos(theta) * np.ones_like(phi), -np.sin(theta) * np.cos(phi)], -1)
    up = np.clip(d[..., 1], 0.0, 1.0)[..., None]
    horizon, zenith = np.array([0.85, 0.88, 0.92]), np.array([0.22, 0.42, 0.85])
    sky = horizon * (1 - up) ** 3 + zenith * (1 - (1 - up) ** 3)
    ground = np.array([0.12, 0.1, 0.08])
    col = np.where(d[..., 1:2] >= 0, sky, ground)
    s = np.asarray(sun, np.float64)
    s /= np.linalg.norm(s)
    col = col + np.where((d @ s)[..., None] > np.cos(sun_angle), sun_radiance, 0.0)
    out = np.ones((height, width, 4), F32)
    out[..., :3] = col
    return out


def terrain_spheres(seed: int = 11, grid: int = 512, spheres: int = 2000, subdivisions: int = 3,
                    sky_size=(1024, 512)):
    """The single-frame workload: a grid x grid fBm heightfield (2 (grid-1)^2 triangles) plus
    `spheres` icospheres (20 * 4^subdivisions triangles each, one instance of a shared mesh each,
    placed by model matrix), a lat-long sky, and one Regular material per object with albedo
    U[0.1, 0.9], roughness U[0.05, 1] and metalness Bernoulli(0.3). (scene, camera factory)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    size, amp = 100.0, 9.0
    heights = fbm_heights(grid, seed)
    scn = S.Scene()
    t_flat = scn.add_texture(S.constant_texture([0.5, 0.5, 1.0, 1.0]))

    def material(albedo, rough, metal):
        a = np.asarray(albedo, F32)
        col = np.concatenate([np.power(a, F32(1.0 / 2.2)), [1.0]]).astype(F32)
        t0 = scn.add_texture(S.constant_texture(col))
        t2 = scn.add_texture(S.constant_texture([rough, metal, 1.0, 1.0]))
        return scn.add_material(S.Material(S.MATERIAL_REGULAR, (t0, t_flat, t2), False))

    m = scn.add_mesh(heightfield_mesh(heights, size, amp))
    scn.objects.append(S.Object(m, material([0.45, 0.4, 0.3], 0.9, 0.0)))
    sph = scn.add_mesh(icosphere(subdivisions, radius=1.0))
    n = grid
    for k in range(spheres):
        r = float(np.exp(rng.uniform(np.log(0.3), np.log(1.6))))
        x, z = rng.uniform(-0.45 * size, 0.45 * size, 2)
        i = int(round((z / size + 0.5) * (n - 1)))
        j = int(round((x / size + 0.5) * (n - 1)))
        y = float(heights[i, j] * amp) + 0.7 * r
        model = np.diag([r, r, r, 1.0]).astype(F32)
        model[:3, 3] = (x, y, z)
        albedo = rng.uniform(0.1, 0.9, 3)
        rough = float(rng.uniform(0.05, 1.0))
        metal = float(rng.random() < 0.3)
        scn.objects.append(S.Object(sph, material(albedo, rough, metal), model))
    scn.background = S.Background(S.BG_LATLONG, texture=scn.add_texture(latlong_sky(*sky_size)))

    def cam(width, height):
        c = Camera()
        c.pose([0.0, 16.0, 58.0], [0.0, 3.0, 0.0], [0.0, 1.0, 0.0])
        c.projection(ratio=width / height, fov=0.85)
        return c

    return scn, cam


def random_job(seed: int):
    """One job of the batched workload: (scene, camera factory, width, height, spp, depth) with a
    triangle count log-uniform on [1e4, 5e5], split as in the single frame: a fBm heightfield
    terrain of about a fifth of the triangles (2 (g-1)^2 of them) and icospheres at subdivision 3
    (1280 triangles each, one shared mesh placed by model matrix) resting on it for the rest; a
    square image whose pixel count is log-uniform on [128^2, 1024^2], spp uniform on {4..256} and
    depth 6. Materials as in terrain_spheres; a 256 x 128 lat-long sky."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tris = float(np.exp(rng.uniform(np.log(1e4), np.log(5e5))))
    side = int(round(np.exp(rng.uniform(np.log(128.0), np.log(1024.0)))))
    spp = int(rng.integers(4, 257))
    g = max(8, int(round(np.sqrt(0.2 * tris / 2.0))) + 1)
    n_sph = max(1, int(round((tris - 2 * (g - 1) ** 2) / 1280.0)))
    size = 8.0 * np.sqrt(n_sph) + 8.0
    amp = 0.12 * size
    heights = fbm_heights(g, seed + 1000)
    scn = S.Scene()
    t_flat = scn.add_texture(S.constant_texture([0.5, 0.5, 1.0, 1.0]))

    def material(albedo, rough, metal):
        a = np.asarray(albedo, F32)
        col = np.concatenate([np.power(a, F32(1.0 / 2.2)), [1.0]]).astype(F32)
        t0 = scn.add_texture(S.constant_texture(col))
        t2 = scn.add_texture(S.constant_texture([rough, metal, 1.0, 1.0]))
        return scn.add_material(S.Material(S.MATERIAL_REGULAR, (t0, t_flat, t2), False))

    m = scn.add_mesh(heightfield_mesh(heights, size, amp))
    scn.objects.append(S.Object(m, material([0.45, 0.4, 0.3], 0.9, 0.0)))
    sph = scn.add_mesh(icosphere(3, radius=1.0))
    for k in range(n_sph):
        r = float(rng.uniform(0.4, 1.2))
        x, z = rng.uniform(-0.42 * size, 0.42 * size, 2)
        i = int(round((z / size + 0.5) * (g - 1)))
        jj = int(round((x / size + 0.5) * (g - 1)))
        model = np.diag([r, r, r, 1.0]).astype(F32)
        model[:3, 3] = (x, float(heights[i, jj] * amp) + 0.7 * r, z)
        scn.objects.append(S.Object(sph, material(rng.uniform(0.1, 0.9, 3), float(rng.uniform(0.05, 1.0)),
                                                  float(rng.random() < 0.3)), model))
    scn.background = S.Background(S.BG_LATLONG, texture=scn.add_texture(latlong_sky(256, 128)))
    e = 0.5 * size

    def cam(width, height):
        c = Camera()
        c.pose([0.0, 0.9 * e, 1.6 * e], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0])
        c.projection(ratio=width / height, fov=0.9)
        return c

    return scn, cam, side, side, spp, 6


def furnace(subdivisions: int = 3, level: float = 0.5, roughness: float = 1.0, metalness: float = 0.0):
    """White furnace: an albedo-1 sphere in a constant `level` cubemap environment."""
    scn = S.Scene()
    faces = [scn.add_texture(S.constant_texture([level, level, level, 1.0])) for _ in range(6)]
    m = scn.add_mesh(icosphere(subdivisions, radius=0.5, center=(0.0, 0.0, 0.0)))
    scn.objects.append(S.Object(m, _regular(scn, [1.0, 1.0, 1.0], roughness, metalness)))
    scn.background = S.Background(S.BG_SKYBOX, cube=tuple(faces))

    def cam(width, height):
        c = Camera()
        c.pose([0.0, 0.0, 2.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0])
        c.projection(ratio=width / height, fov=0.8)
        return c

    return scn, cam


def masked_layers(layers: int, light_radius: float = 100.0):
    """Shadow-walk probe: a 400 x 400 floor at y = 0 seen from (0, 0.5, 0) straight down, a point
    light at (0, 5, 0), `layers` alpha-0 cut-out quads (masked, alpha 0 everywhere) evenly spaced
    between y = 1 and y = 4.5, and an opaque emissive-black quad at y = 6 above the light. Every
    primary ray hits the floor; its light sample crosses every layer, so Rendu's shadow walk
    (PathTracer.cpp:124-157) makes exactly layers + 1 visibility queries, and the light is seen
    only if each step shortens the remaining distance (otherwise the last query reaches y = 6)."""
    scn = S.Scene()
    e = 200.0
    m = scn.add_mesh(quad((-e, 0, e), (e, 0, e), (e, 0, -e), (-e, 0, -e)))
    scn.objects.append(S.Object(m, _regular(scn, [0.6, 0.6, 0.6])))
    t_cut = scn.add_texture(S.constant_texture([1.0, 1.0, 1.0, 0.0]))
    t_flat = scn.add_texture(S.constant_texture([0.5, 0.5, 1.0, 1.0]))
    t_rm = scn.add_texture(S.constant_texture([1.0, 0.0, 1.0, 1.0]))
    cut = scn.add_material(S.Material(S.MATERIAL_REGULAR, (t_cut, t_flat, t_rm), True))
    for k in range(layers):
        y = 1.0 + 3.5 * (k + 0.5) / max(layers, 1)
        m = scn.add_mesh(quad((-e, y, e), (e, y, e), (e, y, -e), (-e, y, -e)))
        scn.objects.append(S.Object(m, cut, use_texcoords=True))
    m = scn.add_mesh(quad((-2, 6, 2), (2, 6, 2), (2, 6, -2), (-2, 6, -2)))
    scn.objects.append(S.Object(m, _emissive(scn, [0.0, 0.0, 0.0])))
    scn.lights.append(S.Light(S.LIGHT_POINT, (4.0, 4.0, 4.0), (0.0, 5.0, 0.0), radius=light_radius))
    scn.background = S.Background(S.BG_COLOR, (0.0, 0.0, 0.0))

    def cam(width, height):
        c = Camera()
        c.pose([0.0, 0.5, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 1.0])
        c.projection(ratio=width / height, fov=1.0)
        return c

    return scn, cam


def background_probe(mode: int, level: float = 1.0):
    """A white rough floor (y = 0, 40 x 40) seen from (0, 1, 4) towards (0, 0.6, 0): the upper
    rows see the background directly, the lower rows see only the floor, which is lit by nothing
    but the background. `mode` is a Background mode; every mode shows the constant `level`
    (IMAGE: a 64 x 64 image whose G channel is the column index / 64, R and B = level)."""
    scn = S.Scene()
    e = 20.0
    m = scn.add_mesh(quad((-e, 0, e), (e, 0, e), (e, 0, -e), (-e, 0, -e)))
    scn.objects.append(S.Object(m, _regular(scn, [1.0, 1.0, 1.0])))
    c = [level, level, level, 1.0]
    if mode == S.BG_COLOR:
        scn.background = S.Background(S.BG_COLOR, (level, level, level))
    elif mode == S.BG_IMAGE:
        img = np.empty((64, 64, 4), F32)
        img[...] = c
        img[..., 1] = (np.arange(64) / 64.0).astype(F32)[None, :]
        scn.background = S.Background(S.BG_IMAGE, texture=scn.add_texture(img))
    elif mode == S.BG_SKYBOX:
        faces = [scn.add_texture(S.constant_texture(c)) for _ in range(6)]
        scn.background = S.Background(S.BG_SKYBOX, cube=tuple(faces))
    elif mode == S.BG_LATLONG:
        scn.background = S.Background(S.BG_LATLONG, texture=scn.add_texture(S.constant_texture(c)))
    else:
        raise ValueError("background_probe: COLOR, IMAGE, SKYBOX or LATLONG")

    def cam(width, height):
        cm = Camera()
        cm.pose([0.0, 1.0, 4.0], [0.0, 0.6, 0.0], [0.0, 1.0, 0.0])
        cm.projection(ratio=width / height, fov=0.9)
        return cm

    return scn, cam


def sky_spheres(seed: int = 5, sun: bool = False, point_light: bool = False, emissive: float = 0.0):
    """Extended-estimator probe: a floor and 5 spheres (roughness 0.2 ... 1, one metal) under a
    64 x 32 lat-long sky (horizon-to-zenith gradient, radiance below 1, plus a sun of radiance 25
    when `sun`), optionally a point light and an emissive quad of radiance `emissive` overhead."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scn = S.Scene()
    m = scn.add_mesh(quad((-6, 0, 6), (6, 0, 6), (6, 0, -6), (-6, 0, -6)))
    scn.objects.append(S.Object(m, _regular(scn, [0.5, 0.5, 0.5], 0.9, 0.0)))
    sph = scn.add_mesh(icosphere(2, radius=1.0))
    for k in range(5):
        model = np.diag([0.6, 0.6, 0.6, 1.0]).astype(F32)
        model[:3, 3] = (-2.4 + 1.2 * k, 0.6, float(rng.uniform(-1.0, 1.0)))
        rough = (0.2, 0.4, 0.6, 0.8, 1.0)[k]
        scn.objects.append(S.Object(sph, _regular(scn, rng.uniform(0.2, 0.8, 3), rough, 1.0 if k == 1 else 0.0), model))
    if emissive > 0.0:
        m = scn.add_mesh(quad((-1, 3, -1), (1, 3, -1), (1, 3, 1), (-1, 3, 1)))
        scn.objects.append(S.Object(m, _emissive(scn, [emissive, 0.9 * emissive, 0.7 * emissive])))
    sky = latlong_sky(64, 32, sun=(0.4, 0.6, -0.7), sun_radiance=25.0 if sun else 0.0, sun_angle=0.12)
    sky[..., :3] *= 0.8
    scn.background = S.Background(S.BG_LATLONG, texture=scn.add_texture(sky))
    if point_light:
        scn.lights.append(S.Light(S.LIGHT_POINT, (3.0, 3.0, 3.0), (1.5, 2.5, 1.5), radius=8.0))

    def cam(width, height):
        c = Camera()
        c.pose([0.0, 2.2, 6.0], [0.0, 0.5, 0.0], [0.0, 1.0, 0.0])
        c.projection(ratio=width / height, fov=0.9)
        return c

    return scn, cam
