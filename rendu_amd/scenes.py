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
