"""Seeded synthetic scenes and ray batches for the raycaster and path tracer workloads.

All generators use numpy's PCG64 with an explicit seed, so every run on every machine with this
numpy rebuilds the same float32 inputs without shipping data files.

Workloads:
  * triangle soup: N triangles, centroids U[0,1]^3, edge length exp(-5 + 0.5 z), z ~ N(0,1),
    equilateral, random rotation from a normalised 4-D Gaussian quaternion; one mesh.
  * soup rays: half coherent pinhole primaries from (0.5, 0.5, -1) looking at (0.5, 0.5, 0.5),
    up (0, 1, 0), fov 1.3, over a W x H grid of pixel centres (Camera::pixelShifts semantics,
    src/engine/input/Camera.cpp:42-49); half incoherent, origins U[0,1]^3 and directions from
    Random::sampleSphere's formula (src/engine/generation/Random.cpp:58-63).
  * icospheres (subdivided icosahedra, 20 * 4^s triangles) and Cornell-style quads for the small
    path-tracing scene.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .camera import Camera

F32 = np.float32


@dataclass
class MeshData:
    """The Mesh fields the path uses (src/engine/resources/Mesh.hpp:126-134)."""
    positions: np.ndarray
    indices: np.ndarray
    normals: Optional[np.ndarray] = None
    tangents: Optional[np.ndarray] = None
    bitangents: Optional[np.ndarray] = None
    texcoords: Optional[np.ndarray] = None

    @property
    def triangle_count(self) -> int:
        return int(self.indices.shape[0] // 3)


def triangle_soup(n: int, seed: int = 1, log_edge: float = -5.0) -> MeshData:
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 3), dtype=np.float64).astype(F32)
    z = rng.standard_normal(n)
    edge = np.exp(log_edge + 0.5 * z)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, zz = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    # Rotation matrix of the unit quaternion (w, x, y, z).
    R = np.stack([
        np.stack([1 - 2 * (y * y + zz * zz), 2 * (x * y - w * zz), 2 * (x * zz + w * y)], -1),
        np.stack([2 * (x * y + w * zz), 1 - 2 * (x * x + zz * zz), 2 * (y * zz - w * x)], -1),
        np.stack([2 * (x * zz - w * y), 2 * (y * zz + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], 1)  # (n, 3, 3)
    k = np.arange(3)
    local = np.stack([np.cos(2 * np.pi * k / 3), np.sin(2 * np.pi * k / 3), np.zeros(3)], -1)  # (3 verts, 3)
    r = (edge / np.sqrt(3.0))[:, None, None]
    verts = c[:, None, :].astype(np.float64) + r * np.einsum("nij,kj->nki", R, local)
    pos = verts.reshape(-1, 3).astype(F32)
    idx = np.arange(3 * n, dtype=np.uint32)
    return MeshData(pos, idx)


def soup_rays(n_coherent: int, n_incoherent: int, seed: int = 2, width: int = 4096, mini: float = 1e-4,
              maxi: float = 1e8) -> np.ndarray:
    """rdu_ray records: coherent pinhole primaries first, then incoherent rays."""
    from .raycaster import make_rays
    rays = []
    if n_coherent:
        height = max(1, n_coherent // width)
        cam = Camera()
        cam.pose([0.5, 0.5, -1.0], [0.5, 0.5, 0.5], [0.0, 1.0, 0.0])
        cam.projection(ratio=width / height, fov=1.3)
        corner, dx, dy = cam.pixelShifts()
        ii = np.arange(n_coherent)
        px = (ii % width).astype(F32) + F32(0.5)
        py = (ii // width).astype(F32) + F32(0.5)
        ndx = px / F32(width)
        ndy = py / F32(height)
        world = (corner[None, :] + ndx[:, None] * dx[None, :]) + ndy[:, None] * dy[None, :]
        d = world - cam.position[None, :]
        o = np.broadcast_to(cam.position, d.shape)
        rays.append(make_rays(o, d, mini, maxi))
    if n_incoherent:
        rng = np.random.Generator(np.random.PCG64(seed))
        o = rng.random((n_incoherent, 3), dtype=F32)
        u = rng.random((n_incoherent, 2), dtype=F32)
        cos_t = F32(2.0) * u[:, 0] - F32(1.0)
        phi = F32(2.0 * np.pi) * u[:, 1]
        sin_t = np.sqrt(F32(1.0) - cos_t * cos_t)
        d = np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], -1).astype(F32)
        rays.append(make_rays(o, d, mini, maxi))
    return np.concatenate(rays) if len(rays) > 1 else rays[0]


def icosphere(subdivisions: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> MeshData:
    """Subdivided icosahedron with shared vertices, outward normals, a tangent frame and spherical UVs."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    unit = np.array(verts)
    pos = (unit * radius + np.asarray(center, np.float64)).astype(F32)
    nrm = unit.astype(F32)
    up = np.where(np.abs(unit[:, 1:2]) > 0.999, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    tan = np.cross(up, unit)
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    bit = np.cross(unit, tan)
    uv = np.stack([0.5 + np.arctan2(unit[:, 2], unit[:, 0]) / (2 * np.pi), 0.5 - np.arcsin(unit[:, 1]) / np.pi], -1)
    return MeshData(pos, np.array(faces, np.uint32).reshape(-1), nrm, tan.astype(F32), bit.astype(F32),
                    uv.astype(F32))


def quad(p0, p1, p2, p3) -> MeshData:
    """Two triangles (p0,p1,p2), (p0,p2,p3) with a flat normal and tangent frame."""
    p = np.array([p0, p1, p2, p3], np.float64)
    n = np.cross(p[1] - p[0], p[2] - p[0])
    n /= np.linalg.norm(n)
    tng = p[1] - p[0]
    tng /= np.linalg.norm(tng)
    b = np.cross(n, tng)
    rep = lambda a: np.repeat(a[None, :], 4, 0).astype(F32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], F32)
    return MeshData(p.astype(F32), np.array([0, 1, 2, 0, 2, 3], np.uint32), rep(n), rep(tng), rep(b), uv)


def box_quads(lo, hi) -> List[MeshData]:
    """Six outward faces of an axis-aligned box."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    return [
        quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # +z
        quad((x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0)),  # -z
        quad((x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1)),  # +x
        quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),  # -x
        quad((x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0)),  # +y
        quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # -y
    ]
