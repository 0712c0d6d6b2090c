"""ORACLE — test infrastructure only (the product never imports it).

Rendu's tangent frames (Mesh::computeTangentsAndBitangents, Mesh.cpp:356-508) as a checker:
* corner tangents from the reference's own mikktspace library, compiled from
  src/libs/mikktspace/mikktspace.c into oracle/_ref/libmikkref.so by oracle/Makefile (with the
  callbacks of mikk_ref.c, as Mesh.cpp:25-91 defines them); available only where the reference
  tree is (the golden vectors under tests/golden/mikktspace carry its outputs elsewhere);
* Rendu's vertex split and bitangents (Mesh.cpp:395-505), restated below in float32.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PATH = os.path.join(HERE, "_ref", "libmikkref.so")
F32 = np.float32


def reference_available() -> bool:
    return os.path.exists(REF_PATH)


def reference_corner_tangents(positions, normals, texcoords, indices):
    """(n_indices, 4) float32: the reference library's tangent and sign per corner, and its return
    value (1 on success)."""
    lib = ctypes.CDLL(REF_PATH)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.ref_mikk_corner_tangents.restype = ctypes.c_int
    lib.ref_mikk_corner_tangents.argtypes = [fp, fp, fp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, fp]
    P = np.ascontiguousarray(positions, F32).reshape(-1)
    N = np.ascontiguousarray(normals, F32).reshape(-1)
    T = None if texcoords is None else np.ascontiguousarray(texcoords, F32).reshape(-1)
    I = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    out = np.zeros((I.size, 4), F32)
    ok = lib.ref_mikk_corner_tangents(P.ctypes.data_as(fp), N.ctypes.data_as(fp),
                                      None if T is None else T.ctypes.data_as(fp),
                                      I.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), I.size // 3,
                                      out.ctypes.data_as(fp))
    return out, int(ok)


def _cross(a, b):
    # glm::cross component order
    return np.stack([a[..., 1] * b[..., 2] - b[..., 1] * a[..., 2],
                     a[..., 2] * b[..., 0] - b[..., 2] * a[..., 0],
                     a[..., 0] * b[..., 1] - b[..., 0] * a[..., 1]], -1).astype(F32)


def _dot(a, b):
    m = (a * b).astype(F32)
    return ((m[..., 0] + m[..., 1]) + m[..., 2]).astype(F32)


def _normalize(v):
    inv = (F32(1.0) / np.sqrt(_dot(v, v))).astype(F32)
    return (v * inv[..., None]).astype(F32)


def _store(tgt, n):
    """Mesh.cpp:456-470 for one corner frame tgt (4 floats) at a vertex of normal n."""
    t = tgt[:3].astype(F32)
    c = _cross(n, t)
    if np.sqrt(_dot(c, c)) < F32(0.01):
        t = np.array([1, 0, 0] if abs(n[2]) > F32(0.01) else [0, 0, 1], F32)
    b = (F32(-tgt[3]) * _cross(n, t)).astype(F32)
    return _normalize(t), _normalize(b)


def split_vertices(corner_tangents, normals, indices):
    """Rendu's split (Mesh.cpp:395-505) of the vertices by corner frame. Returns (indices,
    source_vertex, tangents, bitangents) after the split; source_vertex[v] is the vertex that v
    copies. Pure-Python loops: for meshes of a few thousand corners."""
    ct = np.asarray(corner_tangents, F32)
    N = np.asarray(normals, F32).reshape(-1, 3)
    I = np.asarray(indices, np.uint32).reshape(-1).copy()
    nv = N.shape[0]
    per_vertex = [[] for _ in range(nv)]
    for iid, vid in enumerate(I):
        per_vertex[vid].append(iid)
    # First pass (Mesh.cpp:410-437): an instance reuses the instance just before it when their
    # frames are equal (glm::vec4 ==), else it takes the next new vertex.
    remap, offset, collisions = {}, {}, 0
    for vid in range(nv):
        inst = per_vertex[vid]
        for r in range(1, len(inst)):
            if np.all(ct[inst[r]] == ct[inst[r - 1]]):
                remap[inst[r]] = inst[r - 1]
            else:
                offset[inst[r]] = collisions
                collisions += 1
    source = list(range(nv)) + [0] * collisions
    T = np.zeros((nv + collisions, 3), F32)
    B = np.zeros((nv + collisions, 3), F32)
    # Second pass (Mesh.cpp:472-503), vertices in order, instances in order.
    for vid in range(nv):
        inst = per_vertex[vid]
        if not inst:
            T[vid], B[vid] = _store(np.array([1, 0, 0, -1], F32), N[vid])
            continue
        T[vid], B[vid] = _store(ct[inst[0]], N[vid])
        for r in range(1, len(inst)):
            iid = inst[r]
            if iid in remap:
                I[iid] = I[remap[iid]]
                continue
            fin = nv + offset[iid]
            I[iid] = fin
            source[fin] = vid
            T[fin], B[fin] = _store(ct[iid], N[vid])
    return I, np.array(source, np.uint32), T, B
