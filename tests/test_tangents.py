"""Tangent frames (rdu_mesh_corner_tangents / rdu_mesh_tangent_frames, host code of the library)
against Rendu's: Mesh::computeTangentsAndBitangents (Mesh.cpp:356-508) runs the mikktspace library
and then splits vertices by corner frame.

* Corner tangents equal the reference library's bit for bit: the library itself, compiled from the
  reference tree into oracle/_ref where that tree is present, and its outputs committed as golden
  vectors (tests/golden/mikktspace, written by make_vectors.py) everywhere. Cases: Rendu's sphere,
  vase, leaves and box meshes; a mirrored UV sphere with a seam; welded duplicate vertices;
  degenerate triangles and zero UV area; a non-manifold edge; constant UVs; a random mesh.
* The split equals oracle/tangents.py's restatement of Mesh.cpp:395-505 applied to the reference
  corner tangents (indices, copied vertices, tangents and bitangents bit for bit); the restatement
  is pinned on hand-made cases (a corner compares with the corner just before it only).
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import tangents as OT
from rendu_amd import _lib, ingest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mikktspace", "vectors.npz")
F32 = np.float32


def _cases():
    z = np.load(GOLD)
    names = sorted({k.split("/")[0] for k in z.files})
    out = {}
    for n in names:
        out[n] = (z[f"{n}/positions"], z[f"{n}/normals"], z[f"{n}/texcoords"] if f"{n}/texcoords" in z.files else None,
                  z[f"{n}/indices"], z[f"{n}/corner_tangents"])
    return out


CASES = _cases()


def corner_tangents(P, N, T, I):
    out = np.zeros((I.size, 4), F32)
    P, N = np.ascontiguousarray(P, F32), np.ascontiguousarray(N, F32)
    T = None if T is None else np.ascontiguousarray(T, F32)
    I = np.ascontiguousarray(I, np.uint32)
    _lib.check(_lib.lib.rdu_mesh_corner_tangents(P.ctypes.data, N.ctypes.data, None if T is None else T.ctypes.data,
                                                 P.shape[0], I.ctypes.data, I.size, out.ctypes.data))
    return out


def frames(P, N, T, I, capacity=None):
    P, N = np.ascontiguousarray(P, F32), np.ascontiguousarray(N, F32)
    T = None if T is None else np.ascontiguousarray(T, F32)
    I = np.ascontiguousarray(I, np.uint32).copy()
    cap = capacity if capacity is not None else P.shape[0] + I.size
    src = np.zeros(max(cap, 1), np.uint32)
    tan = np.zeros((max(cap, 1), 3), F32)
    bit = np.zeros((max(cap, 1), 3), F32)
    n = ctypes.c_size_t(0)
    st = _lib.lib.rdu_mesh_tangent_frames(P.ctypes.data, N.ctypes.data, None if T is None else T.ctypes.data,
                                          P.shape[0], I.ctypes.data, I.size, src.ctypes.data, tan.ctypes.data,
                                          bit.ctypes.data, cap, ctypes.byref(n))
    return st, I, src[:n.value], tan[:n.value], bit[:n.value]


@pytest.mark.parametrize("name", sorted(CASES))
def test_corner_tangents_equal_the_reference_library(name):
    P, N, T, I, expected = CASES[name]
    got = corner_tangents(P, N, T, I)
    assert np.array_equal(got.view(np.uint32), expected.view(np.uint32)), (
        name, int(np.sum(np.any(got != expected, axis=1))))
    if OT.reference_available():
        live, ok = OT.reference_corner_tangents(P, N, T, I)
        assert ok == 1 and np.array_equal(live.view(np.uint32), expected.view(np.uint32))


@pytest.mark.parametrize("name", sorted(CASES))
def test_vertex_split_equals_rendus(name):
    P, N, T, I, expected = CASES[name]
    st, idx, src, tan, bit = frames(P, N, T, I)
    assert st == 0
    e_idx, e_src, e_tan, e_bit = OT.split_vertices(expected, N, I)
    assert np.array_equal(idx, e_idx) and np.array_equal(src, e_src)
    assert np.array_equal(tan.view(np.uint32), e_tan.view(np.uint32))
    assert np.array_equal(bit.view(np.uint32), e_bit.view(np.uint32))
    # Geometry is unchanged: every corner still sees its original position.
    assert np.array_equal(np.asarray(P)[src[idx]], np.asarray(P)[I])
    assert np.allclose(np.linalg.norm(tan, axis=1), 1, atol=1e-6)


def test_cases_exercise_the_branches():
    """The vectors hold mirrored (sign -1 and +1), split vertices, and untouched default frames."""
    _, _, _, _, ct = CASES["uv_sphere_mirrored"]
    assert set(np.unique(ct[:, 3])) == {-1.0, 1.0}
    P, N, T, I, ct = CASES["sphere_obj"]
    _, _, src, _, _ = frames(P, N, T, I)
    assert src.size > P.shape[0]  # seam vertices split
    _, _, _, _, ct = CASES["vase_no_uv"]
    assert np.all(ct == np.array([1, 0, 0, -1], F32))  # no usable derivatives: the library's default frame


def test_split_restatement_on_hand_cases():
    """Mesh.cpp:410-437: an instance compares with the instance just before it only."""
    A, B = np.array([1, 0, 0, 1], F32), np.array([0, 1, 0, 1], F32)
    N = np.array([[0, 0, 1]] * 2, F32)
    # Vertex 0 used by corners 0, 1, 2 with frames A, B, A: two new vertices (the third A does not
    # find the first one); vertex 1 used by corners 3, 4 with frames A, A: none.
    I = np.array([0, 0, 0, 1, 1], np.uint32)
    idx, src, tan, bit = OT.split_vertices(np.stack([A, B, A, A, A]), N, I)
    assert list(idx) == [0, 2, 3, 1, 1] and list(src) == [0, 1, 0, 0]
    assert np.array_equal(tan[2], [0, 1, 0]) and np.array_equal(tan[3], [1, 0, 0])
    # bitangent = -sign * cross(n, t): n = +z, t = +x, sign +1 -> -y.
    assert np.array_equal(bit[0], [0, -1, 0])
    # Frames A, A, B: the second reuses vertex 0, the third is new.
    idx, src, _, _ = OT.split_vertices(np.stack([A, A, B, A, A]), N, I)
    assert list(idx) == [0, 0, 2, 1, 1] and list(src) == [0, 1, 0]
    # A tangent parallel to the normal is replaced: +x where |n.z| > 0.01.
    _, _, tan, bit = OT.split_vertices(np.array([[0, 0, 1, -1]] * 5, F32), N, I)
    assert np.array_equal(tan[0], [1, 0, 0]) and np.array_equal(bit[0], [0, 1, 0])


def test_single_triangle_frame_by_hand():
    """One triangle with the OBJ loader's flipped v: positions (0,0,0), (1,0,0), (0,1,0), UVs
    (0,1), (1,1), (0,0) (v' = 1 - v). Texture area 2A = 1 * (-1) - 0 * 0 < 0: not orientation
    preserving, sign -1; Os = t31y d1 - t21y d2 = -(1,0,0), scaled by sign / |Os| gives +x. Rendu's
    bitangent -sign * cross(n, t) = cross(+z, +x) = +y."""
    m = ingest.load_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n")
    ingest.compute_normals(m)
    ct = corner_tangents(m.positions, m.normals, m.texcoords, m.indices)
    assert np.array_equal(ct, np.array([[1, 0, 0, -1]] * 3, F32))
    ingest.compute_tangents(m)
    assert np.array_equal(m.tangents, [[1, 0, 0]] * 3) and np.array_equal(m.bitangents, [[0, 1, 0]] * 3)


def test_edge_cases_and_errors():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F32)
    N = np.array([[0, 0, 1]] * 3, F32)
    # No triangle: fixed axes (Mesh.cpp:388-392).
    st, idx, src, tan, bit = frames(P, N, None, np.zeros(0, np.uint32))
    assert st == 0 and list(src) == [0, 1, 2]
    assert np.array_equal(tan, [[1, 0, 0]] * 3) and np.array_equal(bit, [[0, 1, 0]] * 3)
    assert frames(P, N, None, np.array([0, 1, 3], np.uint32))[0] == _lib.RDU_ERR_INVALID
    assert frames(P, N, None, np.array([0, 1], np.uint32))[0] == _lib.RDU_ERR_INVALID
    assert frames(P, N, None, np.array([0, 1, 2], np.uint32), capacity=5)[0] == _lib.RDU_ERR_INVALID
    # A vertex no triangle uses keeps the library's default frame.
    P4, N4 = np.concatenate([P, [[5, 5, 5]]]).astype(F32), np.concatenate([N, [[0, 1, 0]]]).astype(F32)
    st, _, src, tan, bit = frames(P4, N4, None, np.array([0, 1, 2], np.uint32))
    assert st == 0 and np.array_equal(tan[3], [1, 0, 0]) and np.array_equal(bit[3], [0, 0, -1])
