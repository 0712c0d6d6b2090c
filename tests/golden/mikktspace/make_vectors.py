"""Writes vectors.npz: inputs and the reference mikktspace library's corner tangents (oracle/_ref,
built from the reference's src/libs/mikktspace/mikktspace.c by oracle/Makefile) for the tangent
tests, so that they also run where the reference tree is absent.

Cases: Rendu's OBJ meshes under tests/golden (loaded as Mesh(stream, Indexed) does, normals from
the file), and synthetic meshes for the algorithm's branches: a UV sphere with a seam and mirrored
halves, a grid with welded duplicate vertices, degenerate triangles (repeated positions, zero UV
area), a non-manifold edge shared by three triangles, constant UVs (the forced frame without
texture coordinates), and a random shared-vertex mesh.

Run from the repository root: python tests/golden/mikktspace/make_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
sys.path.insert(0, ROOT)
from oracle import tangents as OT  # noqa: E402
from rendu_amd import ingest  # noqa: E402

F32 = np.float32
GOLD = os.path.join(ROOT, "tests", "golden")


def obj(path):
    m = ingest.load_obj(open(os.path.join(GOLD, path)).read())
    if m.normals is None:
        ingest.compute_normals(m)
    return m.positions, m.normals, m.texcoords, m.indices


def uv_sphere(nu=24, nv=12, mirror=True):
    P, N, T, I = [], [], [], []
    for j in range(nv + 1):
        th = np.pi * j / nv
        for i in range(nu + 1):  # column nu duplicates column 0 with u = 1: a UV seam
            ph = 2 * np.pi * i / nu
            n = [np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)]
            P.append(n)
            N.append(n)
            u = i / nu
            if mirror:  # the second half mirrors the first: orientation flips
                u = 2 * u if u <= 0.5 else 2 * (1 - u)
            T.append([u, j / nv])
    for j in range(nv):
        for i in range(nu):
            a, b = j * (nu + 1) + i, j * (nu + 1) + i + 1
            c, d = a + nu + 1, b + nu + 1
            I += [a, c, b, b, c, d]
    return np.array(P, F32), np.array(N, F32), np.array(T, F32), np.array(I, np.uint32)


def grid_with_duplicates(n=6, seed=3):
    rng = np.random.default_rng(seed)
    P, T, I = [], [], []
    for y in range(n):
        for x in range(n):
            P.append([x, y, 0.1 * np.sin(x + y)])
            T.append([x / (n - 1), y / (n - 1)])
    P, T = np.array(P, F32), np.array(T, F32)
    N = np.tile(np.array([[0, 0, 1]], F32), (P.shape[0], 1))
    # Every vertex once more at the end with identical attributes: the library welds them.
    P, N, T = np.concatenate([P, P]), np.concatenate([N, N]), np.concatenate([T, T])
    for y in range(n - 1):
        for x in range(n - 1):
            a, b, c, d = y * n + x, y * n + x + 1, (y + 1) * n + x, (y + 1) * n + x + 1
            q = [a, b, d, a, d, c]
            I += [v + (n * n if rng.random() < 0.5 else 0) for v in q]
    return P, N, T, np.array(I, np.uint32)


def degenerate_and_zero_area():
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0], [2, 1, 0], [0, 0, 0]], F32)
    N = np.tile(np.array([[0, 0, 1]], F32), (7, 1))
    T = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [2, 0], [2, 0], [0.5, 0.5]], F32)
    I = np.array([0, 1, 2, 0, 2, 3,      # healthy quad
                  1, 4, 5, 1, 5, 2,      # zero UV area (T[4] == T[5]): no usable derivatives
                  0, 6, 1,               # repeated position (P[6] == P[0]): degenerate
                  2, 2, 3], np.uint32)   # repeated index: degenerate
    return P, N, T, I


def non_manifold():
    P = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0], [0.5, 0, 1], [0.5, 0, -1]], F32)
    N = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 1, 0], [0, 1, 0]], F32)
    T = np.array([[0, 0], [1, 0], [0.5, 1], [0.5, -1], [0.5, 0.7], [0.3, -0.6]], F32)
    I = np.array([0, 1, 2, 1, 0, 3, 0, 1, 4, 1, 0, 5], np.uint32)  # edge (0, 1) in four triangles
    return P, N, T, I


def random_mesh(nv=300, nt=500, seed=7):
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(nv, 3)).astype(F32)
    N = rng.normal(size=(nv, 3))
    N = (N / np.linalg.norm(N, axis=1, keepdims=True)).astype(F32)
    T = rng.random((nv, 2)).astype(F32)
    I = np.concatenate([rng.choice(nv, 3, replace=False) for _ in range(nt)]).astype(np.uint32)
    return P, N, T, I


def cases():
    out = {
        "sphere_obj": obj("common/sphere.obj"),
        "vase_obj": obj("cornellbox/cornellbox_vase.obj"),
        "leaves_obj": obj("cornellbox/cornellbox_leaves.obj"),
        "large_box_obj": obj("cornellbox/cornellbox_largeBox_60.obj"),
        "uv_sphere_mirrored": uv_sphere(),
        "grid_duplicates": grid_with_duplicates(),
        "degenerate_zero_area": degenerate_and_zero_area(),
        "non_manifold": non_manifold(),
        "random": random_mesh(),
    }
    P, N, _, I = out["vase_obj"]
    out["vase_no_uv"] = (P, N, None, I)
    return out


def main():
    if not OT.reference_available():
        raise SystemExit("oracle/_ref/libmikkref.so is not built (make -C oracle, with the reference tree present)")
    arrays = {}
    for name, (P, N, T, I) in cases().items():
        ct, ok = OT.reference_corner_tangents(P, N, T, I)
        assert ok == 1, name
        arrays[f"{name}/positions"] = np.asarray(P, F32)
        arrays[f"{name}/normals"] = np.asarray(N, F32)
        if T is not None:
            arrays[f"{name}/texcoords"] = np.asarray(T, F32)
        arrays[f"{name}/indices"] = np.asarray(I, np.uint32)
        arrays[f"{name}/corner_tangents"] = ct
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "vectors.npz"), **arrays)


if __name__ == "__main__":
    main()
