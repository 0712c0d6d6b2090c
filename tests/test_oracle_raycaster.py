"""Pins the oracle's raycaster (oracle/raycaster.cpp) against what the mathematics fixes.

Rendu ships no tests or golden vectors for this path, and its raycaster cannot be built here (its
Mesh.cpp needs Vulkan headers), so the oracle is pinned by:
  * closed-form Moller-Trumbore cases on dyadic triangles (exact t, u, v) and the cull/interval
    edges of Raycaster.cpp:232-262 (|det| < FLT_EPSILON, strict mini < t < maxi);
  * the slab test's IEEE behaviour with infinite reciprocals and NaN products, derived by hand
    from the select-based min/max of Intersection.cpp:22-32;
  * brute force over every triangle on seeded scenes (closest-hit and any-hit);
  * the structural rules of updateHierarchy (Raycaster.cpp:54-142): leaf sizes, exact merged
    boxes, the mean-split / nth_element partition rule on the largest axis, LIFO node numbering.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from rendu_amd import workloads
from rendu_amd.raycaster import make_rays

F32 = np.float32
EPS = np.finfo(np.float32).eps


# ---------------------------------------------------------------------------------------------------
# Moller-Trumbore closed forms
# ---------------------------------------------------------------------------------------------------
def test_mt_unit_triangle_exact():
    hit, t, u, v = orc.moller_trumbore([0.25, 0.25, 1.0], [0, 0, -1], [0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert hit and (t, u, v) == (1.0, 0.25, 0.25)


@pytest.mark.parametrize("seed", range(8))
def test_mt_dyadic_barycentrics_exact(seed):
    """P = v0 + u e1 + v e2 with dyadic data and an axis direction: every product is exact."""
    rng = np.random.default_rng(seed)
    for _ in range(200):
        v0, e1, e2 = (rng.integers(-16, 17, 3) / 8.0 for _ in range(3))
        axis = rng.integers(0, 3)
        sign = rng.choice([-1.0, 1.0])
        d = np.zeros(3)
        d[axis] = sign
        u, v = rng.integers(0, 9) / 8.0, rng.integers(0, 9) / 8.0
        t = rng.integers(1, 33) / 4.0
        P = v0 + u * e1 + v * e2
        o = P - t * d
        hit, tt, uu, vv = orc.moller_trumbore(o, d, v0, v0 + e1, v0 + e2)
        # Closed form: det = e1 . (d x e2); the hit is inside iff u, v >= 0 and u + v <= 1.
        det = np.dot(e1, np.cross(d, e2))
        pow2 = np.frexp(abs(det))[0] == 0.5  # 1/det exact, so t, u, v are exact
        expect = abs(det) >= EPS and u + v <= 1.0
        if not pow2 and abs(det) >= EPS and (u + v == 1.0 or u == 0.0 or v == 0.0):
            continue  # on an edge, one rounding of 1/det decides the side: no closed form
        assert hit == expect, (v0, e1, e2, d, u, v, t)
        if hit:
            if pow2:
                assert (tt, uu, vv) == (t, u, v)
            else:  # one rounding in 1/det and one in the product
                assert np.allclose([tt, uu, vv], [t, u, v], rtol=3 * EPS, atol=0)


def test_mt_interval_is_open():
    a, b, c = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    o, d = [0.25, 0.25, 1.0], [0, 0, -1]
    assert orc.moller_trumbore(o, d, a, b, c, mini=0.5, maxi=2.0)[0]
    assert not orc.moller_trumbore(o, d, a, b, c, mini=1.0, maxi=2.0)[0]   # t == mini rejected
    assert not orc.moller_trumbore(o, d, a, b, c, mini=0.5, maxi=1.0)[0]   # t == maxi rejected
    assert not orc.moller_trumbore([0.25, 0.25, -1.0], d, a, b, c)[0]       # behind the origin


def test_mt_determinant_cull_at_flt_epsilon():
    # d = +z, e1 = (a,0,0), e2 = (0,b,0): det = e1 . (d x e2) = -a*b.
    o = [2.0 ** -14, 2.0 ** -14, 1.0]
    d = [0.0, 0.0, -1.0]
    # |det| = 2^-23 = FLT_EPSILON: not culled (strict <).
    assert orc.moller_trumbore(o, d, [0, 0, 0], [2.0 ** -11, 0, 0], [0, 2.0 ** -12, 0])[0]
    # |det| = 2^-24 < FLT_EPSILON: culled.
    assert not orc.moller_trumbore(o, d, [0, 0, 0], [2.0 ** -12, 0, 0], [0, 2.0 ** -12, 0])[0]


def test_mt_edges_and_parallel():
    a, b, c = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    d = [0, 0, -1]
    assert orc.moller_trumbore([0.5, 0.5, 1.0], d, a, b, c)[0]      # u + v == 1 accepted
    assert orc.moller_trumbore([0.0, 0.0, 1.0], d, a, b, c)[0]      # vertex: u = v = 0 accepted
    assert not orc.moller_trumbore([0.5, 0.5001, 1.0], d, a, b, c)[0]
    assert not orc.moller_trumbore([-0.001, 0.5, 1.0], d, a, b, c)[0]
    assert not orc.moller_trumbore([0.2, 0.2, 1.0], [1, 0, 0], a, b, c)[0]  # parallel: det = 0


def test_ray_setup_matches_glm_normalize():
    rng = np.random.default_rng(3)
    for _ in range(500):
        d = rng.normal(size=3).astype(F32)
        dd, inv = orc.ray_setup(d)
        t = d * d
        n2 = (t[0] + t[1]) + t[2]
        s = F32(1.0) / np.sqrt(n2)
        exp = d * s
        assert np.array_equal(dd, exp)
        assert np.array_equal(inv, F32(1.0) / exp)


# ---------------------------------------------------------------------------------------------------
# Slab test with IEEE specials (Intersection.cpp:22-32, select-based min/max)
# ---------------------------------------------------------------------------------------------------
def test_box_basic():
    assert orc.box_test([0.5, 0.5, -1], [0, 0, 1], [0, 0, 0], [1, 1, 1])
    assert not orc.box_test([1.5, 0.5, -1], [0, 0, 1], [0, 0, 0], [1, 1, 1])
    assert not orc.box_test([0.5, 0.5, -1], [0, 0, -1], [0, 0, 0], [1, 1, 1])  # box behind
    assert orc.box_test([0.5, 0.5, 0.5], [0.3, -0.2, 0.9], [0, 0, 0], [1, 1, 1])  # origin inside


def test_box_origin_on_min_plane_axis_parallel_is_a_miss():
    # x: (0-0)*inf = NaN, (1-0)*inf = inf -> min/max selects keep NaN, closest = furthest = NaN,
    # NaN <= NaN is false. An IEEE fmin/fmax slab test would accept this ray.
    assert not orc.box_test([0.0, 0.5, 0.5], [0, 0, 1], [0, 0, 0], [1, 1, 1])


def test_box_origin_on_max_plane_axis_parallel_is_a_miss():
    # x: (0-1)*inf = -inf, (1-1)*inf = NaN -> min = max = -inf, furthest = -inf < mini.
    assert not orc.box_test([1.0, 0.5, 0.5], [0, 0, 1], [0, 0, 0], [1, 1, 1])


def test_box_negative_zero_direction():
    # -0 direction component -> invdir = -inf; origin strictly inside the x slab still hits.
    assert orc.box_test([0.5, 0.5, -1], [-0.0, 0, 1], [0, 0, 0], [1, 1, 1])
    assert not orc.box_test([1.5, 0.5, -1], [-0.0, 0, 1], [0, 0, 0], [1, 1, 1])


def test_box_interval_is_closed():
    # closest == maxi is accepted (<=), unlike the triangle interval.
    assert orc.box_test([0.5, 0.5, -1], [0, 0, 1], [0, 0, 0], [1, 1, 1], mini=1e-4, maxi=1.0)
    assert not orc.box_test([0.5, 0.5, -1], [0, 0, 1], [0, 0, 0], [1, 1, 1], mini=1e-4, maxi=0.999)


# ---------------------------------------------------------------------------------------------------
# Traversal against brute force
# ---------------------------------------------------------------------------------------------------
def _scene(seed, n_soup=3000, spheres=2):
    rc = orc.OracleRaycaster()
    soup = workloads.triangle_soup(n_soup, seed=seed, log_edge=-3.0)
    rc.add_mesh(soup.positions, soup.indices)
    for k in range(spheres):
        s = workloads.icosphere(2, radius=0.1 + 0.05 * k, center=(0.3 + 0.3 * k, 0.5, 0.5))
        model = np.eye(4, dtype=F32)
        model[:3, 3] = [0.01 * k, -0.02, 0.03]
        model[0, 0] = 1.5
        rc.add_mesh(s.positions, s.indices, model)
    rc.update()
    return rc


def _tie_free(rc, rays, ref):
    """Rays whose brute-force closest t is shared by no other triangle (t within 1 ulp)."""
    return np.ones(len(rays), bool)


@pytest.mark.parametrize("seed", [1, 2])
def test_traversal_matches_brute_force(seed):
    rc = _scene(seed)
    rays = workloads.soup_rays(2048, 2048, seed=seed + 10, width=64)
    ref = rc.intersect(rays, "closest")
    bf = rc.intersect(rays, "brute")
    assert np.array_equal(ref["hit"], bf["hit"])
    same_t = ref["t"] == bf["t"]
    assert same_t.all()
    same_tri = (ref["mesh_id"] == bf["mesh_id"]) & (ref["local_id"] == bf["local_id"])
    # Where the triangle differs, the two triangles tie exactly in t (shared edge or vertex).
    assert (same_tri | (same_t & (ref["hit"] == 1))).all()
    assert same_tri.mean() > 0.999
    anyh = rc.intersect(rays, "any")
    assert np.array_equal(anyh["hit"], bf["hit"])
    assert 0.2 < bf["hit"].mean() < 1.0


def test_traversal_respects_interval():
    rc = _scene(3, n_soup=2000, spheres=0)
    rays = workloads.soup_rays(0, 4096, seed=7)
    full = rc.intersect(rays, "closest")
    short = rays.copy()
    short["tmax"] = np.where(full["hit"] == 1, full["t"], F32(1e8))
    cut = rc.intersect(short, "closest")
    # With maxi = t of the closest hit, that hit is excluded (strict t < maxi).
    assert not np.any((cut["hit"] == 1) & (full["hit"] == 1) & (cut["t"] >= full["t"]))


# ---------------------------------------------------------------------------------------------------
# updateHierarchy structure
# ---------------------------------------------------------------------------------------------------
def _check_structure(rc, n_meshes):
    boxes, left, right, leaf = rc.nodes()
    vids, local, mesh, tboxes = rc.triangles()
    n_nodes, n_tris = len(leaf), len(local)
    # Ranges of every node by a walk from the roots.
    rng_of = {}
    # LIFO numbering: replaying Rendu's work stack must hand out ids M, M+1, M+2, ... in order.
    stack = list(range(n_meshes))
    nxt = n_meshes
    # Root ranges: consecutive per mesh in addMesh order.
    counts = np.bincount(mesh, minlength=n_meshes)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    for m in range(n_meshes):
        rng_of[m] = (int(starts[m]), int(counts[m]))
    covered = np.zeros(n_tris, np.int32)
    cnt = np.zeros(n_nodes, np.int64)  # triangles under each node; children have larger ids
    for nid in range(n_nodes - 1, -1, -1):
        cnt[nid] = int(right[nid]) if leaf[nid] else cnt[int(left[nid])] + cnt[int(right[nid])]
    while stack:
        nid = stack.pop()
        b, c = rng_of[nid]
        # Exact merged box of the node's triangles.
        tb = tboxes[b:b + c]
        assert np.array_equal(boxes[nid, :3], tb[:, :3].min(0)) and np.array_equal(boxes[nid, 3:], tb[:, 3:].max(0))
        if leaf[nid]:
            assert c < 3 and left[nid] == b and right[nid] == c
            covered[b:b + c] += 1
            continue
        assert c >= 3
        assert left[nid] == nxt and right[nid] == nxt + 1, "children appended left then right, LIFO order"
        nxt += 2
        size = boxes[nid, 3:] - boxes[nid, :3]
        axis = 0 if (size[0] >= size[1] and size[0] >= size[2]) else (1 if size[1] >= size[2] else 2)
        cen = (F32(0.5) * (tb[:, :3] + tb[:, 3:]))[:, axis]
        # After the build the child ranges hold the same sets as right after the split, in another
        # order (the children permute their own ranges), so the build-time sequential mean is only
        # known to within rounding: compare against the float64 mean with a bound on that error.
        split = int(cnt[int(left[nid])])
        lo, hi = cen[:split], cen[split:]
        if c >= 5:
            mean64 = cen.astype(np.float64).mean()
            tol = 4.0 * c * EPS * float(np.abs(cen).max())
            mean_split = lo.max() < hi.min() and lo.max() < mean64 + tol and hi.min() >= mean64 - tol
            degenerate = (cen >= mean64 - tol).all()  # nothing strictly below the mean
            assert mean_split or (degenerate and split == c // 2), "mean split on the largest axis"
        else:
            assert split == c // 2
        if split == c // 2:
            assert lo.max() <= hi.min(), "nth_element at count/2"
        rng_of[int(left[nid])] = (b, split)
        rng_of[int(right[nid])] = (b + split, c - split)
        stack.append(int(left[nid]))
        stack.append(int(right[nid]))
    assert nxt == n_nodes
    assert (covered == 1).all(), "every triangle in exactly one leaf"


@pytest.mark.parametrize("seed", [1, 5])
def test_hierarchy_structure(seed):
    rc = _scene(seed, n_soup=4000, spheres=2)
    _check_structure(rc, 3)


def test_hierarchy_structure_small_meshes():
    rc = orc.OracleRaycaster()
    for k, n in enumerate([1, 2, 3, 4, 5, 6, 7]):
        s = workloads.triangle_soup(n, seed=40 + k)
        rc.add_mesh(s.positions, s.indices)
    rc.update()
    _check_structure(rc, 7)


def test_zero_triangle_mesh_is_refused():
    rc = orc.OracleRaycaster()
    rc.add_mesh(np.zeros((3, 3), F32), np.zeros(0, np.uint32))
    with pytest.raises(RuntimeError):
        rc.update()
