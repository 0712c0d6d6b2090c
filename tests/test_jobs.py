"""Batched render jobs (rendu_amd/jobs.py): job generator on the CPU; concurrent runs on the GPU
give each job exactly the image it gets alone, and one small job matches the oracle."""
import numpy as np
import pytest

from rendu_amd import scenes


def test_random_job_distribution():
    tris, sides, spps = [], [], []
    for s in range(64):
        scn, cam, w, h, spp, depth = scenes.random_job(s)
        t = sum(scn.meshes[o.mesh].triangle_count for o in scn.objects)
        tris.append(t)
        sides.append(w)
        spps.append(spp)
        assert depth == 6 and w == h
    assert 1e4 * 0.85 <= min(tris) and max(tris) <= 5e5 * 1.01
    assert 128 <= min(sides) and max(sides) <= 1024
    assert 4 <= min(spps) and max(spps) <= 256
    # Log-uniform: the median triangle count sits near the geometric mean of the range.
    assert 2e4 < np.median(tris) < 2e5
    a = scenes.random_job(5)
    b = scenes.random_job(5)
    assert a[2:] == b[2:] and len(a[0].objects) == len(b[0].objects)


def test_estimated_cost_orders_by_work():
    from rendu_amd import jobs as J
    scn, cam, w, h, spp, depth = scenes.random_job(3)
    small = J.Job(scn, cam(64, 64), 64, 64, 4, depth)
    wide = J.Job(scn, cam(512, 512), 512, 512, 4, depth)
    deep = J.Job(scn, cam(64, 64), 64, 64, 256, depth)
    gpu = J.Job(scn, cam(64, 64), 64, 64, 4, depth, gpu_build=True)
    assert J.estimated_cost(small) < J.estimated_cost(wide)
    assert J.estimated_cost(small) < J.estimated_cost(deep)
    assert J.estimated_cost(gpu) < J.estimated_cost(small)  # the GPU build is the cheaper build
    # samples round to a power of two as the render does
    odd = J.Job(scn, cam(64, 64), 64, 64, 5, depth)
    assert J.estimated_cost(odd) == J.estimated_cost(small)
    with pytest.raises(ValueError):
        J.run_batch([], order="shortest")
    with pytest.raises(ValueError):
        J.run_batch([], build_workers=-1)


def test_merge_groups_partition():
    from rendu_amd import jobs as J
    js = []
    for s in range(12):
        scn, cam, w, h, spp, depth = scenes.random_job(s)
        js.append(J.Job(scn, cam(w, h), w, h, spp, depth, seed=s))
    cost = [J.estimated_cost(j) for j in js]
    order = sorted(range(12), key=lambda i: -cost[i])
    for groups in (1, 2, 3, 5, 12, 20):
        parts = J.merge_groups(js, groups)
        assert 1 <= len(parts) <= min(groups, 12) and all(parts)
        # consecutive slices of the longest-first order, covering every job once
        assert [i for g in parts for i in g] == order
        if groups == 1:
            assert parts == [order]
        # a job's group is the share of the total cost its start falls in (the last share takes
        # the rest): one share per group, increasing along the order
        tot, acc, shares = sum(cost), 0.0, []
        for g in parts:
            sh = set()
            for i in g:
                sh.add(min(groups - 1, int(acc / tot * groups)))
                acc += cost[i]
            assert len(sh) == 1
            shares.append(sh.pop())
        assert shares == sorted(set(shares))
    assert J.merge_groups([], 3) == []
    with pytest.raises(ValueError):
        J.merge_groups(js, 0)


@pytest.mark.gpu
def test_longest_first_batch_matches_sequential():
    from conftest import log_case
    from rendu_amd import jobs as J
    log_case("5 small jobs, longest first, 4 workers vs alone")
    js = []
    for s in range(5):
        scn, cam, w, h, spp, depth = scenes.random_job(200 + s)
        side = 32 + 16 * s
        js.append(J.Job(scn, cam(side, side), side, side, min(spp, 8), depth, seed=s))
    alone = [J.run_job(j).image for j in js]
    res, wall = J.run_batch(js, workers=4, order="lpt")
    assert [r.index for r in res] == list(range(5))
    for r, a in zip(res, alone):
        np.testing.assert_array_equal(r.image, a)


@pytest.mark.gpu
def test_pipelined_builds_match_sequential():
    from conftest import log_case
    from rendu_amd import jobs as J
    log_case("5 small jobs, builds in their own pool of 2 ahead of 4 render threads, vs alone")
    js = []
    for s in range(5):
        scn, cam, w, h, spp, depth = scenes.random_job(300 + s)
        side = 40 + 8 * s
        js.append(J.Job(scn, cam(side, side), side, side, min(spp, 8), depth, seed=s))
    alone = [J.run_job(j).image for j in js]
    res, wall = J.run_batch(js, workers=4, order="lpt", build_workers=2)
    assert [r.index for r in res] == list(range(5))
    for r, a in zip(res, alone):
        np.testing.assert_array_equal(r.image, a)
        assert r.latency_s >= r.build_s + r.render_s - 1e-6


@pytest.mark.gpu
def test_concurrent_jobs_match_sequential():
    from conftest import log_case
    from rendu_amd import jobs as J
    log_case("6 small jobs, 3 workers vs alone")
    js = []
    for s in range(6):
        scn, cam, w, h, spp, depth = scenes.random_job(100 + s)
        side = 48 + 8 * s
        js.append(J.Job(scn, cam(side, side), side, side, min(spp, 16), depth, seed=s))
    alone = [J.run_job(j).image for j in js]
    res, wall = J.run_batch(js, workers=3)
    for r, a in zip(res, alone):
        np.testing.assert_array_equal(r.image, a)
    summ = J.summary(res, wall)
    assert summ["jobs"] == 6 and summ["jobs_per_s"] > 0


@pytest.mark.gpu
def test_small_job_matches_oracle():
    from conftest import log_case
    from oracle import oracle as orc
    from rendu_amd import jobs as J
    from rendu_amd import scene as S
    log_case("job 7 at 40x40 8spp vs oracle")
    scn, cam, w, h, spp, depth = scenes.random_job(7)
    c = cam(40, 40)
    got = J.run_job(J.Job(scn, c, 40, 40, 8, depth, seed=3)).image
    exp = np.zeros((40, 40, 3), np.float32)
    orc.OraclePathTracer(scn).render(S.camera_desc(c), 8, depth, exp, S.RenderOpts(3, 0, 3, 0))
    d = np.abs(got.astype(np.float64) - exp)
    assert (d <= 1e-4 * np.abs(exp) + 1e-6).mean() >= 0.98
    assert d.mean() <= 1e-3 * max(np.abs(exp).mean(), 1e-6)


@pytest.mark.gpu
def test_render_working_sets_move_between_scenes():
    """A render's working memory comes from a pool shared by all scenes: a scene rendered after
    others (smaller and larger images, a deeper or shallower tree, graph replay on) gets the same
    image as the first time."""
    from conftest import log_case
    from rendu_amd.pathtracer import PathTracer
    log_case("two scenes alternating renders over shared working sets, one with graph replay")
    scn_a, cam_a, _, _, _, depth_a = scenes.random_job(11)
    scn_b, cam_b = scenes.cornell_spheres(seed=3, subdivisions=2)
    a, b = PathTracer(scn_a), PathTracer(scn_b)
    b.set_graph(True)

    def render(pt, cam, w, h, spp, depth, seed):
        img = np.zeros((h, w, 3), np.float32)
        pt.render(cam(w, h), spp, depth, img, seed=seed, max_paths=w * h * 2)
        return img

    a1 = render(a, cam_a, 96, 96, 8, depth_a, 1)
    b1 = render(b, cam_b, 40, 40, 8, 5, 2)
    a2 = render(a, cam_a, 96, 96, 8, depth_a, 1)
    b2 = render(b, cam_b, 40, 40, 8, 5, 2)
    b3 = render(b, cam_b, 40, 40, 8, 5, 2)
    a3 = render(a, cam_a, 160, 160, 4, depth_a, 1)  # a larger pass: the shared set grows
    b4 = render(b, cam_b, 40, 40, 8, 5, 2)
    np.testing.assert_array_equal(a1, a2)
    for img in (b2, b3, b4):
        np.testing.assert_array_equal(img, b1)
    assert a3.mean() > 0
    assert b.last_stats()["graph_passes"] > 0


@pytest.mark.gpu
def test_memory_is_reused_across_scenes():
    """Scenes built, rendered and dropped one after another reuse the library's device memory
    (render working sets and buffers come from pools): after the first cycles the device's free
    memory stops falling."""
    import torch
    from conftest import log_case
    from rendu_amd.pathtracer import PathTracer
    log_case("12 create / render / destroy cycles, device free memory stable")
    scn, cam, w, h, spp, depth = scenes.random_job(21)
    free = []
    for k in range(12):
        pt = PathTracer(scn)
        img = np.zeros((160, 160, 3), np.float32)
        pt.render(cam(160, 160), 8, depth, img, seed=k)
        del pt
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
    # Later cycles allocate nothing new from the device: free memory moves by less than 64 MiB.
    assert max(free[3:]) - min(free[3:]) < 64 << 20, [f >> 20 for f in free]
