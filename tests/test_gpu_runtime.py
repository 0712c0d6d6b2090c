"""Host runtime of the library on the GPU: failure paths and stream ordering of shared memory.

- A render working set whose growth failed half-way is grown again in full by the next render that
  leases it (it is not reused with freed buffers). Run in a child process, where that set is the
  only one, so the next render is sure to lease it.
- A scene whose commit failed can be committed again (its raycaster is dropped on failure).
- Asynchronous batches on two streams share one raycaster's scratch memory in stream order: both
  give the results of the same batches run one after the other.
Allocation failures are forced from the host with RDU_FAIL_ALLOC=1 (no kernel runs on them).
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, log_case
from rendu_amd import _lib
from rendu_amd import scene as S
from rendu_amd import scenes, workloads
from rendu_amd._lib import check, lib
from rendu_amd.pathtracer import PathTracer
from rendu_amd.raycaster import Raycaster

pytestmark = pytest.mark.gpu

CHILD = r"""
import os, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from rendu_amd import scenes
from rendu_amd.pathtracer import PathTracer
scn, cam_of = scenes.cornell_spheres(seed=3, subdivisions=2)
pt = PathTracer(scn)
a = np.zeros((32, 32, 3), np.float32)
pt.render(cam_of(32, 32), 4, 3, a, seed=1)
print("small render ok", flush=True)
os.environ["RDU_FAIL_ALLOC"] = "1"
big = np.zeros((256, 256, 3), np.float32)
try:
    pt.render(cam_of(256, 256), 16, 3, big, seed=1)
    print("grow unexpectedly succeeded", flush=True)
    sys.exit(3)
except Exception as e:
    print("grow failed as forced:", str(e)[:80], flush=True)
del os.environ["RDU_FAIL_ALLOC"]
b = np.zeros((32, 32, 3), np.float32)
pt.render(cam_of(32, 32), 4, 3, b, seed=1)
print("after failure", "equal" if np.array_equal(a, b) else "DIFFERENT", flush=True)
sys.exit(0 if np.array_equal(a, b) else 4)
"""


def test_failed_growth_is_regrown_by_the_next_render():
    log_case("child process: small render, forced failed growth, small render again")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "after failure equal" in r.stdout


def _scene_handle(scn):
    h = ctypes.c_void_p()
    check(lib.rdu_scene_create(ctypes.byref(h)))
    keep, tex, meshes, mats, objs, lights, bg = scn.descriptors()
    i = ctypes.c_int32()
    for t in tex:
        check(lib.rdu_scene_add_texture(h, ctypes.byref(t), ctypes.byref(i)))
    for m in meshes:
        check(lib.rdu_scene_add_mesh(h, ctypes.byref(m), ctypes.byref(i)))
    for m in mats:
        check(lib.rdu_scene_add_material(h, ctypes.byref(m), ctypes.byref(i)))
    for o in objs:
        check(lib.rdu_scene_add_object(h, ctypes.byref(o)))
    check(lib.rdu_scene_set_background(h, ctypes.byref(bg)))
    return h, keep


def test_commit_can_be_retried_after_a_failure():
    scn, cam_of = scenes.cornell_spheres(seed=5, subdivisions=2)
    h, keep = _scene_handle(scn)
    log_case("commit with forced allocation failure, then again")
    os.environ["RDU_FAIL_ALLOC"] = "1"
    try:
        assert lib.rdu_scene_commit(h) == _lib.RDU_ERR_ALLOC
    finally:
        del os.environ["RDU_FAIL_ALLOC"]
    assert not lib.rdu_scene_raycaster(h)
    check(lib.rdu_scene_commit(h))
    cam = S.camera_desc(cam_of(24, 24))
    got = np.zeros((24, 24, 3), np.float32)
    opts = S.RenderOpts(2, 0, 3, 0)
    check(lib.rdu_pathtracer_render(h, ctypes.byref(cam), 4, 3, got.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    24, 24, 3, ctypes.byref(opts)))
    lib.rdu_scene_destroy(h)
    exp = np.zeros((24, 24, 3), np.float32)
    PathTracer(scn).render(cam_of(24, 24), 4, 3, exp, seed=2)
    assert np.array_equal(got, exp)
    del keep


def test_async_batches_on_two_streams_share_scratch_in_order():
    import torch
    soup = workloads.triangle_soup(200000, seed=3, log_edge=-4.0)
    rc = Raycaster()
    rc.addMesh(soup)
    rc.updateHierarchy()
    n = 1 << 17  # above the size where incoherent batches are sorted (scratch: keys, permutation)
    ra = workloads.soup_rays(0, n, seed=4)
    rb = workloads.soup_rays(0, n, seed=5)
    ea = rc.intersect_batch(ra, incoherent=True)
    eb = rc.intersect_batch(rb, incoherent=True)
    log_case("two async incoherent batches on two torch streams")
    da = torch.from_numpy(np.ascontiguousarray(ra).view(np.uint8)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(rb).view(np.uint8)).cuda()
    ha = torch.empty(n * 24, dtype=torch.uint8, device="cuda")
    hb = torch.empty(n * 24, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        rc.intersect_device(da.data_ptr(), n, ha.data_ptr(), sa.cuda_stream, incoherent=True)
        rc.intersect_device(db.data_ptr(), n, hb.data_ptr(), sb.cuda_stream, incoherent=True)
    torch.cuda.synchronize()
    ga = ha.cpu().numpy().view(_lib.HIT_DTYPE)
    gb = hb.cpu().numpy().view(_lib.HIT_DTYPE)
    for got, exp in ((ga, ea), (gb, eb)):
        for k in ("t", "u", "v", "hit", "mesh_id", "local_id"):
            assert np.array_equal(got[k], exp[k]), k
