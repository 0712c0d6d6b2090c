"""The allocation-free batch entry (rdu_raycaster_intersect_noalloc) against rdu_raycaster_intersect.

Both are called through the C ABI on device buffers with the same flags; every hit field must be
bit-identical. Each case makes two calls in a row on the same raycaster, the second with a smaller
batch (the shared scratch array keeps the first call's contents). Cases: incoherent batches that
are Morton-ordered in the scratch array, Rendu's visiting order with the spill stack forced into the
scratch array, and a batch too large for the array, which is rejected before anything runs.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import log_case
from rendu_amd import _lib, workloads
from rendu_amd._lib import lib
from rendu_amd.raycaster import Raycaster

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def soup():
    mesh = workloads.triangle_soup(300000, seed=21, log_edge=-4.2)
    rc = Raycaster()
    rc.addMesh(mesh)
    rc.updateHierarchy()
    return rc


def _hits(fn, rc, rays, flags):
    import torch
    n = rays.shape[0]
    d_rays = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
    d_hits = torch.full((n * 24,), 0xAB, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    st = fn(rc._h, ctypes.c_void_p(d_rays.data_ptr()), n, ctypes.c_void_p(d_hits.data_ptr()),
            flags | _lib.RDU_DEVICE_POINTERS, ctypes.c_void_p(stream.cuda_stream))
    assert st == 0, lib.rdu_last_error().decode()
    torch.cuda.synchronize()
    return d_hits.cpu().numpy()


def _compare_two_calls(rc, sizes, flags, seed):
    for n in sizes:
        rays = workloads.soup_rays(0, n, seed=seed + n)
        log_case(f"no-alloc entry vs intersect: n={n} flags={flags}")
        exp = _hits(lib.rdu_raycaster_intersect, rc, rays, flags)
        got = _hits(lib.rdu_raycaster_intersect_noalloc, rc, rays, flags)
        assert np.array_equal(got, exp), f"n={n}"
        h = exp.view(_lib.HIT_DTYPE)
        assert h["hit"].mean() > 0.2


def test_incoherent_ordered_batches(soup):
    _compare_two_calls(soup, [1 << 19, (1 << 16) + 37], _lib.RDU_INCOHERENT, seed=1)


def test_unordered_then_small(soup):
    _compare_two_calls(soup, [200000, 1000], 0, seed=2)


def test_reference_order_with_spill_stack(soup):
    os.environ["RDU_DEBUG_LDS_STACK"] = "3"  # keep 3 levels in LDS: deeper levels use the scratch array
    try:
        _compare_two_calls(soup, [50000, 777], _lib.RDU_REFERENCE_ORDER, seed=3)
    finally:
        del os.environ["RDU_DEBUG_LDS_STACK"]


def test_batch_beyond_the_scratch_array_is_rejected(soup):
    import torch
    n = 1 << 23  # ordering buffers alone: 4 x 32 MiB
    log_case("no-alloc entry rejects a 2^23-ray incoherent batch")
    d_rays = torch.zeros(n * 32, dtype=torch.uint8, device="cuda")
    d_hits = torch.zeros(n * 24, dtype=torch.uint8, device="cuda")
    st = lib.rdu_raycaster_intersect_noalloc(soup._h, ctypes.c_void_p(d_rays.data_ptr()), n,
                                             ctypes.c_void_p(d_hits.data_ptr()),
                                             _lib.RDU_INCOHERENT | _lib.RDU_DEVICE_POINTERS, None)
    assert st == _lib.RDU_ERR_UNSUPPORTED
    assert b"scratch" in lib.rdu_last_error()
    torch.cuda.synchronize()
    assert int(d_hits.count_nonzero()) == 0  # nothing ran
    st = lib.rdu_raycaster_intersect_noalloc(soup._h, ctypes.c_void_p(d_rays.data_ptr()), 16,
                                             ctypes.c_void_p(d_hits.data_ptr()), _lib.RDU_INCOHERENT, None)
    assert st == _lib.RDU_ERR_INVALID  # host pointers are refused
