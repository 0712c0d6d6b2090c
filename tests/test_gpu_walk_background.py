"""GPU path tracer against the oracle on the shadow walk's and the backgrounds' reference rules.

- Rendu's shadow walk (PathTracer.cpp:124-157) crosses any number of alpha-0 layers: with 20 layers
  (more than one round of walk steps) every floor point still sees the light, and the GPU makes the
  same layers + 1 visibility queries per light sample as the oracle.
- COLOR and IMAGE backgrounds light direct hits only, cube maps and lat-long maps light the floor
  too (PathTracer.cpp:23-55).
Bars as in test_gpu_pathtracer.py.
"""
import numpy as np
import pytest

from conftest import log_case
from rendu_amd import scene as S
from rendu_amd import scenes
from test_gpu_pathtracer import assert_close, render_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layers", [3, 20])
def test_shadow_walk_through_many_layers(layers):
    scn, cam_of = scenes.masked_layers(layers)
    w, h, spp = 24, 16, 4
    log_case(f"masked layers {layers}, {w}x{h} {spp}spp depth 2")
    got, exp, st = render_both(scn, cam_of(w, h), w, h, spp, 2)
    assert np.all(exp > 0)
    assert_close(got, exp, frac=0.999)
    assert st["shadow_rays"] == w * h * spp * (layers + 1)


def test_shadow_walk_many_layers_with_graph_replay_enabled():
    """Scenes with cut-outs render eagerly even with graph replay on (the walk reads a count on
    the host between rounds of steps): same image, no pass replayed."""
    from rendu_amd.pathtracer import PathTracer
    scn, cam_of = scenes.masked_layers(18)
    cam = cam_of(16, 16)
    log_case("masked layers 18 with graph replay enabled, 4 passes")
    pt = PathTracer(scn)
    a = np.zeros((16, 16, 3), np.float32)
    pt.render(cam, 4, 2, a, seed=3, max_paths=16 * 16)
    pt.set_graph(True)
    b = np.zeros((16, 16, 3), np.float32)
    st = pt.render(cam, 4, 2, b, seed=3, max_paths=16 * 16)
    assert st["graph_passes"] == 0
    assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [S.BG_COLOR, S.BG_IMAGE, S.BG_SKYBOX, S.BG_LATLONG])
def test_background_direct_and_indirect_rules(mode):
    scn, cam_of = scenes.background_probe(mode)
    w, h = 32, 24
    log_case(f"background probe mode {mode}")
    got, exp, _ = render_both(scn, cam_of(w, h), w, h, 16, 4)
    assert_close(got, exp)
    if mode in (S.BG_COLOR, S.BG_IMAGE):
        assert np.all(got[-6:] == 0)
    else:
        assert got[-6:].mean() > 0.3
