"""GPU parity of the viewer's other programs (viewer/shaders.hxx): every
output of every store against the oracle, bit for bit, on the key path and
on the in-order path, and OBJ2S lit through a shadow map that a frame
drawn from the light stored."""
import numpy as np
import pytest
import torch  # noqa: F401

import program_scenes as P
from test_inorder_gpu import _check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("build", P.ALL, ids=lambda f: f().name)
def test_program_scene_matches_oracle(build):
    sc = build()
    ref, rst, gst = _check(sc)
    assert rst["fragments"] > 1000


def test_obj2s_with_the_stored_shadow_map():
    from oracle_backend import render_calls as oracle_calls
    from rsr_amd import rgl
    # frame 1: the light's depth pass; the GPU's stored depth equals the oracle's
    sm = P.scene_shadow_map()
    print(f"[{sm.name}]", flush=True)
    (ref_map,), _ = oracle_calls(sm, threads=8)
    (gpu_map,), _ = rgl.render_calls(sm)
    assert np.array_equal(ref_map.view(np.uint32), gpu_map.view(np.uint32))
    # frame 2: OBJ2S reading that depth through the depth texture unit
    for inorder in (False, True):
        sc = P.scene_obj2s(gpu_map, inorder=inorder)
        ref, rst, gst = _check(sc)
        assert rst["fragments"] > 1000
