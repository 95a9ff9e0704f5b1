"""GPU parity: the HIP pipeline against the CPU oracle on the same seeded
scenes, through the C ABI.  Coverage, depth and primitive-ID buffers must
be bit-exact; colour must be bit-exact too (the plain path's tolerance is
1/255 per channel, asserted separately so a failure says which)."""
import numpy as np
import torch  # noqa: F401  (first import on a fresh box is slow; do it at collection)
import pytest

from rsr_amd import scenes

pytestmark = pytest.mark.gpu


def _gpu_render(scene):
    import torch  # noqa: F401
    from rsr_amd import rgl
    return rgl.render(scene)


def _compare(name, ref, got):
    rc, rd, rp, rst = ref
    gc, gd, gp, gst = got
    print(f"[{name}] oracle stats {rst} gpu stats {gst}", flush=True)
    cov_r, cov_g = rd < 1.0, gd < 1.0
    assert np.array_equal(cov_r, cov_g), f"{name}: coverage differs at {np.argwhere(cov_r != cov_g)[:8].tolist()}"
    assert np.array_equal(rd.view(np.uint32), gd.view(np.uint32)), \
        f"{name}: depth bits differ at {np.argwhere(rd.view(np.uint32) != gd.view(np.uint32))[:8].tolist()}"
    assert np.array_equal(rp, gp), f"{name}: primitive ids differ at {np.argwhere(rp != gp)[:8].tolist()}"
    ch = lambda c, s: ((c >> s) & 255).astype(np.int32)
    for s in (16, 8, 0):
        assert np.abs(ch(rc, s) - ch(gc, s)).max() <= 1, f"{name}: colour channel beyond 1/255"
    assert np.array_equal(rc, gc), f"{name}: colour differs at {np.argwhere(rc != gc)[:8].tolist()}"
    assert rst["fragments"] == gst["fragments"], f"{name}: fragment counts differ"


def test_sphere_ground_640x360():
    from oracle_backend import render as oracle_render
    s = scenes.scene_sphere_ground()
    _compare(s.name, oracle_render(s), _gpu_render(s))


def test_icospheres_1080p():
    from oracle_backend import render as oracle_render
    s = scenes.scene_icospheres()
    _compare(s.name, oracle_render(s, threads=8), _gpu_render(s))
