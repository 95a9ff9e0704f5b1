"""GPU parity of the float colour stores (full, downsampled, quad canvas)
under each colour attachment and of the Kawase blur pass: bit for bit
against the oracle."""
import numpy as np
import pytest
import torch

import float_store_scenes as F
from test_inorder_gpu import _check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("build", F.ALL, ids=lambda f: f().name)
def test_float_store_scene_matches_oracle(build):
    ref, rst, gst = _check(build())
    assert rst["fragments"] > 1000


@pytest.mark.parametrize("h,w", [(360, 640), (23, 37), (7, 5)])
def test_kawase_passes_match_oracle(h, w):
    from oracle_backend import lib as olib
    from rsr_amd import rgl
    L = olib()
    rng = np.random.default_rng(h * 1000 + w)
    img = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    src = torch.from_numpy(img).cuda()
    cur = img
    for d in (0, 1, 2, 2, 3):   # a glow chain's distances
        print(f"[kawase {w}x{h} d={d}]", flush=True)
        dst = torch.full_like(src, float("nan"))
        rgl.kawase_blur(src, dst, d)
        torch.cuda.synchronize()
        want = np.full_like(cur, np.nan)
        L.rglo_kawase_blur(np.ascontiguousarray(cur).ctypes.data, want.ctypes.data, w, h, w, w, d)
        got = dst.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), d
        src, cur = dst, want
