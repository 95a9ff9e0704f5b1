"""CPU pins of the oracle's float colour stores and Kawase blur against
what fixes them independently: the truecolor store of the same frame
(the float store's pixels encode to its bytes), box means and quad
layout recomputed from the full float store, the depth store (the
RB_COLOR_DEPTH quad alpha), and closed forms of the blur."""
import ctypes

import numpy as np
import pytest

import float_store_scenes as F
from oracle_backend import lib as olib, render_calls as oracle_calls
from rsr_amd.scenes import RB_COLOR_DEPTH, RB_RGBAF32, RB_RGBF32


def _srgb8(v):
    L = olib()
    return np.vectorize(lambda x: L.rglo_to_srgb8(float(x)), otypes=[np.uint32])(v)


@pytest.mark.parametrize("attach", [RB_RGBF32, RB_RGBAF32, RB_COLOR_DEPTH])
def test_float_stores_agree_with_the_truecolor_and_depth_stores(attach):
    sc = F.scene_float_stores(attach, 328, 186, (4, 4))
    color, depth, prim, rgba, half, quads = oracle_calls(sc, threads=4)[0]
    h, w = depth.shape
    # full store: R, G, B encode to the truecolor bytes; alpha 0
    tc = (_srgb8(rgba[..., 0]) << 16) | (_srgb8(rgba[..., 1]) << 8) | _srgb8(rgba[..., 2])
    assert np.array_equal(tc, color)
    assert np.all(rgba[..., 3] == 0.0)
    # downsampled: the 2 x 2 box mean in lane order, alpha 0
    f = rgba[..., :3]
    box = (((f[0::2, 0::2] + f[0::2, 1::2]) + f[1::2, 0::2]) + f[1::2, 1::2]) * np.float32(0.25)
    assert np.array_equal(half[..., :3], box) and np.all(half[..., 3] == 0.0)
    # quad canvas: r[4] g[4] b[4] a[4] per quad, lane = (y & 1) * 2 + (x & 1)
    for c in range(3):
        for lane in range(4):
            assert np.array_equal(quads[..., 4 * c + lane], f[lane >> 1::2, lane & 1::2, c])
    alpha = quads[..., 12:16]
    if attach == RB_RGBF32 or attach == RB_RGBAF32:   # RGBA: the clear's alpha 1, never drawn
        assert np.all(alpha == 1.0)
    else:                                              # RB_COLOR_DEPTH: the depth
        for lane in range(4):
            assert np.array_equal(alpha[..., lane], depth[lane >> 1::2, lane & 1::2])


def _kawase(src, d):
    L = olib()
    src = np.ascontiguousarray(src, np.float32)
    h, w = src.shape[:2]
    dst = np.full_like(src, np.nan)
    L.rglo_kawase_blur(src.ctypes.data, dst.ctypes.data, w, h, w, w, d)
    return dst


def test_kawase_closed_forms():
    # KawaseBlurFilter (rglr_kawase.cxx:21-85): the mean of four 2 x 2 boxes
    # at (x - d - 1 .. x - d) and (x + d .. x + d + 1) in each axis, clamped
    rng = np.random.default_rng(3)
    h, w = 41, 57
    const = np.full((h, w, 4), 0.375, np.float32)
    for d in range(4):
        assert np.all(_kawase(const, d) == np.float32(0.375))
    # an affine image (dyadic values, exact sums) is reproduced away from the
    # clamped border: the boxes' centroid is the pixel
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float32)
    aff = np.stack([xs * 0.25 + ys * 0.5, xs * -0.125 + 3.0, ys * 0.0625, xs * 0 + 1.0], -1).astype(np.float32)
    for d in range(4):
        m = d + 2
        out = _kawase(aff, d)
        assert np.array_equal(out[m:h - m, m:w - m], aff[m:h - m, m:w - m]), d
    # everywhere, clamped sampling evaluated in float64
    img = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    for d in range(4):
        want = np.zeros((h, w, 4))
        for oy in (-d - 1, d):
            for ox in (-d - 1, d):
                for dy in (0, 1):
                    for dx in (0, 1):
                        yy = np.clip(np.arange(h) + oy + dy, 0, h - 1)
                        xx = np.clip(np.arange(w) + ox + dx, 0, w - 1)
                        want += img[yy][:, xx].astype(np.float64)
        np.testing.assert_allclose(_kawase(img, d), want / 16, rtol=1e-6, atol=1e-7)
