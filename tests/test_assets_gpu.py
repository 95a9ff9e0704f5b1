"""The asset path's device side: RGBA8 images to float textures (sRGB
decode, premultiplied alpha, the power-of-two resample, the mip chain)
against the host restatements, and a decoded PNG drawn on the GPU against
the oracle."""
import io

import numpy as np
import pytest
import torch

from rsr_amd import assets, scenes

pytestmark = pytest.mark.gpu


def _host_texture(rgba8, premultiply=False):
    h, w = rgba8.shape[:2]
    tab = scenes.srgb8_to_float_table()
    f = np.empty((h, w, 4), np.float32)
    f[..., :3] = tab[rgba8[..., :3]]
    f[..., 3] = rgba8[..., 3].astype(np.float32) / np.float32(255.0)
    if premultiply:
        f[..., :3] *= f[..., 3:4]
    dim = 1 << max(1, (max(w, h) - 1).bit_length())
    if (w, h) != (dim, dim):   # ensurePowerOf2 (rglr_texture.cxx:13-31)
        iy = (np.arange(dim) / (dim - 1) * (h - 1)).astype(np.int64)
        ix = (np.arange(dim) / (dim - 1) * (w - 1)).astype(np.int64)
        f = f[iy][:, ix]
    out = np.zeros((2 * dim, dim, 4), np.float32)
    out[:dim] = f
    from rsr_amd import rgl
    flat = np.ascontiguousarray(out.reshape(-1))
    rgl.make_mipmap(flat, dim)            # the host Texture::maybe_make_mipmap
    return flat.reshape(2 * dim, dim, 4), dim


@pytest.mark.parametrize("h,w,pm", [(64, 64, False), (64, 64, True), (60, 100, False), (17, 5, True), (256, 256, False)])
def test_texture_from_rgba8_matches_host(h, w, pm):
    from rsr_amd import rgl
    rgba8 = np.random.default_rng(h * w).integers(0, 256, (h, w, 4), dtype=np.uint8)
    print(f"[texture {w}x{h} premultiply={pm}]", flush=True)
    got, dim = rgl.texture_from_rgba8(rgba8, premultiply=pm)
    want, wdim = _host_texture(rgba8, pm)
    assert dim == wdim
    assert np.array_equal(got.cpu().numpy().view(np.uint32), want.view(np.uint32))


def test_png_texture_drawn_matches_oracle():
    # a PNG decoded on the host, its texture made on the device, drawn by
    # AmyProgram over the sphere; the oracle draws the host-made texture
    PIL = pytest.importorskip("PIL.Image")
    from test_inorder_gpu import _check
    import inorder_scenes as S
    from rsr_amd import rgl
    rng = np.random.default_rng(5)
    img = (scenes.value_noise_rgba8(128, 31)[:, :, :3] // 2 + rng.integers(0, 64, (128, 128, 3))).astype(np.uint8)
    buf = io.BytesIO()
    PIL.fromarray(img, "RGB").save(buf, format="PNG")
    rgba8 = assets.decode_png(buf.getvalue())
    tex, dim = rgl.texture_from_rgba8(rgba8)
    host, _ = _host_texture(rgba8)
    assert np.array_equal(tex.cpu().numpy().view(np.uint32), host.view(np.uint32))
    sc, pro, _ = S.base(name="png_texture")
    sc.arrays["png"] = np.ascontiguousarray(tex.cpu().numpy().reshape(-1))
    sc.calls = pro + [("clear", S.CB | S.DB)] + S.lambert(sc) + S.draw(sc, "ground") + \
        [("use_program", scenes.PGM_AMY), ("bind_texture", 0, "@png", dim, dim, dim, scenes.GL_LINEAR_MIPMAP_NEAREST)] + \
        S.draw(sc, "sphere") + S.stores()
    sc.outputs = S.OUT3
    _check(sc)
