"""CPU checks of the oracle's in-order tile semantics (the checker of the
GPU's in-order path): scissor, alpha blending, depth test and write masks,
clears and stores between draws, each against what the reference's command
semantics fix independently of the oracle's code."""
import numpy as np
import pytest

import inorder_scenes as S
from oracle_backend import render_calls as oracle_calls
from rsr_amd import scenes


@pytest.mark.parametrize("build", S.ALL, ids=lambda f: getattr(f, "__name__", "variant"))
def test_oracle_draws_every_call_scene(build):
    sc = build()
    outs, st = oracle_calls(sc, threads=4)
    for kind, o in zip(sc.outputs, outs):
        if kind == "depth":
            assert not np.isnan(o).any()
        else:
            assert not (o == 0x5A5A5A5A).all()


def _full(sc):
    return oracle_calls(sc, threads=4)[0]


def test_scissor_keeps_exactly_the_rect():
    # GPU::ScissorRect (rglv_gpu.hxx:258-263): columns [x, x + w), rows
    # [H - y - 1 - h, H - y - 1); inside it the scissored frame equals the
    # unscissored one, outside it holds the clear values
    for (x, y, w, h, W, H, tb) in ((37, 23, 301, 201, 640, 360, (8, 8)), (11, 7, 250, 151, 333, 181, (4, 4))):
        cut = _full(S.scene_scissor(x, y, w, h, W, H, tb))
        whole = _full(S.scene_scissor(0, -1, W, H + 1, W, H, tb))   # the rect covering the target
        top, bottom = H - y - 1 - h, H - y - 1
        inside = np.zeros((H, W), bool)
        inside[max(top, 0):bottom, x:x + w] = True
        for k, kind in enumerate(("color", "depth", "prim")):
            a, b = cut[k].view(np.uint32), whole[k].view(np.uint32)
            assert np.array_equal(a[inside], b[inside]), kind
        assert np.all(cut[1][~inside] == 1.0) and np.all(cut[2][~inside] == 0xFFFFFFFF)


def test_depth_test_off_last_draw_wins():
    # without a depth test every covered pixel takes the later draw
    off = _full(S.scene_depth_test_off())
    sc, pro, _ = S.base(name="ground_alone")
    sc.calls = pro + [("clear", S.CB | S.DB)] + S.lambert(sc) + S.draw(sc, "ground") + S.stores()
    sc.outputs = S.OUT3
    ground = _full(sc)
    gcov = ground[2] != 0xFFFFFFFF
    sphere_tris = 960
    assert np.array_equal(off[0][gcov], ground[0][gcov])
    assert np.array_equal(off[2][gcov], ground[2][gcov] + sphere_tris)     # ids count from the frame's first triangle
    assert np.array_equal(off[1].view(np.uint32)[gcov], ground[1].view(np.uint32)[gcov])


def test_prepass_equal_equals_single_pass():
    # a depth pre-pass then an EQUAL colour pass draws what one LESS pass of
    # the colour program draws, where the pre-pass and the colour pass agree
    eq = _full(S.scene_prepass_equal())
    sc, pro, _ = S.base(name="single_pass_amy")
    sc.calls = pro + [("clear", S.CB | S.DB)] + S.amy_alpha(sc) + S.draw(sc, "ground") + S.draw(sc, "sphere") + S.stores()
    sc.outputs = S.OUT3
    one = _full(sc)
    assert np.array_equal(eq[1].view(np.uint32), one[1].view(np.uint32))
    # the EQUAL pass redraws the same triangles: ids are those of the second
    # pair of draws (offset by the first pair's triangles)
    cov = one[2] != 0xFFFFFFFF
    assert np.array_equal(eq[2][cov], one[2][cov] + 962)
    assert np.array_equal(eq[0][cov], one[0][cov])


def test_blend_alpha_closed_form():
    # one constant-colour layer of alpha a over the clear colour c0:
    # c = src * a + c0 * (1 - a); alpha 1 replaces (ManyProgram)
    W, H = 64, 48
    tri = np.array([[-3, -2, -5], [3, -2, -5], [0, 3, -5]], np.float32)
    b = {k: scenes.pad4(tri[:, k]) for k in range(3)}
    b[9] = scenes.pad4(np.full(3, 0.25, np.float32)); b[10] = scenes.pad4(np.full(3, 0.5, np.float32))
    rgba = np.zeros((4, 4, 4), np.uint8); rgba[..., 0] = 255; rgba[..., 1] = 128; rgba[..., 3] = 64
    tex = scenes.make_mipmap_stack(rgba)
    sc = scenes.CallScene(name="blend_closed_form", width=W, height=H, tile_blocks=(2, 2),
                          arrays={"t.0": b[0], "t.1": b[1], "t.2": b[2], "t.9": b[9], "t.10": b[10],
                                  "t.idx": np.arange(3, dtype=np.uint16), "tex": tex})
    m = scenes.Mesh({k: b[k] for k in (0, 1, 2, 9, 10)}, np.arange(3, dtype=np.uint16))
    sc.calls = [("renderbuffer_type", scenes.GL_COLOR_ATTACHMENT0, scenes.RB_RGBF32),
                ("renderbuffer_type", scenes.GL_DEPTH_ATTACHMENT, scenes.RB_F32),
                ("clear_color", 0.0, 0.0, 1.0), ("clear_depth", 1.0), ("clear", S.CB | S.DB),
                ("view_matrix", scenes.mat4_rows(np.eye(4, dtype=np.float32))),
                ("projection_matrix", scenes.perspective2(60.0, W / H, 0.5, 50.0)),
                ("enable", scenes.GL_BLEND), ("use_program", scenes.PGM_AMY),
                ("bind_texture", 0, "@tex", 4, 4, 4, scenes.GL_NEAREST_MIPMAP_NEAREST)] + \
        scenes.mesh_calls("t", m) + [("use_program", 1), ("store_color_tc", ("out", 0), False)]
    sc.outputs = ["color"]
    col = _full(sc)[0]
    a = np.float32(64) / np.float32(255)
    tab = scenes.srgb8_to_float_table()
    want_r = tab[255] * a + np.float32(0.0) * (np.float32(1) - a)
    want_g = tab[128] * a
    want_b = np.float32(0.0) * a + np.float32(1.0) * (np.float32(1) - a)
    lin = lambda v: int(np.float32(min(max(v, 0), 1)) * np.float32(255))
    covered = col != 0x0000FF
    assert covered.sum() > 300
    assert np.all(col[covered] == ((lin(want_r) << 16) | (lin(want_g) << 8) | lin(want_b)))


def test_mid_frame_stores_snapshot_the_tile():
    # a store between draws writes the frame as it was at that point: the
    # first colour store equals a frame of the ground alone
    outs = _full(S.scene_mid_frame_clears())
    sc, pro, _ = S.base(name="ground_alone")
    sc.calls = pro + [("clear", S.CB | S.DB)] + S.lambert(sc) + S.draw(sc, "ground") + S.stores()
    sc.outputs = S.OUT3
    ground = _full(sc)
    assert np.array_equal(outs[0], ground[0])
    # after the depth-only clear the sphere covers the ground wherever it is
    # drawn, and its ids follow the ground's two triangles
    cov = (outs[3] != 0xFFFFFFFF) & (outs[3] >= 2)
    assert cov.sum() > 1000 and np.all(outs[2][cov] < 1.0)


def test_tile_size_invariance_up_to_128px():
    # the raster is invariant to the tile grid: 128, 96 and 64 px tiles
    # store the same buffers (no scissor in the frame)
    from test_tiles128_gpu import _plain
    base = _full(_plain((8, 8)))
    for tb in ((16, 16), (16, 8), (12, 12)):
        got = _full(_plain(tb))
        for a, b in zip(base, got):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), tb
