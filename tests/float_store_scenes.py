"""Call-list scenes for the float colour stores (GL::StoreColor of a
FloatingPointCanvas, full or downsampled, and of a QFloat4Canvas;
rglv_gpu.cxx:345-384) under each colour attachment, next to the truecolor,
depth and id stores of the same frame."""
from __future__ import annotations

from rsr_amd import scenes
from rsr_amd.scenes import GL_BLEND, GL_COLOR_ATTACHMENT0, GL_DEPTH_ATTACHMENT, RB_COLOR_DEPTH, RB_F32, RB_RGBAF32, RB_RGBF32

import inorder_scenes as S

CB, DB = S.CB, S.DB
OUTS = ["color", "depth", "prim", "rgba", "rgba_half", "quads"]


def float_stores(first=0):
    return S.stores(first=first) + [("store_color_fp", ("out", first + 3), 0, False),
                                    ("store_color_fp", ("out", first + 4), 0, True),
                                    ("store_color_quads", ("out", first + 5), 0)]


def scene_float_stores(attach=RB_RGBF32, width=640, height=360, tile_blocks=(8, 8), blend=False):
    names = {RB_RGBF32: "rgbf32", RB_RGBAF32: "rgbaf32", RB_COLOR_DEPTH: "color_depth"}
    sc, pro, _ = S.base(width, height, tile_blocks,
                        name=f"float_stores_{names[attach]}_{width}x{height}_t{tile_blocks[0] * 8}" + ("_blend" if blend else ""))
    depth_type = RB_COLOR_DEPTH if attach == RB_COLOR_DEPTH else RB_F32
    pro = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, attach), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, depth_type)] + pro[2:]
    body = S.lambert(sc) + S.draw(sc, "ground") + ([("enable", GL_BLEND)] + S.amy_alpha(sc) if blend else []) + \
        S.draw(sc, "sphere")
    sc.calls = pro + [("clear", CB | DB)] + body + float_stores()
    sc.outputs = OUTS
    return sc


def scene_float_stores_mid_frame():
    # a float store between draws, then more drawing and the final stores
    sc, pro, _ = S.base(name="float_stores_mid_frame")
    sc.calls = pro + [("clear", CB | DB)] + S.lambert(sc) + S.draw(sc, "ground") + \
        [("store_color_fp", ("out", 0), 0, True)] + S.draw(sc, "sphere") + float_stores(first=1)
    sc.outputs = ["rgba_half"] + OUTS
    return sc


ALL = [scene_float_stores, lambda: scene_float_stores(RB_RGBAF32), lambda: scene_float_stores(RB_COLOR_DEPTH),
       lambda: scene_float_stores(RB_RGBF32, 328, 186, (4, 4)), lambda: scene_float_stores(RB_RGBAF32, blend=True),
       scene_float_stores_mid_frame]
