"""Call-list scenes for the parity tests of states beyond a single pass:
alpha blending, scissor, depth test and write masks, GL_EQUAL, clears and
stores between draws, RB_COLOR_DEPTH, several programs in one frame,
DrawArrays, viewports, the NEAREST-mip and non-power-of-two samplers, the
Exposure post program and the linear store.  Each builder returns a
scenes.CallScene that both the GPU and the oracle draw."""
from __future__ import annotations

import numpy as np

from rsr_amd import scenes
from rsr_amd.scenes import (GL_BLEND, GL_COLOR_ATTACHMENT0, GL_COLOR_BUFFER_BIT, GL_CULL_FACE, GL_DEPTH_ATTACHMENT,
                            GL_DEPTH_BUFFER_BIT, GL_DEPTH_TEST, GL_EQUAL, GL_FRONT, GL_BACK, GL_LEQUAL, GL_LESS,
                            GL_SCISSOR_TEST, GL_TRIANGLES, RB_COLOR_DEPTH, RB_F32, RB_RGBAF32, RB_RGBF32)

CB, DB = GL_COLOR_BUFFER_BIT, GL_DEPTH_BUFFER_BIT


def _rgba_texture(dim, seed, alpha_lo=60, alpha_hi=230):
    """Value-noise colour with a value-noise alpha in [alpha_lo, alpha_hi]."""
    t = scenes.value_noise_rgba8(dim, seed)
    a = scenes.value_noise_rgba8(dim, seed + 17)[..., 0].astype(np.float64) / 255.0
    t[..., 3] = np.round(alpha_lo + (alpha_hi - alpha_lo) * a).astype(np.uint8)
    return t


def base(width=640, height=360, tile_blocks=(8, 8), name="calls"):
    """The sphere-and-ground geometry with its arrays; returns (scene, the
    prologue calls: attachments, clear values, camera)."""
    sg = scenes.scene_sphere_ground(width, height)
    arrays = {}
    arrays.update(scenes.mesh_arrays("ground", sg.meshes[0]))
    arrays.update(scenes.mesh_arrays("sphere", sg.meshes[1]))
    arrays["tex"] = sg.texture
    arrays["alpha_tex"] = scenes.make_mipmap_stack(_rgba_texture(128, 5))
    arrays["uni"] = sg.uniforms
    sc = scenes.CallScene(name=name, width=width, height=height, tile_blocks=tile_blocks, arrays=arrays)
    pro = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_RGBF32), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_F32),
           ("clear_color", 0.1, 0.1, 0.15), ("clear_depth", 1.0),
           ("view_matrix", sg.view), ("projection_matrix", sg.projection),
           ("enable", GL_CULL_FACE), ("cull_face", GL_BACK)]
    return sc, pro, sg


def lambert(sc):
    return [("use_program", scenes.PGM_LAMBERT_TEX), ("alloc_uniforms", "@uni"),
            ("bind_texture", 0, "@tex", 256, 256, 256, scenes.GL_LINEAR_MIPMAP_NEAREST)]


def amy_alpha(sc):
    return [("use_program", scenes.PGM_AMY), ("bind_texture", 0, "@alpha_tex", 128, 128, 128, scenes.GL_LINEAR_MIPMAP_NEAREST)]


def draw(sc, mesh):
    m = scenes.Mesh({int(k.split(".")[1]): v for k, v in sc.arrays.items()
                     if k.startswith(mesh + ".") and k.split(".")[1].isdigit()}, sc.arrays[mesh + ".idx"])
    return scenes.mesh_calls(mesh, m)


def stores(post=scenes.PGM_DEFAULT_POST, gamma=True, first=0):
    return [("use_program", post), ("store_color_tc", ("out", first), gamma), ("store_depth", ("out", first + 1)),
            ("store_primitive_id", ("out", first + 2))]


OUT3 = ["color", "depth", "prim"]


def scene_blend():
    # opaque ground, then the sphere alpha-blended over it (AmyProgram's
    # texture alpha, BlendAlpha)
    sc, pro, _ = base(name="blend_sphere_over_ground")
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + [("enable", GL_BLEND)] + amy_alpha(sc) + \
        draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_blend_no_depth_write():
    # transparent pass without depth writes after the opaque one (the usual
    # order), with the sphere drawn twice under different cull faces
    sc, pro, _ = base(name="blend_nodepthwrite")
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + [("enable", GL_BLEND), ("depth_write_mask", False)] + \
        amy_alpha(sc) + [("cull_face", GL_FRONT)] + draw(sc, "sphere") + [("cull_face", GL_BACK)] + draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_scissor(x=37, y=23, w=301, h=201, width=640, height=360, tile_blocks=(8, 8), name="scissor_odd"):
    sc, pro, _ = base(width, height, tile_blocks, name=name)
    sc.calls = pro + [("clear", CB | DB), ("enable", GL_SCISSOR_TEST), ("scissor", x, y, w, h)] + lambert(sc) + \
        draw(sc, "ground") + draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_depth_test_off():
    # the ground drawn last without a depth test covers the sphere
    sc, pro, _ = base(name="depth_test_off")
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "sphere") + [("disable", GL_DEPTH_TEST)] + \
        draw(sc, "ground") + stores()
    sc.outputs = OUT3
    return sc


def scene_prepass_equal():
    # depth pre-pass with colour writes off, then the colour pass with
    # GL_EQUAL and depth writes off
    sc, pro, _ = base(name="prepass_equal")
    sc.calls = pro + [("clear", CB | DB), ("color_write_mask", False)] + lambert(sc) + draw(sc, "ground") + \
        draw(sc, "sphere") + [("color_write_mask", True), ("depth_func", GL_EQUAL), ("depth_write_mask", False)] + \
        amy_alpha(sc) + draw(sc, "ground") + draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_mid_frame_clears():
    # ground -> store; depth-only clear; sphere over everything -> store;
    # colour clear with another colour; sphere again with LEQUAL -> store
    sc, pro, _ = base(name="mid_frame_clears")
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + \
        [("use_program", scenes.PGM_DEFAULT_POST), ("store_color_tc", ("out", 0), True)] + \
        [("clear", DB)] + lambert(sc) + draw(sc, "sphere") + stores(first=1) + \
        [("clear_color", 0.6, 0.2, 0.1), ("clear", CB), ("depth_func", GL_LEQUAL)] + lambert(sc) + draw(sc, "sphere") + \
        [("use_program", scenes.PGM_EXPOSURE_POST), ("alloc_uniforms", "@exposure"), ("store_color_tc", ("out", 4), False),
         ("store_depth", ("out", 5))]
    sc.arrays["exposure"] = np.array([1.7] + [0.0] * 31, np.float32)
    sc.outputs = ["color", "color", "depth", "prim", "color", "depth"]
    return sc


def scene_shadow_pass():
    # the $layer shadow-map pass (node/gllayer.cxx:44-45, 152-181):
    # BaseProgram, colour writes off, depth-only clear, front faces culled,
    # depth stored
    sc, pro, _ = base(256, 256, name="shadow_pass")
    sc.calls = pro + [("color_write_mask", False), ("depth_write_mask", True), ("depth_func", GL_LESS),
                      ("enable", GL_CULL_FACE), ("cull_face", GL_FRONT), ("clear_depth", 1.0), ("clear", DB),
                      ("use_program", scenes.PGM_BASE)] + \
        draw(sc, "ground") + draw(sc, "sphere") + [("store_depth", ("out", 0)), ("store_primitive_id", ("out", 1))]
    sc.outputs = ["depth", "prim"]
    return sc


def scene_color_depth():
    # the viewer's default attachments (node/gpu.cxx:130-136): depth in the
    # colour buffer's alpha
    sc, pro, _ = base(name="rb_color_depth")
    pro = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_COLOR_DEPTH), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_COLOR_DEPTH)] + pro[2:]
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_color_depth_blend():
    sc, pro, _ = base(name="rb_color_depth_blend")
    pro = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_RGBAF32), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_F32)] + pro[2:]
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + [("enable", GL_BLEND)] + amy_alpha(sc) + \
        draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_mixed_depth_funcs():
    sc, pro, _ = base(name="mixed_depth_funcs")
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + [("depth_func", GL_LEQUAL)] + \
        amy_alpha(sc) + draw(sc, "sphere") + draw(sc, "ground") + stores()
    sc.outputs = OUT3
    return sc


def scene_mixed_programs():
    # several programs and state changes between draws on the key path
    sc, pro, _ = base(name="mixed_programs")
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + amy_alpha(sc) + draw(sc, "sphere") + \
        [("disable", GL_CULL_FACE)] + lambert(sc) + draw(sc, "ground") + stores()
    sc.outputs = OUT3
    return sc


def scene_exposure_linear(inorder=False):
    sc, pro, _ = base(name="exposure_linear" + ("_inorder" if inorder else ""))
    extra = [("enable", GL_BLEND)] + amy_alpha(sc) if inorder else []
    sc.arrays["exposure"] = np.array([1.7] + [0.0] * 31, np.float32)
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + extra + draw(sc, "sphere") + \
        [("use_program", scenes.PGM_EXPOSURE_POST), ("alloc_uniforms", "@exposure"), ("store_color_tc", ("out", 0), False),
         ("store_depth", ("out", 1)), ("store_primitive_id", ("out", 2))]
    sc.outputs = OUT3
    return sc


def scene_draw_arrays():
    # the sphere de-indexed and drawn by DrawArrays, then an instanced
    # DrawArrays with ManyProgram
    sc, pro, sg = base(name="draw_arrays")
    sp = sg.meshes[1]
    idx = sp.indices.astype(np.int64)
    flat = {k: scenes.pad4(v[idx]) for k, v in sp.buffers.items()}
    for k, v in flat.items():
        sc.arrays[f"flat.{k}"] = v
    mats = scenes.instance_matrices(3, 4, (0.0, 0.5, 0.0), 4.0, 0.3, 0.6)
    sc.arrays["flat.15"] = mats.reshape(-1)
    sc.arrays["many_uni"] = np.array([0.3] + [0.0] * 31, np.float32)
    n = len(idx)
    use = [("use_buffer", s, f"@flat.{s}" if f"flat.{s}" in sc.arrays else None) for s in range(16)]
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + use + [("draw_arrays", GL_TRIANGLES, 0, n)] + \
        [("use_program", scenes.PGM_MANY), ("alloc_uniforms", "@many_uni")] + use + \
        [("draw_arrays_instanced", GL_TRIANGLES, 0, n, 3)] + stores()
    sc.outputs = OUT3
    return sc


def scene_viewport():
    sc, pro, _ = base(name="viewport")
    sc.calls = pro + [("clear", CB | DB), ("viewport", 40, 30, 400, 250)] + lambert(sc) + draw(sc, "ground") + \
        draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_samplers():
    # NEAREST_MIPMAP_NEAREST on the ground, a 200 x 120 non-power-of-two
    # texture (stride 208) on the sphere
    sc, pro, _ = base(name="samplers_nearest_npot")
    npot = scenes.value_noise_rgba8(256, 9)[:120, :208]
    tab = scenes.srgb8_to_float_table()
    t = np.zeros((120, 208, 4), np.float32)
    t[..., :3] = tab[npot[..., :3]]
    t[..., 3] = 1.0
    sc.arrays["npot"] = np.ascontiguousarray(t.reshape(-1))
    sc.calls = pro + [("clear", CB | DB), ("use_program", scenes.PGM_LAMBERT_TEX), ("alloc_uniforms", "@uni"),
                      ("bind_texture", 0, "@tex", 256, 256, 256, scenes.GL_NEAREST_MIPMAP_NEAREST)] + draw(sc, "ground") + \
        [("bind_texture", 0, "@npot", 200, 120, 208, scenes.GL_LINEAR_MIPMAP_NEAREST)] + draw(sc, "sphere") + stores()
    sc.outputs = OUT3
    return sc


def scene_dense_blend(width=640, height=360, instances=160, seed=31):
    # many small icospheres in one screen region, alpha-blended over an
    # opaque ground: tiles hold more than 4096 (triangle, tile) records, so
    # the in-order sort takes its LDS-runs-and-merge form
    sc, pro, sg = base(width, height, name=f"dense_blend{instances}")
    pos, nrm, uv, idx = scenes.icosphere(3)
    b = scenes._mesh_buffers(pos, nrm, uv)
    b[15] = scenes.instance_matrices(instances, seed, (0.6, 1.2, 0.0), 1.2, 0.25, 0.5).reshape(-1)
    m = scenes.Mesh(b, idx, instances)
    sc.arrays.update(scenes.mesh_arrays("ico", m))
    sc.arrays["many_uni"] = np.array([0.45] + [0.0] * 31, np.float32)
    sc.calls = pro + [("clear", CB | DB)] + lambert(sc) + draw(sc, "ground") + \
        [("enable", GL_BLEND), ("use_program", scenes.PGM_MANY), ("alloc_uniforms", "@many_uni")] + \
        scenes.mesh_calls("ico", m) + stores()
    sc.outputs = OUT3
    return sc


ALL = [scene_blend, scene_blend_no_depth_write, scene_scissor, scene_depth_test_off, scene_prepass_equal,
       scene_mid_frame_clears, scene_shadow_pass, scene_color_depth, scene_color_depth_blend, scene_mixed_depth_funcs,
       scene_mixed_programs, scene_exposure_linear, lambda: scene_exposure_linear(True), scene_draw_arrays,
       scene_viewport, scene_samplers,
       lambda: scene_scissor(-20, 10, 700, 150, name="scissor_past_edges"),
       lambda: scene_scissor(11, 7, 250, 151, 333, 181, (4, 4), name="scissor_odd_target_tile32")]
