"""Call-list scenes for the viewer's other programs (viewer/shaders.hxx):
AlphaTexture (65) and Text (26) with their late depth test and alpha
discard, Depth (5) reading the depth texture unit, Pattern (41) sampling at
gl_FragCoord, OBJ1 (7), OBJ2 (8) and OBJ2S (9) with the view and normal
matrices, and OBJ2S lit by a spot light whose shadow map is the depth
stored by an earlier frame drawn from the light.  Each builder returns a
scenes.CallScene that both the GPU and the oracle draw; the "_inorder"
variants enable a full-target scissor, which routes the same draws through
the in-order tile path instead of the visibility-key path."""
from __future__ import annotations

import numpy as np

from rsr_amd import scenes
from rsr_amd.scenes import (GL_BLEND, GL_COLOR_ATTACHMENT0, GL_CULL_FACE, GL_DEPTH_ATTACHMENT, GL_FRONT,
                            GL_SCISSOR_TEST, RB_F32, RB_RGBF32)

import inorder_scenes as S

CB, DB = S.CB, S.DB


def _cutout_texture(dim, seed, threshold=0.45, alpha=200):
    """Value-noise colour, alpha 0 where a second noise is below threshold
    and alpha/255 elsewhere: holes for the alpha discard."""
    t = scenes.value_noise_rgba8(dim, seed)
    a = scenes.value_noise_rgba8(dim, seed + 23, lattice=8)[..., 0].astype(np.float64) / 255.0
    t[..., 3] = np.where(a < threshold, 0, alpha).astype(np.uint8)
    return scenes.make_mipmap_stack(t)


def _vertex_colours(sc, mesh, seed, lo, hi):
    """Per-vertex colours in slots 6-8 of *mesh* (the kd / colour inputs)."""
    n = len(sc.arrays[f"{mesh}.0"])
    rng = np.random.RandomState(seed)
    for k in range(3):
        sc.arrays[f"{mesh}.{6 + k}"] = rng.uniform(lo, hi, n).astype(np.float32)


def _inorder(calls, sc):
    """A scissor covering the whole target: nothing is cut, but the frame
    takes the in-order path."""
    return [("enable", GL_SCISSOR_TEST), ("scissor", 0, 0, sc.width, sc.height)] + calls


def _finish(sc, pro, body, inorder, post=None):
    sc.calls = pro + [("clear", CB | DB)] + (_inorder(body, sc) if inorder else body)
    sc.calls += post if post is not None else S.stores()
    sc.outputs = sc.outputs or S.OUT3
    if inorder:
        sc.name += "_inorder"
    return sc


def scene_alpha_texture(blend=False):
    # the sphere with AlphaTextureProgram and culling off: where its texture
    # alpha is 0 the fragments are discarded (before the late depth test),
    # so the ground and the sphere's inner faces show through the holes
    sc, pro, _ = S.base(name="alpha_texture" + ("_blend" if blend else ""))
    sc.arrays["cutout"] = _cutout_texture(128, 41)
    body = S.lambert(sc) + S.draw(sc, "ground") + [("disable", GL_CULL_FACE)] + \
        ([("enable", GL_BLEND)] if blend else []) + \
        [("use_program", scenes.PGM_ALPHA_TEXTURE),
         ("bind_texture", 0, "@cutout", 128, 128, 128, scenes.GL_LINEAR_MIPMAP_NEAREST)] + S.draw(sc, "sphere")
    return _finish(sc, pro, body, False)


def scene_text():
    # TextProgram: vertex colours times the texture, alpha-blended, with
    # discard; the ground drawn by it without colour buffers (white)
    sc, pro, _ = S.base(name="text")
    sc.arrays["cutout"] = _cutout_texture(128, 43, threshold=0.4, alpha=230)
    _vertex_colours(sc, "sphere", 3, 0.2, 1.0)
    body = [("use_program", scenes.PGM_TEXT),
            ("bind_texture", 0, "@cutout", 128, 128, 128, scenes.GL_LINEAR_MIPMAP_NEAREST)] + S.draw(sc, "ground") + \
        [("enable", GL_BLEND), ("disable", GL_CULL_FACE)] + S.draw(sc, "sphere")
    return _finish(sc, pro, body, False)


def _depth_map(dim, seed):
    """A smooth dim x dim float map in [-1, 1] (a stand-in depth texture)."""
    n = scenes.value_noise_rgba8(dim, seed, lattice=8)[..., 0].astype(np.float32) / np.float32(127.5)
    return np.ascontiguousarray((n - np.float32(1.0)).reshape(-1))


def scene_depth_texture(inorder=False):
    # DepthProgram: the depth unit at the vertex uv, linearised; the
    # ground's uv runs to 4, outside [0, 1) the unit returns -1
    sc, pro, _ = S.base(name="depth_texture")
    sc.arrays["dmap"] = _depth_map(128, 12)
    body = [("use_program", scenes.PGM_DEPTH), ("bind_texture3", "@dmap", 128)] + S.draw(sc, "ground") + \
        S.draw(sc, "sphere")
    return _finish(sc, pro, body, inorder)


def _pattern_uniforms(dim_y, ox, oy):
    u = np.zeros(32, np.float32)
    u[0], u[1] = ox, oy
    u[4], u[5] = dim_y * 1.75, dim_y  # dim = (width, height); the shader reads dim.y
    return u


def scene_pattern(inorder=False):
    # PatternProgram: texture 0 at gl_FragCoord / dim.y + offset
    sc, pro, _ = S.base(name="pattern")
    sc.arrays["pat_uni"] = _pattern_uniforms(200.0, 0.13, 0.37)
    body = [("use_program", scenes.PGM_PATTERN), ("alloc_uniforms", "@pat_uni"),
            ("bind_texture", 0, "@tex", 256, 256, 256, scenes.GL_LINEAR_MIPMAP_NEAREST)] + S.draw(sc, "ground") + \
        S.draw(sc, "sphere")
    return _finish(sc, pro, body, inorder)


def scene_pattern_clipped(inorder=False):
    # PatternProgram on the near-plane-crossing floor: clipped triangles are
    # drawn by the scalar rasteriser, whose gl_FragCoord rows differ from
    # the vector rasteriser's
    cf = scenes.scene_clip_floor(640, 360, 24)
    arrays = scenes.mesh_arrays("floor", cf.meshes[0])
    arrays["tex"] = cf.texture
    arrays["pat_uni"] = _pattern_uniforms(96.0, -0.21, 0.05)
    sc = scenes.CallScene(name="pattern_clipped", width=640, height=360, arrays=arrays)
    pro = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_RGBF32), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_F32),
           ("clear_color", 0.0, 0.0, 0.0), ("clear_depth", 1.0),
           ("view_matrix", cf.view), ("projection_matrix", cf.projection)]
    body = [("use_program", scenes.PGM_PATTERN), ("alloc_uniforms", "@pat_uni"),
            ("bind_texture", 0, "@tex", 256, 256, 256, scenes.GL_LINEAR_MIPMAP_NEAREST)] + \
        scenes.mesh_calls("floor", cf.meshes[0])
    return _finish(sc, pro, body, inorder)


def _obj_scene(program, name, inorder):
    sc, pro, _ = S.base(name=name)
    # kd small enough that the eye light's 100 / distance keeps the colour
    # below 1 over most of the image
    _vertex_colours(sc, "sphere", 5, 0.02, 0.09)
    _vertex_colours(sc, "ground", 6, 0.02, 0.06)
    body = [("use_program", program)] + S.draw(sc, "ground") + S.draw(sc, "sphere")
    return _finish(sc, pro, body, inorder)


def scene_obj1(inorder=False):
    return _obj_scene(scenes.PGM_OBJ1, "obj1", inorder)


def scene_obj2(inorder=False):
    return _obj_scene(scenes.PGM_OBJ2, "obj2", inorder)


def scene_obj_white_exposure():
    # OBJ1 and OBJ2 without colour buffers (kd defaults to white), stored
    # linear through ExposurePost at a low exposure
    sc, pro, _ = S.base(name="obj_white_exposure")
    sc.arrays["exposure"] = np.array([0.09] + [0.0] * 31, np.float32)
    body = [("use_program", scenes.PGM_OBJ1)] + S.draw(sc, "ground") + [("use_program", scenes.PGM_OBJ2)] + \
        S.draw(sc, "sphere")
    post = [("use_program", scenes.PGM_EXPOSURE_POST), ("alloc_uniforms", "@exposure"), ("store_color_tc", ("out", 0), False),
            ("store_depth", ("out", 1)), ("store_primitive_id", ("out", 2))]
    sc.outputs = S.OUT3
    return _finish(sc, pro, body, False, post)


# the spot light of the OBJ2S scenes (world space)
LIGHT_POS = (3.0, 6.0, 4.0)
LIGHT_AT = (0.0, 0.5, 0.0)
SHADOW_DIM = 256


def _light_matrices():
    view = scenes.look_at(LIGHT_POS, LIGHT_AT, (0, 1, 0))
    proj = scenes.perspective2(50.0, 1.0, 1.0, 30.0)
    return view, proj


def _col(m):
    return np.asarray(m, np.float64).reshape(4, 4).T  # column-major flat -> matrix


def scene_shadow_map():
    # the light's depth pass as $layer draws it (node/gllayer.cxx:44-45,
    # 152-181): BaseProgram, colour writes off, front faces culled, the
    # depth stored as the OBJ2S frame's shadow map
    sc, pro, _ = S.base(SHADOW_DIM, SHADOW_DIM, name="obj2s_shadow_map")
    view, proj = _light_matrices()
    pro = [p for p in pro if p[0] not in ("view_matrix", "projection_matrix")] + \
        [("view_matrix", view), ("projection_matrix", proj)]
    sc.calls = pro + [("color_write_mask", False), ("cull_face", GL_FRONT), ("clear", DB),
                      ("use_program", scenes.PGM_BASE)] + S.draw(sc, "ground") + S.draw(sc, "sphere") + \
        [("store_depth", ("out", 0))]
    sc.outputs = ["depth"]
    return sc


def scene_obj2s(shadow_map, inorder=False, cone_deg=22.0):
    # OBJ2SProgram: spot light at LIGHT_POS toward LIGHT_AT with the shadow
    # map in the depth texture unit; positions and the light in view space
    sc, pro, sg = S.base(name="obj2s")
    _vertex_colours(sc, "sphere", 7, 0.03, 0.08)
    _vertex_colours(sc, "ground", 8, 0.03, 0.06)
    lview, lproj = _light_matrices()
    vm = _col(sg.view)
    shadow = _col(lproj) @ _col(lview)
    lpos_v = vm @ np.array([*LIGHT_POS, 1.0])
    ldir_v = vm[:3, :3] @ (np.array(LIGHT_AT) - np.array(LIGHT_POS))
    u = np.zeros(32, np.float32)
    u[0:16] = shadow.T.reshape(16)
    u[16:19] = lpos_v[:3]
    u[19:22] = ldir_v
    u[22] = np.cos(np.radians(cone_deg))
    sc.arrays["obj2s_uni"] = u
    sc.arrays["shadow"] = np.ascontiguousarray(np.asarray(shadow_map, np.float32).reshape(-1))
    body = [("use_program", scenes.PGM_OBJ2S), ("alloc_uniforms", "@obj2s_uni"), ("bind_texture3", "@shadow", SHADOW_DIM)] + \
        S.draw(sc, "ground") + S.draw(sc, "sphere")
    return _finish(sc, pro, body, inorder)


def scene_envmap(two_pass=False):
    # EnvmapProgram: the sphere-map uv per vertex from the reflected view
    # vector; alone (the LESS install) or as the blended second pass of
    # shaders_envmap.cxx:22-25 (GL_EQUAL, no depth write, BlendAlpha with
    # its alpha 0.5) over a Lambert pass, in the viewer's RB_COLOR_DEPTH
    sc, pro, _ = S.base(name="envmap" + ("_two_pass" if two_pass else ""))
    pro = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, scenes.RB_COLOR_DEPTH),
           ("renderbuffer_type", GL_DEPTH_ATTACHMENT, scenes.RB_COLOR_DEPTH)] + pro[2:]
    sc.arrays["spheremap"] = scenes.make_mipmap_stack(scenes.value_noise_rgba8(256, 29, lattice=8))
    _vertex_colours(sc, "sphere", 9, -1.0, 1.0)   # face normals: weight 0 in the mix
    env = [("use_program", scenes.PGM_ENVMAP),
           ("bind_texture", 0, "@spheremap", 256, 256, 256, scenes.GL_LINEAR_MIPMAP_NEAREST)]
    if two_pass:
        body = S.lambert(sc) + S.draw(sc, "ground") + S.draw(sc, "sphere") + \
            [("depth_func", scenes.GL_EQUAL), ("depth_write_mask", False), ("enable", GL_BLEND)] + env + \
            S.draw(sc, "ground") + S.draw(sc, "sphere")
    else:
        body = env + S.draw(sc, "ground") + S.draw(sc, "sphere")
    return _finish(sc, pro, body, False)


def scene_iq_post(inorder=False, gamma=True, width=640, height=360, tile_blocks=(8, 8)):
    # IQPostProgram on the store: per-channel gamma, vignette and hash of
    # the canvas coordinate, which FilterTile steps from each tile's corner
    sc, pro, _ = S.base(width, height, tile_blocks, name=f"iq_post_{width}x{height}_t{tile_blocks[0] * 8}" +
                        ("" if gamma else "_linear"))
    body = S.lambert(sc) + S.draw(sc, "ground") + S.draw(sc, "sphere")
    return _finish(sc, pro, body, inorder, post=S.stores(post=scenes.PGM_IQ_POST, gamma=gamma))


# scenes that need no input from another frame
ALL = [scene_alpha_texture, lambda: scene_alpha_texture(True), scene_text, scene_depth_texture,
       lambda: scene_depth_texture(True), scene_pattern, lambda: scene_pattern(True), scene_pattern_clipped,
       lambda: scene_pattern_clipped(True), scene_obj1, lambda: scene_obj1(True), scene_obj2, lambda: scene_obj2(True),
       scene_obj_white_exposure, scene_envmap, lambda: scene_envmap(True), scene_iq_post, lambda: scene_iq_post(True), lambda: scene_iq_post(gamma=False),
       lambda: scene_iq_post(False, True, 328, 184, (4, 4))]
