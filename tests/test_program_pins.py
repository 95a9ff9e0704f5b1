"""CPU pins of the oracle's restatement of the viewer's other programs
(viewer/shaders.hxx: AlphaTexture, Text, Depth, Pattern, OBJ1, OBJ2, OBJ2S)
by checks other than a retyping of their code: the alpha discard against
frames without the discarded draw and against AmyProgram, the depth unit
against a texel-index map, gl_FragCoord of both rasterisers against the
pixel grid through PatternProgram, and the lighting of OBJ1 / OBJ2 / OBJ2S
against a float64 evaluation with an exact matrix inverse."""
from __future__ import annotations

import ctypes
import math

import numpy as np

import inorder_scenes as S
import program_scenes as P
from oracle_backend import lib as olib, render_calls as oracle_calls
from rsr_amd import scenes
from rsr_amd.scenes import GL_COLOR_ATTACHMENT0, GL_CULL_FACE, GL_DEPTH_ATTACHMENT, RB_F32, RB_RGBF32

CB, DB = S.CB, S.DB
KMAXVARY = 16


def _buffers(buffers):
    keep, ptrs = [], (ctypes.c_void_p * 16)()
    for k, a in (buffers or {}).items():
        a = np.ascontiguousarray(a, np.float32)
        keep.append(a)
        ptrs[k] = a.ctypes.data
    return keep, ptrs


def shade_ex(pid, vary, uniforms=None, buffers=None, tex=None, tex_dim=0, filt=1, vm=None, tu3=None, tu3_dim=0, fc=None):
    """ShadeFragment of one quad: (colour[4][4], mask[4])."""
    L = olib()
    vp = ctypes.c_void_p
    L.rglo_shade_ex.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, vp, vp, vp, vp]
    keep, ptrs = _buffers(buffers)
    arrs = {}
    for name, a in (("u", uniforms), ("tex", tex), ("vm", vm), ("tu3", tu3), ("fc", fc)):
        arrs[name] = None if a is None else np.ascontiguousarray(a, np.float32)
    ptr = {k: (None if a is None else a.ctypes.data) for k, a in arrs.items()}
    vv = np.zeros((KMAXVARY, 4), np.float32)
    vary = np.asarray(vary, np.float32)
    vv[:len(vary)] = vary
    out = np.zeros(16, np.float32)
    mask = np.zeros(4, np.int32)
    rc = L.rglo_shade_ex(pid, ptr["u"], ptrs, ptr["tex"], tex_dim, filt, ptr["vm"], ptr["tu3"], tu3_dim, ptr["fc"],
                         vv.ctypes.data, out.ctypes.data, mask.ctypes.data)
    assert rc >= 0
    return out.reshape(4, 4), mask


def shade_vertex(pid, idx, buffers, vpm, vm, uniforms=None, iid=0):
    """ShadeVertex: (clip position[4], varyings[16])."""
    L = olib()
    vp = ctypes.c_void_p
    L.rglo_shade_vertex.argtypes = [ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp]
    keep, ptrs = _buffers(buffers)
    u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float32)
    vpm = np.ascontiguousarray(vpm, np.float32)
    vm = np.ascontiguousarray(vm, np.float32)
    pos = np.zeros(4, np.float32)
    vary = np.zeros(KMAXVARY, np.float32)
    rc = L.rglo_shade_vertex(pid, None if u is None else u.ctypes.data, ptrs, vpm.ctypes.data, vm.ctypes.data, iid, idx,
                             pos.ctypes.data, vary.ctypes.data)
    assert rc >= 0
    return pos, vary


def _rgb(c):
    return (c >> 16) & 255, (c >> 8) & 255, c & 255


# ---------------------------------------------------------------------------
# alpha discard (AlphaTexture, Text)
# ---------------------------------------------------------------------------
def _sphere_over_ground(program, alpha, colours=None, name="x"):
    sc, pro, _ = S.base(name=name)
    t = scenes.value_noise_rgba8(128, 41)
    t[..., 3] = alpha
    sc.arrays["t"] = scenes.make_mipmap_stack(t)
    if colours is not None:
        n = len(sc.arrays["sphere.0"])
        for k in range(3):
            sc.arrays[f"sphere.{6 + k}"] = np.full(n, colours[k], np.float32)
    body = S.lambert(sc) + S.draw(sc, "ground") + [("disable", GL_CULL_FACE)]
    if program is not None:
        body += [("use_program", program), ("bind_texture", 0, "@t", 128, 128, 128, scenes.GL_LINEAR_MIPMAP_NEAREST)] + \
            S.draw(sc, "sphere")
    sc.calls = pro + [("clear", CB | DB)] + body + S.stores()
    sc.outputs = S.OUT3
    return oracle_calls(sc, threads=4)[0]


def test_zero_alpha_discards_every_fragment():
    # AlphaTextureProgram / TextProgram (viewer/shaders.hxx:225, :313):
    # gl_triMask &= alpha > 0 before the late depth test, so a sphere whose
    # texture alpha is 0 everywhere leaves colour, depth and ids untouched
    ground = _sphere_over_ground(None, 255)
    for pid in (scenes.PGM_ALPHA_TEXTURE, scenes.PGM_TEXT):
        got = _sphere_over_ground(pid, 0)
        for a, b in zip(ground, got):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), pid


def test_full_alpha_equals_amy_program():
    # with alpha 1 nothing is discarded and the late depth test of an opaque
    # LESS draw keeps what AmyProgram's early test keeps; TextProgram with
    # white vertex colours multiplies by 1
    amy = _sphere_over_ground(scenes.PGM_AMY, 255)
    assert (amy[2] != _sphere_over_ground(None, 255)[2]).sum() > 5000
    for pid, col in ((scenes.PGM_ALPHA_TEXTURE, None), (scenes.PGM_TEXT, None), (scenes.PGM_TEXT, (1.0, 1.0, 1.0))):
        got = _sphere_over_ground(pid, 255, col)
        for a, b in zip(amy, got):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (pid, col)


def test_text_vertex_colour_channels():
    # TextProgram's colour slots 6, 7, 8 scale red, green, blue
    amy = _sphere_over_ground(scenes.PGM_AMY, 255)
    sphere = amy[2] != _sphere_over_ground(None, 255)[2]
    ar, ag, ab = _rgb(amy[0])
    for col in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)):
        r, g, b = _rgb(_sphere_over_ground(scenes.PGM_TEXT, 255, col)[0])
        for ch, want, c in ((r, ar, col[0]), (g, ag, col[1]), (b, ab, col[2])):
            assert np.array_equal(ch[sphere], want[sphere] if c else np.zeros_like(want[sphere])), col


def test_alpha_mask_follows_texel_alpha():
    # NEAREST lookups of a 4 x 4 texture whose alpha is 0 in its top-right
    # 2 x 2 texels; lanes 0 and 1 one texel apart (LOD 0)
    t = np.full((4, 4, 4), 200, np.uint8)
    t[0:2, 2:4, 3] = 0  # rows 0-1 (top), columns 2-3
    tex = scenes.make_mipmap_stack(t)
    u = np.array([0.375, 0.625, 0.375, 0.625], np.float32)
    v = np.array([0.875, 0.875, 0.125, 0.125], np.float32)  # v = 0.875 addresses row 0
    for pid, vary in ((scenes.PGM_ALPHA_TEXTURE, [u, v]), (scenes.PGM_TEXT, [np.ones(4), np.ones(4), np.ones(4), u, v])):
        col, mask = shade_ex(pid, vary, tex=tex, tex_dim=4, filt=scenes.GL_NEAREST_MIPMAP_NEAREST)
        assert mask.tolist() == [-1, 0, -1, -1], pid
        assert col[3].tolist() == [np.float32(200) / np.float32(255), 0.0] + [np.float32(200) / np.float32(255)] * 2


# ---------------------------------------------------------------------------
# the depth texture unit (DepthProgram)
# ---------------------------------------------------------------------------
def test_depth_program_reads_texel_and_border():
    # DepthTextureUnit::sample (rglr_texture_sampler.hxx:70-79): texel
    # (column floor(x dim), row floor((1 - y) dim)), -1 outside [0, 1)^2;
    # DepthProgram linearises with zNear 10, zFar 1000
    dim = 16
    dmap = (np.arange(dim * dim, dtype=np.float32) / np.float32(dim * dim)).astype(np.float32)
    rng = np.random.default_rng(5)
    for _ in range(100):
        x = rng.uniform(-0.3, 1.3, 4).astype(np.float32)
        y = rng.uniform(-0.3, 1.3, 4).astype(np.float32)
        col, _ = shade_ex(scenes.PGM_DEPTH, [x, y], tu3=dmap, tu3_dim=dim)
        for l in range(4):
            t = float(col[0, l])
            assert col[1, l] == t and col[2, l] == t and col[3, l] == 1.0
            inside = 0 <= x[l] < 1 and 0 <= y[l] < 1
            d = (1010.0 - 20.0 / t) / 990.0  # invert 2 n / ((f + n) - d (f - n))
            if not inside:
                assert col[0, l] == np.float32(20) / np.float32(2000)
                continue
            ex, ey = float(x[l]) * dim, (1.0 - float(y[l])) * dim
            if min(abs(ex - round(ex)), abs(ey - round(ey))) < 1e-3:
                continue
            want = (int(ey) * dim + int(ex)) / (dim * dim)
            assert abs(d - want) < 1e-4, (x[l], y[l], d, want)


# ---------------------------------------------------------------------------
# gl_FragCoord through PatternProgram
# ---------------------------------------------------------------------------
def _index_srgb_texture():
    # texel (column x, row y) stores sRGB8 codes (x, y, x ^ y) decoded to float
    x = np.arange(256)
    t = np.zeros((256, 256, 4), np.uint8)
    t[..., 0] = x[None, :]
    t[..., 1] = x[:, None]
    t[..., 2] = x[None, :] ^ x[:, None]
    t[..., 3] = 255
    return scenes.make_mipmap_stack(t)


def _pattern_frame(positions, view, proj, cull=False):
    # a 256 x 256 target; PatternProgram with dim.y 256, offset 0, NEAREST
    arrays = {"tex": _index_srgb_texture(), "u": P._pattern_uniforms(256.0, 0.0, 0.0)}
    pos = np.asarray(positions, np.float32)
    for k in range(3):
        arrays[f"q.{k}"] = scenes.pad4(pos[:, k])
    arrays["q.idx"] = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    sc = scenes.CallScene(name="pattern_fc", width=256, height=256, arrays=arrays)
    m = scenes.Mesh({k: arrays[f"q.{k}"] for k in range(3)}, arrays["q.idx"])
    sc.calls = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_RGBF32), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_F32),
                ("clear_color", 0.0, 0.0, 0.0), ("clear_depth", 1.0), ("view_matrix", view), ("projection_matrix", proj),
                ("disable", GL_CULL_FACE), ("clear", CB | DB), ("use_program", scenes.PGM_PATTERN), ("alloc_uniforms", "@u"),
                ("bind_texture", 0, "@tex", 256, 256, 256, scenes.GL_NEAREST_MIPMAP_NEAREST)] + \
        scenes.mesh_calls("q", m) + S.stores()
    sc.outputs = S.OUT3
    outs, st = oracle_calls(sc, threads=4)
    return outs, st


def _srgb_roundtrip_ok():
    L = olib()
    tab = scenes.srgb8_to_float_table()
    return all(L.rglo_to_srgb8(float(tab[k])) == k for k in range(256))


def test_fragcoord_of_the_vector_rasteriser():
    # VTriangleRasterizer (rglv_triangle.hxx:297): gl_FragCoord = (x + 0.5,
    # H - y - 0.5) for pixel column x, row y from the top; with the texture
    # spanning 256 pixels PatternProgram shows texel (x, y) at pixel (x, y)
    assert _srgb_roundtrip_ok()
    eye = scenes.mat4_rows(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)
    (col, depth, prim), st = _pattern_frame([(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)], eye, eye)
    assert st["clipped"] == 0
    r, g, b = _rgb(col)
    ys, xs = np.mgrid[0:256, 0:256]
    assert (prim != 0xFFFFFFFF).all()
    assert np.array_equal(r, xs) and np.array_equal(g, ys) and np.array_equal(b, xs ^ ys)


def test_fragcoord_of_the_clipping_rasteriser():
    # TriangleRasterizer, which draws clipped triangles (rglv_triangle.hxx:
    # 160): gl_FragCoord.y = H - y + 0.5, one row above the vector
    # rasteriser's, so pixel (x, y) shows texel (x, y - 1)
    eye = scenes.mat4_rows(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)
    proj = scenes.perspective2(90.0, 1.0, 0.1, 100.0)
    (col, depth, prim), st = _pattern_frame([(-50, -1, 5), (50, -1, 5), (50, -1, -60), (-50, -1, -60)], eye, proj)
    assert st["clipped"] == 2
    drawn = prim != 0xFFFFFFFF
    assert drawn.sum() > 20000
    r, g, b = _rgb(col)
    ys, xs = np.mgrid[0:256, 0:256]
    row = (ys - 1) % 256
    assert np.array_equal(r[drawn], xs[drawn]) and np.array_equal(g[drawn], row[drawn])
    assert np.array_equal(b[drawn], (xs ^ row)[drawn])


# ---------------------------------------------------------------------------
# OBJ1 / OBJ2 / OBJ2S lighting against float64
# ---------------------------------------------------------------------------
def _diffuse64(p, n):
    # lightPos (0, 0, 0, 1) - p as the qfloat4 difference forms it: w is
    # 1 - p.z (rmlv_soa.hxx:163); |l| over four components
    l = np.array([-p[0], -p[1], -p[2], 1.0 - p[2]])
    dist = np.sqrt(l @ l)
    return max(float(np.dot(n, l / dist)), 0.1) * (100.0 / dist)


def _view(rng):
    eye = rng.uniform(-5, 5, 3)
    at = rng.uniform(-1, 1, 3)
    return scenes.look_at(eye, at, (0, 1, 0))


def test_obj2_fragment_diffuse():
    rng = np.random.default_rng(31)
    for _ in range(50):
        sp = np.c_[rng.uniform(-4, 4, (4, 2)), rng.uniform(-20, -1, 4), np.ones(4)].astype(np.float32)
        sn = rng.normal(size=(4, 3))
        sn /= np.linalg.norm(sn, axis=1, keepdims=True)
        sn = np.c_[sn, np.zeros(4)].astype(np.float32)
        kd = rng.uniform(0, 1, (4, 3)).astype(np.float32)
        vary = np.concatenate([sp.T, sn.T, kd.T])
        col, _ = shade_ex(scenes.PGM_OBJ2, vary)
        for l in range(4):
            want = kd[l].astype(np.float64) * _diffuse64(sp[l].astype(np.float64), sn[l].astype(np.float64))
            np.testing.assert_allclose(col[:3, l], want, rtol=2e-3)
            assert col[3, l] == 1.0


def _mesh(rng, n):
    pos = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    kd = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    b = {k: scenes.pad4(pos[:, k]) for k in range(3)}
    b.update({3 + k: scenes.pad4(nrm[:, k]) for k in range(3)})
    b.update({6 + k: scenes.pad4(kd[:, k]) for k in range(3)})
    return pos, nrm, kd, b


def test_obj1_vertex_diffuse_and_defaults():
    rng = np.random.default_rng(32)
    pos, nrm, kd, b = _mesh(rng, 16)
    for trial in range(6):
        vm = _view(rng)
        proj = scenes.perspective2(60.0, 1.5, 0.1, 100.0)
        V = vm.reshape(4, 4).T.astype(np.float64)
        vpm = (proj.reshape(4, 4).T.astype(np.float64) @ V).T.reshape(16)
        for i in range(16):
            cp, vary = shade_vertex(scenes.PGM_OBJ1, i, b, vpm, vm)
            p = V @ np.r_[pos[i], 1.0]
            n = V @ np.r_[nrm[i], 0.0]
            np.testing.assert_allclose(vary[:3], kd[i] * _diffuse64(p, n), rtol=2e-3)
            np.testing.assert_allclose(cp, vpm.reshape(4, 4).T @ np.r_[pos[i], 1.0], rtol=1e-4, atol=1e-4)
        # no colour buffers: kd (1, 1, 1); no normals: (0, 0, 1)
        nb = {k: b[k] for k in range(3)}
        _, vary = shade_vertex(scenes.PGM_OBJ1, 3, nb, vpm, vm)
        d = _diffuse64(V @ np.r_[pos[3], 1.0], V @ np.array([0.0, 0.0, 1.0, 0.0]))
        np.testing.assert_allclose(vary[:3], [d, d, d], rtol=2e-3)


def test_obj2s_vertex_normal_matrix_is_inverse_transpose():
    # MakeMatrices (rglv_gpu_impl.hxx:44-51): nm = transpose(inverse(vm));
    # OBJ2S's sn = nm (n, 1), lp = modelToShadow (p, 1)
    rng = np.random.default_rng(33)
    pos, nrm, kd, b = _mesh(rng, 8)
    for trial in range(8):
        vm = _view(rng)
        # a non-rigid view (scale) so that the inverse-transpose matters
        S_ = np.diag([rng.uniform(0.5, 2), rng.uniform(0.5, 2), rng.uniform(0.5, 2), 1.0])
        V = S_ @ vm.reshape(4, 4).T.astype(np.float64)
        vmf = V.T.reshape(16).astype(np.float32)
        shadow = rng.uniform(-1, 1, (4, 4))
        u = np.zeros(32, np.float32)
        u[:16] = shadow.T.reshape(16)
        nm = np.linalg.inv(V.astype(np.float32).astype(np.float64)).T
        for i in range(8):
            _, vary = shade_vertex(scenes.PGM_OBJ2S, i, b, vmf, vmf, uniforms=u)
            np.testing.assert_allclose(vary[0:4], V @ np.r_[pos[i], 1.0], rtol=1e-4, atol=1e-4)
            np.testing.assert_allclose(vary[4:8], nm @ np.r_[nrm[i], 1.0], rtol=1e-3, atol=1e-4)
            np.testing.assert_allclose(vary[8:11], kd[i], rtol=0)
            np.testing.assert_allclose(vary[11:15], u[:16].reshape(4, 4).T.astype(np.float64) @ np.r_[pos[i], 1.0],
                                       rtol=1e-4, atol=1e-4)


def test_obj2s_fragment_spot_and_shadow():
    rng = np.random.default_rng(34)
    dim = 32
    smap = rng.uniform(-1, 1, dim * dim).astype(np.float32)
    checked = shadowed = coned = 0
    for _ in range(200):
        u = np.zeros(32, np.float32)
        u[16:19] = rng.uniform(-3, 3, 3)
        u[19:22] = rng.normal(size=3)
        u[22] = rng.uniform(0.3, 0.95)
        sp = np.c_[rng.uniform(-3, 3, (4, 3)), np.ones(4)].astype(np.float32)
        sn = np.c_[rng.normal(size=(4, 3)), np.zeros(4)].astype(np.float32)
        kd = rng.uniform(0.1, 1, (4, 3)).astype(np.float32)
        lp = np.c_[rng.uniform(-1.2, 1.2, (4, 3)), rng.uniform(0.5, 2, 4)].astype(np.float32)
        vary = np.concatenate([sp.T, sn.T, kd.T, lp.T])
        col, _ = shade_ex(scenes.PGM_OBJ2S, vary, uniforms=u, tu3=smap, tu3_dim=dim)
        for l in range(4):
            stl = u[16:19].astype(np.float64) - sp[l, :3]
            dist = np.linalg.norm(stl)
            stl /= dist
            att = 100.0 / dist
            ldir = u[19:22].astype(np.float64) / np.linalg.norm(u[19:22])
            ang = -stl @ ldir
            if abs(ang - u[22]) < 2e-3:
                continue
            if ang < u[22]:
                att = 0.111
                coned += 1
            n = sn[l, :3].astype(np.float64) / np.linalg.norm(sn[l, :3])
            diff = max(0.0, n @ stl)
            lx, ly, lz = (lp[l, :3].astype(np.float64) / lp[l, 3])
            x, y = lx * 0.5 + 0.5, ly * 0.5 + 0.5
            ex, ey = x * dim, (1 - y) * dim
            if min(abs(ex - round(ex)), abs(ey - round(ey))) < 1e-3:
                continue
            depth = smap[int(ey) * dim + int(ex)] if (0 <= x < 1 and 0 <= y < 1) else -1.0
            if abs(lz - 0.0005 - depth) < 1e-4:
                continue
            if lz - 0.0005 > depth:
                att = 0.111
                shadowed += 1
            want = att * kd[l].astype(np.float64) * diff
            np.testing.assert_allclose(col[:3, l], want, rtol=3e-3, atol=1e-6)
            checked += 1
    assert checked > 600 and shadowed > 100 and coned > 100


# ---------------------------------------------------------------------------
# the scenes the GPU parity tests draw, and programs added at build time
# ---------------------------------------------------------------------------
def test_oracle_draws_every_program_scene():
    scs = [b() for b in P.ALL]
    (smap,), st = oracle_calls(P.scene_shadow_map(), threads=4)
    assert st["fragments"] > 2000 and (smap < 1).mean() > 0.03
    scs += [P.scene_obj2s(smap), P.scene_obj2s(smap, inorder=True)]
    for sc in scs:
        (col, depth, prim), st = oracle_calls(sc, threads=4)
        r, g, b = _rgb(col)
        drawn = prim != 0xFFFFFFFF
        assert drawn.sum() > 20000, sc.name
        # not a flat or saturated image
        assert len(np.unique(col[drawn])) > 50, sc.name
        assert (np.maximum(np.maximum(r, g), b)[drawn] == 255).mean() < 0.05, sc.name


def test_user_programs_header_builds_into_the_library(tmp_path):
    # python -m rsr_amd.build --programs HEADER: the header's programs join
    # the built-in list, so every per-program kernel is generated for them
    import os
    import shutil
    import subprocess
    from rsr_amd import build
    if not os.path.exists(build.HIPCC):
        import pytest
        pytest.skip("hipcc not present")
    here = os.path.dirname(os.path.abspath(__file__))
    out = str(tmp_path / "librgl_user.so")
    build.build(programs=os.path.join(here, "user_programs_example.h"), out=out)
    nm = shutil.which("nm")
    syms = subprocess.run([nm, "-C", out], capture_output=True, text=True, check=True).stdout
    for kern in ("k_vertex<rgl::ProgFlatScaled>", "k_clip<rgl::ProgFlatScaled>", "k_resolve<rgl::ShadeOne<rgl::ProgFlatScaled> >"):
        assert kern in syms, kern


# ---------------------------------------------------------------------------
# the IQ post program (viewer/shaders.hxx:56-67)
# ---------------------------------------------------------------------------
def test_sse_pow_accuracy_and_fixed_points():
    # Fonseca's exp2 / log2 minimax fits (sse_pow.h): relative error of a
    # few 1e-4 against pow; pow(1, y) = exp2(0) = the exp2 fit's constant
    # term, 0.9999252
    L = olib()
    for x in np.geomspace(1e-4, 16.0, 400).astype(np.float32):
        for y in (0.2, 0.45, 0.5, 0.55, 1.0, 2.0):
            got = L.rglo_sse_pow(float(x), y)
            want = float(x) ** float(np.float32(y))
            assert abs(got / want - 1.0) < 3e-4, (x, y, got, want)
    for y in (0.2, 0.45, 3.0):
        assert L.rglo_sse_pow(1.0, y) == np.float32(9.9992520e-1)
    assert L.rglo_sse_pow(0.0, 0.45) < 1e-16        # log2(0) = -127 in the fit
    assert L.rglo_sse_pow(-0.25, 0.5) == L.rglo_sse_pow(0.25, 0.5)  # the sign bit is not read


def test_approx_sin_accuracy_and_symmetry():
    # the wrapped parabola y (3.1 + 3.6 |y|): |error| < 1.1e-3, odd, 2 pi
    # periodic up to the wrap's rounding
    L = olib()
    xs = np.linspace(-60.0, 60.0, 20001).astype(np.float32)
    got = np.array([L.rglo_approx_sin(float(x)) for x in xs])
    assert np.max(np.abs(got - np.sin(xs.astype(np.float64)))) < 1.1e-3
    for x in xs[::97]:
        assert L.rglo_approx_sin(float(-x)) == -L.rglo_approx_sin(float(x))
    assert L.rglo_approx_sin(0.0) == 0.0
    assert abs(L.rglo_approx_sin(math.pi / 2) - 1.0) < 1e-3


def test_fract_truncates_toward_zero():
    L = olib()
    for x, want in ((2.75, 0.75), (-2.75, -0.75), (0.0, 0.0), (12345.5, 0.5)):
        assert L.rglo_fract(x) == want
    assert L.rglo_fract(3e9) == np.float32(3e9) - np.float32(-2147483648.0)  # cvttps2dq overflow


def _iq(c, qx, qy):
    L = olib()
    cin = np.ascontiguousarray(c, np.float32)
    out = np.zeros(3, np.float32)
    L.rglo_iq_post(cin.ctypes.data, qx, qy, out.ctypes.data)
    return out


def test_iq_post_composes_gamma_vignette_and_hash():
    # built from the pinned blocks: pow per channel with (0.45, 0.5, 0.55),
    # vignette 0.2 + 0.8 pow(16 x y (1 - x) (1 - y), 0.2), plus 1/255 of
    # fract(sin(q.x + 13 q.y + k) 43758.5453123) for channel k (in (-1, 1))
    L = olib()
    rng = np.random.default_rng(41)
    for _ in range(300):
        c = rng.uniform(0.0, 1.5, 3).astype(np.float32)
        qx, qy = (float(np.float32(v)) for v in rng.uniform(0.0, 1.0, 2))
        got = _iq(c, qx, qy)
        t = np.float32(np.float32(np.float32(np.float32(16.0) * np.float32(qx)) * np.float32(qy)) * np.float32(1.0 - np.float32(qx)))
        t = np.float32(t * np.float32(np.float32(1.0) - np.float32(qy)))
        v = np.float32(np.float32(0.2) + np.float32(0.8) * np.float32(L.rglo_sse_pow(float(t), 0.2)))
        n = np.float32(np.float32(qx) + np.float32(13.0) * np.float32(qy))
        for k, g in enumerate((0.45, 0.5, 0.55)):
            base = np.float32(np.float32(L.rglo_sse_pow(float(c[k]), g)) * v)
            h = np.float32(L.rglo_fract(float(np.float32(L.rglo_approx_sin(float(n + np.float32(k)))) * np.float32(43758.5453123))))
            assert -1.0 < h < 1.0   # fract truncates toward zero: negative for a negative sine
            want = np.float32(base + np.float32(1.0 / 255.0) * h)
            assert got[k] == want, (c, qx, qy, k)
    # the vignette is 1 (up to the exp2 fit) at the centre and 0.2 at the rim
    ctr = _iq(np.ones(3, np.float32), 0.5, 0.5)
    assert np.all(np.abs(ctr - 1.0) < 5e-3)
    rim = _iq(np.ones(3, np.float32), 0.0, 0.7)
    assert np.all((rim > 0.2 - 1 / 255 - 1e-4) & (rim < 0.2 + 1 / 255 + 1e-4))


def _flat_iq_frame(w, h, tile_blocks, gamma):
    sc = scenes.CallScene(name="iq_flat", width=w, height=h, tile_blocks=tile_blocks)
    sc.calls = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_RGBF32), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_F32),
                ("clear_color", 0.3, 0.5, 0.7), ("clear", CB | DB), ("use_program", scenes.PGM_IQ_POST),
                ("store_color_tc", ("out", 0), gamma)]
    sc.outputs = ["color"]
    return oracle_calls(sc, threads=4)[0][0]


def test_iq_post_frame_canvas_coordinates():
    # FilterTile (rglr_algorithm.hxx:31-62) gives the post program
    # q = ((x + 0.5) / W, (H - y - 0.5) / H) for pixel (x, y), stepped per
    # quad from each tile's corner; on a flat frame the stored colour lies
    # within 1/255 of the hash-free value computed with the exact q
    L = olib()
    w, h = 200, 96
    out = _flat_iq_frame(w, h, (4, 4), False)           # linear store, 32 px tiles
    ys, xs = np.mgrid[0:h, 0:w]
    qx = (xs + 0.5) / w
    qy = (h - ys - 0.5) / h
    vig = 0.2 + 0.8 * (16 * qx * qy * (1 - qx) * (1 - qy)) ** 0.2
    chans = _rgb(out)
    for k, (c, g) in enumerate(((0.3, 0.45), (0.5, 0.5), (0.7, 0.55))):
        c32 = float(np.float32(c))
        base = c32 ** g * vig
        lo = np.floor(np.clip(base * (1 - 1e-3) - 1 / 255, 0, 1) * 255)
        hi = np.floor(np.clip(base * (1 + 1e-3) + 1 / 255, 0, 1) * 255)
        got = chans[k]
        assert np.all((got >= lo) & (got <= hi)), k
        # the hash term spreads over (-1/255, 1/255): pixels land a code
        # below, at and above the hash-free value
        mid = np.floor(base * 255)
        for frac in ((got < mid).mean(), (got == mid).mean(), (got > mid).mean()):
            assert frac > 0.1, (k, frac)


def test_envmap_vertex_sphere_map_uv():
    # EnvmapProgram (shaders_envmap.hxx:100-112): e = unit view position,
    # n = unit (transpose(inverse(vm)) normal), r = e - 2 (n . e) n,
    # uv = r.xy / (2 sqrt(r.x^2 + r.y^2 + (r.z + 1)^2)) + 0.5; the face
    # normal enters with weight 0
    rng = np.random.default_rng(35)
    pos, nrm, face, b = _mesh(rng, 12)
    eye = scenes.mat4_rows(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)
    # the closed form: a point straight ahead with its normal toward the eye
    one = {0: scenes.pad4(np.array([0.0], np.float32)), 1: scenes.pad4(np.array([0.0], np.float32)),
           2: scenes.pad4(np.array([-5.0], np.float32)), 3: scenes.pad4(np.array([0.0], np.float32)),
           4: scenes.pad4(np.array([0.0], np.float32)), 5: scenes.pad4(np.array([1.0], np.float32))}
    _, vary = shade_vertex(scenes.PGM_ENVMAP, 0, one, eye, eye)
    assert vary[0] == 0.5 and vary[1] == 0.5
    for trial in range(6):
        vm = _view(rng)
        V = vm.reshape(4, 4).T.astype(np.float64)
        nm = np.linalg.inv(V).T
        for i in range(12):
            _, vary = shade_vertex(scenes.PGM_ENVMAP, i, b, vm, vm)
            e = (V @ np.r_[pos[i], 1.0])[:3]
            e /= np.linalg.norm(e)
            n = nm[:3, :3] @ nrm[i]
            n /= np.linalg.norm(n)
            r = e - 2 * (n @ e) * n
            m = 2 * np.sqrt(r[0] ** 2 + r[1] ** 2 + (r[2] + 1) ** 2)
            np.testing.assert_allclose(vary[:2], r[:2] / m + 0.5, atol=2e-3)
            # other face normals, same uv
            b2 = dict(b)
            for k in range(3):
                b2[6 + k] = -b[6 + k]
            assert np.array_equal(shade_vertex(scenes.PGM_ENVMAP, i, b2, vm, vm)[1][:2], vary[:2])
