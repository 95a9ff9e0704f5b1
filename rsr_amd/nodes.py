"""The viewer's scene-graph callers of the rgl pipeline, as call lists.

Each class mirrors one node of the reference viewer and issues, in the same
order, the rglv::GL calls its draw code makes:

  GpuNode       $gpu   (viewer/node/gpu.cxx:107-160): Reset, RB_COLOR_DEPTH
                       attachments, clear to the first layer's colour, then
                       every layer's Render for pass 0 and pass 1
  LayerNode     $layer (viewer/node/gllayer.cxx:112-188): camera matrices,
                       spot lights, one shadow-map frame per light on the
                       layer's own GPU (depth only, front faces culled,
                       BaseProgram, StoreDepth), then its gl nodes' Draw
  MeshNode      $mesh  (viewer/node/mesh.cxx:58-112): OBJ2Program, or
                       OBJ2SProgram with the light's shadow map
  ManyNode      $many  (viewer/node/many.cxx:148-230): three instanced
                       batches with ManyProgram's uniform 0.222 / 0.555 / 0.888
  TrueColorNode $truecolor (viewer/node/truecolor.cxx:90-108): the post
                       program, its uniform, StoreColor

`build_frames(gpu_node, truecolor)` returns the frames in the order they
run: the layers' shadow-map frames, then the main frame, as
scenes.CallScene objects that any backend (the HIP library or the oracle)
draws with scenes.render_calls.  An argument "@name" names an array of the
frame; "@frame:k/out:j" is output j of frame k (a shadow map), which the
caller substitutes once that frame has run.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import scenes
from .scenes import (GL_BACK, GL_COLOR_ATTACHMENT0, GL_COLOR_BUFFER_BIT, GL_CULL_FACE, GL_DEPTH_ATTACHMENT,
                     GL_DEPTH_BUFFER_BIT, GL_FRONT, GL_LESS, GL_TRIANGLES, GL_UNSIGNED_SHORT, RB_COLOR_DEPTH, RB_F32,
                     RB_RGBF32)


def mat(m) -> np.ndarray:
    """A column-major flat float32 mat4 as a 4x4 (row, column) array."""
    return np.asarray(m, np.float64).reshape(4, 4).T


def flat(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, np.float64).T.reshape(16)).astype(np.float32)


@dataclass
class Camera:
    eye: tuple
    at: tuple
    fovy: float = 60.0
    near: float = 0.1
    far: float = 100.0

    def view(self):
        return mat(scenes.look_at(self.eye, self.at, (0, 1, 0)))

    def projection(self, aspect):
        return mat(scenes.perspective2(self.fovy, aspect, self.near, self.far))


@dataclass
class SpotLight:
    """A spot light: its model matrix (light to world) from a position and
    a target, cone angle in degrees, shadow-map size."""
    pos: tuple
    at: tuple
    angle: float = 50.0
    size: int = 256

    def model(self):
        return np.linalg.inv(mat(scenes.look_at(self.pos, self.at, (0, 1, 0))))


@dataclass
class LightPack:
    pmat: list = field(default_factory=list)
    vmat: list = field(default_factory=list)
    pos: list = field(default_factory=list)
    dir: list = field(default_factory=list)
    cos: list = field(default_factory=list)
    size: list = field(default_factory=list)
    map: list = field(default_factory=list)     # "@frame:k/out:0" references

    @property
    def cnt(self):
        return len(self.pmat)


class MeshNode:
    """$mesh: one indexed mesh, positions (slots 0-2), normals (3-5), kd
    colours (6-8)."""

    def __init__(self, name, pos, nrm, kd, idx, make_shadows=True, take_shadows=True):
        self.name, self.idx = name, np.asarray(idx, np.uint16)
        self.arrays = {f"{name}.{k}": scenes.pad4(np.asarray(a, np.float32)[:, j])
                       for a, k0 in ((pos, 0), (nrm, 3), (kd, 6)) for j, k in enumerate(range(k0, k0 + 3))}
        self.arrays[f"{name}.idx"] = self.idx
        self.make_shadows, self.take_shadows = make_shadows, take_shadows

    def _buffers(self):
        return [("use_buffer", s, f"@{self.name}.{s}") for k0 in (0, 3, 6) for s in range(k0, k0 + 3)]

    def draw_depth(self, calls, pmat, mvmat):           # mesh.cxx:58-72
        if not self.make_shadows:
            return
        calls += [("view_matrix", flat(mvmat)), ("projection_matrix", flat(pmat))] + self._buffers()
        calls.append(("draw_elements", GL_TRIANGLES, len(self.idx), GL_UNSIGNED_SHORT, f"@{self.name}.idx", 0))

    def draw(self, pass_, lights, calls, pmat, vmat, mmat, arrays):   # mesh.cxx:74-112
        if pass_ != 1:
            return
        if self.take_shadows and lights.cnt > 0:
            calls.append(("use_program", scenes.PGM_OBJ2S))
            u = np.zeros(32, np.float32)
            u[0:16] = flat(lights.pmat[0] @ lights.vmat[0] @ mmat)
            u[16:19] = lights.pos[0]
            u[19:22] = lights.dir[0]
            u[22] = lights.cos[0]
            key = f"{self.name}.uniforms"
            arrays[key] = u
            calls += [("alloc_uniforms", "@" + key), ("bind_texture3", lights.map[0], lights.size[0])]
        else:
            calls.append(("use_program", scenes.PGM_OBJ2))
        calls += [("view_matrix", flat(vmat @ mmat)), ("projection_matrix", flat(pmat))] + self._buffers()
        calls.append(("draw_elements", GL_TRIANGLES, len(self.idx), GL_UNSIGNED_SHORT, f"@{self.name}.idx", 0))


class ManyNode:
    """$many: a mesh (positions, normals, uv) drawn in three instanced
    batches of `batch` instances, instance matrices in slot 15."""
    MAGIC = (0.222, 0.555, 0.888)

    def __init__(self, name, pos, nrm, uv, idx, mats, program=scenes.PGM_MANY):
        self.name, self.idx, self.program = name, np.asarray(idx, np.uint16), program
        mats = np.asarray(mats, np.float32).reshape(-1, 16)
        assert len(mats) % 3 == 0
        self.batch = len(mats) // 3
        self.arrays = {f"{name}.{k}": scenes.pad4(np.asarray(a, np.float32)[:, j])
                       for a, k0 in ((pos, 0), (nrm, 3)) for j, k in enumerate(range(k0, k0 + 3))}
        self.arrays[f"{name}.9"] = scenes.pad4(np.asarray(uv, np.float32)[:, 0])
        self.arrays[f"{name}.10"] = scenes.pad4(np.asarray(uv, np.float32)[:, 1])
        for b in range(3):   # mats_.data() + b * batch
            self.arrays[f"{name}.mats{b}"] = np.ascontiguousarray(mats[b * self.batch:(b + 1) * self.batch].reshape(-1))
        self.arrays[f"{name}.idx"] = self.idx

    def _batches(self, calls, arrays, which):
        for b in which:
            key = f"{self.name}.magic{b}"
            arrays[key] = np.array([self.MAGIC[b]] + [0.0] * 31, np.float32)
            calls.append(("alloc_uniforms", "@" + key))
            calls += [("use_buffer", s, f"@{self.name}.{s}") for s in (0, 1, 2, 3, 4, 5, 9, 10)]
            calls.append(("use_buffer", 15, f"@{self.name}.mats{b}"))
            calls.append(("draw_elements_instanced", GL_TRIANGLES, len(self.idx), GL_UNSIGNED_SHORT, f"@{self.name}.idx",
                          self.batch))

    def draw_depth(self, calls, pmat, mvmat, arrays):   # many.cxx:148-164: the first batch
        calls += [("projection_matrix", flat(pmat)), ("view_matrix", flat(mvmat))]
        self._batches(calls, arrays, (0,))

    def draw(self, pass_, lights, calls, pmat, vmat, mmat, arrays):   # many.cxx:188-230
        if pass_ != 1:
            return
        calls += [("use_program", self.program), ("projection_matrix", flat(pmat)), ("view_matrix", flat(vmat @ mmat))]
        self._batches(calls, arrays, (0, 1, 2))


class LayerNode:
    """$layer: a camera, spot lights and gl nodes ($mesh, $many)."""

    def __init__(self, color, camera, gls, lights=()):
        self.color, self.camera, self.gls, self.lights = color, camera, list(gls), list(lights)

    def render(self, pass_, calls, arrays, target, aspect, frames):   # gllayer.cxx:112-188
        pmat = self.camera.projection(aspect) if self.camera else np.eye(4)
        vmat = self.camera.view() if self.camera else np.eye(4)
        mmat = np.eye(4)
        lights = LightPack()
        for li, L in enumerate(self.lights):
            light_model = L.model()
            light_to_view = vmat @ light_model
            view_to_light = np.linalg.inv(light_to_view)
            lights.pmat.append(mat(scenes.perspective2(L.angle, 1.0, 20.0, 500.0)))
            lights.vmat.append(np.linalg.inv(light_model))
            lights.pos.append(light_to_view[:3, 3].astype(np.float32))
            lights.dir.append((-view_to_light[2, :3]).astype(np.float32))
            lights.cos.append(np.float32(math.cos(L.angle / 2.0 * 3.141592 / 180)))
            lights.size.append(L.size)
            # the light's depth frame on the layer's own GPU
            sc = scenes.CallScene(name=f"shadow{len(frames)}_light{li}_pass{pass_}", width=L.size, height=L.size,
                                  tile_blocks=(8, 8))
            sc.calls = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_RGBF32), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_F32),
                        ("color_write_mask", False), ("depth_write_mask", True), ("depth_func", GL_LESS),
                        ("enable", GL_CULL_FACE), ("cull_face", GL_FRONT), ("clear_depth", 1.0),
                        ("clear", GL_DEPTH_BUFFER_BIT), ("use_program", scenes.PGM_BASE)]
            for gl in self.gls:
                if isinstance(gl, ManyNode):
                    gl.draw_depth(sc.calls, lights.pmat[li], lights.vmat[li], sc.arrays)
                else:
                    gl.draw_depth(sc.calls, lights.pmat[li], lights.vmat[li])
                sc.arrays.update(gl.arrays)
            sc.calls.append(("store_depth", ("out", 0)))
            sc.outputs = ["depth"]
            frames.append(sc)
            lights.map.append(f"@frame:{len(frames) - 1}/out:0")
        for gl in self.gls:
            gl.draw(pass_, lights, calls, pmat, vmat, mmat, arrays)
            arrays.update(gl.arrays)


class GpuNode:
    """$gpu: target size, tile size in blocks, layers."""

    def __init__(self, size=(256, 256), tile_blocks=(8, 8), layers=()):
        self.size, self.tile_blocks, self.layers = tuple(size), tuple(tile_blocks), list(layers)


class TrueColorNode:
    """$truecolor: post program (1 default, 2 exposure, 3 IQ), its uniform
    0, and sRGB encoding."""

    def __init__(self, program=scenes.PGM_DEFAULT_POST, uf0=None, srgb=True):
        self.program, self.uf0, self.srgb = program, uf0, srgb


def build_frames(gpu: GpuNode, tc: TrueColorNode, extra_stores: bool = True):
    """The frames of one viewer frame, in running order: every shadow-map
    frame, then the main frame (outputs: the truecolor canvas, and, with
    extra_stores, the depth and primitive ids stored after it)."""
    w, h = gpu.size
    aspect = float(w) / h
    main = scenes.CallScene(name=f"viewer_{w}x{h}", width=w, height=h, tile_blocks=gpu.tile_blocks)
    frames = []
    color = gpu.layers[0].color if gpu.layers else (0.0, 0.0, 0.0)
    calls = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_COLOR_DEPTH), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_COLOR_DEPTH),
             ("clear_color", *color), ("clear_depth", 1.0), ("clear", GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT)]
    for pass_ in (0, 1):
        for layer in gpu.layers:
            layer.render(pass_, calls, main.arrays, gpu.size, aspect, frames)
    calls.append(("use_program", tc.program))
    if tc.uf0 is not None:
        main.arrays["truecolor.uf0"] = np.array([tc.uf0] + [0.0] * 31, np.float32)
        calls.append(("alloc_uniforms", "@truecolor.uf0"))
    calls.append(("store_color_tc", ("out", 0), tc.srgb))
    main.outputs = ["color"]
    if extra_stores:
        calls += [("store_depth", ("out", 1)), ("store_primitive_id", ("out", 2))]
        main.outputs += ["depth", "prim"]
    main.calls = calls
    frames.append(main)
    return frames


def run_frames(frames, draw):
    """Draws frames in order with draw(scene) -> outputs, substituting each
    "@frame:k/out:j" argument by that frame's output (an array of the
    backend); returns every frame's outputs."""
    outs = []
    for sc in frames:
        for i, c in enumerate(sc.calls):
            args = list(c)
            for j, a in enumerate(args):
                if isinstance(a, str) and a.startswith("@frame:"):
                    k, o = a[len("@frame:"):].split("/out:")
                    key = f"frame{k}.out{o}"
                    sc.arrays[key] = outs[int(k)][int(o)]
                    args[j] = "@" + key
            sc.calls[i] = tuple(args)
        outs.append(draw(sc))
    return outs
