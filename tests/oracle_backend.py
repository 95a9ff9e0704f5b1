"""ctypes binding of the CPU oracle (oracle/librgl_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg use it, always as the checker.  Method names mirror the
rglv::GPU / rglv::GL surface so render_scene() drives it exactly like the
HIP pipeline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# RGL_ORACLE_SO points the tests at another build of the oracle (the
# mutation check in oracle/mutation_check.py uses it)
ORACLE_SO = os.environ.get("RGL_ORACLE_SO", os.path.join(ORACLE_DIR, "librgl_oracle.so"))

_lib = None


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        vp, i, f, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint32
        L.rglo_gpu_create.restype = vp
        L.rglo_gpu_create.argtypes = [i]
        L.rglo_gpu_destroy.argtypes = [vp]
        L.rglo_gpu_set_threads.argtypes = [vp, i]
        L.rglo_gpu_set_guard_band.argtypes = [vp, i]
        L.rglo_gpu_last_error.restype = ctypes.c_char_p
        L.rglo_gpu_last_error.argtypes = [vp]
        L.rglo_gpu_install.argtypes = [vp, i, u32, i]
        L.rglo_gpu_reset.argtypes = [vp, i, i, i, i]
        L.rglo_gpu_run.argtypes = [vp]
        L.rglo_gpu_stats.argtypes = [vp, vp]
        for name in ("enable", "disable", "depth_func", "depth_write_mask", "color_write_mask",
                     "cull_face", "use_program", "use_uniforms", "clear"):
            getattr(L, "rglo_gl_" + name).argtypes = [vp, i]
        L.rglo_gl_scissor.argtypes = [vp, i, i, i, i]
        L.rglo_gl_viewport.argtypes = [vp, i, i, i, i]
        L.rglo_gl_alloc_uniform_buffer.argtypes = [vp, vp, i]
        L.rglo_gl_alloc_uniform_buffer.restype = i
        L.rglo_gl_renderbuffer_type.argtypes = [vp, i, i]
        L.rglo_gl_bind_texture.argtypes = [vp, i, vp, i, i, i, i]
        L.rglo_gl_use_buffer.argtypes = [vp, i, vp]
        L.rglo_gl_bind_texture3.argtypes = [vp, vp, i]
        L.rglo_gl_clear_color.argtypes = [vp, f, f, f]
        L.rglo_gl_clear_depth.argtypes = [vp, f]
        for name in ("view_matrix", "projection_matrix", "normal_matrix"):
            getattr(L, "rglo_gl_" + name).argtypes = [vp, vp]
        L.rglo_gl_draw_elements.argtypes = [vp, i, i, i, vp, i]
        L.rglo_gl_draw_elements_instanced.argtypes = [vp, i, i, i, vp, i]
        L.rglo_gl_draw_arrays.argtypes = [vp, i, i, i]
        L.rglo_gl_draw_arrays_instanced.argtypes = [vp, i, i, i, i]
        L.rglo_gl_store_color_tc.argtypes = [vp, vp, i]
        L.rglo_gl_store_depth.argtypes = [vp, vp]
        L.rglo_gl_store_primitive_id.argtypes = [vp, vp]
        L.rglo_gl_store_color_fp.argtypes = [vp, vp, i, i]
        L.rglo_gl_store_color_quads.argtypes = [vp, vp, i]
        L.rglo_kawase_blur.argtypes = [vp, vp, i, i, i, i, i]
        L.rglo_gl_reset.argtypes = [vp]
        L.rglo_oneover.argtypes = [f]
        L.rglo_oneover.restype = f
        L.rglo_rcpps.argtypes = [f]
        L.rglo_rcpps.restype = f
        L.rglo_rsqrtps.argtypes = [f]
        L.rglo_rsqrtps.restype = f
        for name in ("sse_pow", "approx_sin", "fract"):
            fn = getattr(L, "rglo_" + name)
            fn.argtypes = [f] * (2 if name == "sse_pow" else 1)
            fn.restype = f
        L.rglo_iq_post.argtypes = [vp, f, f, vp]
        L.rglo_cvtt.argtypes = [f]
        L.rglo_cvtt.restype = ctypes.c_int32
        L.rglo_to_srgb8.argtypes = [f]
        L.rglo_to_srgb8.restype = u32
        L.rglo_to_linear8.argtypes = [f]
        L.rglo_to_linear8.restype = u32
        L.rglo_mat4_mul.argtypes = [vp, vp, vp]
        L.rglo_xform4.argtypes = [vp, vp, vp]
        L.rglo_scalar_triangle.argtypes = [i, i, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return a


class OracleGPU:
    """The oracle behind the rglv::GPU / rglv::GL method names."""

    def __init__(self, threads: int = 1):
        self.L = lib()
        self.h = self.L.rglo_gpu_create(threads)
        self._keep = []

    def close(self):
        if self.h:
            self.L.rglo_gpu_destroy(self.h)
            self.h = None

    __del__ = close

    def set_threads(self, n: int):
        self.L.rglo_gpu_set_threads(self.h, n)

    def set_guard_band(self, fixed: bool):
        self.L.rglo_gpu_set_guard_band(self.h, int(fixed))

    def check(self, rc: int):
        if rc != 0:
            raise RuntimeError("oracle: " + self.L.rglo_gpu_last_error(self.h).decode())

    def install(self, program_id, state_key, kind):
        self.check(self.L.rglo_gpu_install(self.h, program_id, state_key, kind))

    def reset(self, w, h, tbx=8, tby=8):
        self._keep = []
        self.check(self.L.rglo_gpu_reset(self.h, w, h, tbx, tby))

    def run(self):
        self.check(self.L.rglo_gpu_run(self.h))

    def stats(self):
        out = np.zeros(5, np.int64)
        self.L.rglo_gpu_stats(self.h, out.ctypes.data)
        return dict(submitted=int(out[0]), clipped=int(out[1]), tile_pairs=int(out[2]), fragments=int(out[3]),
                    clipped_tile_pairs=int(out[4]))

    def __getattr__(self, name):
        if not name.startswith("gl_"):
            raise AttributeError(name)
        fn = getattr(self.L, "rglo_" + name)

        def call(*args):
            conv = []
            for a in args:
                if isinstance(a, np.ndarray):
                    self._keep.append(a)
                    conv.append(_p(a))
                elif isinstance(a, bool):
                    conv.append(int(a))
                else:
                    conv.append(a)
            return fn(self.h, *conv)
        return call

    def gl_use_buffer(self, slot, ptr, count=0):
        if isinstance(ptr, np.ndarray):
            self._keep.append(ptr)
            ptr = ptr.ctypes.data
        self.L.rglo_gl_use_buffer(self.h, slot, ptr)

    def gl_alloc_uniform_buffer(self, values):
        v = np.zeros(32, np.float32)
        if values is not None:
            v[:len(values)] = values
        return self.L.rglo_gl_alloc_uniform_buffer(self.h, v.ctypes.data, 32)


def render(scene, threads: int = 1, prim: bool = True):
    """Render a Scene with the oracle; returns (color u32, depth f32, prim u32)."""
    from rsr_amd.scenes import render_scene
    g = OracleGPU(threads)
    color = np.zeros((scene.height, scene.width), np.uint32)
    depth = np.zeros((scene.height, scene.width), np.float32)
    primid = np.zeros((scene.height, scene.width), np.uint32) if prim else None
    render_scene(g, scene, color, depth, primid)
    g.run()
    st = g.stats()
    g.close()
    return color, depth, primid, st


def render_calls(sc, threads: int = 4):
    """Draw a CallScene with the oracle; returns (outputs, stats)."""
    from rsr_amd.scenes import render_calls as issue
    g = OracleGPU(threads)
    outs = []
    from rsr_amd.scenes import output_array
    for kind in sc.outputs:
        outs.append(output_array(kind, sc.width, sc.height))
    issue(g, sc, outs)
    g.run()
    st = g.stats()
    g.close()
    return outs, st
