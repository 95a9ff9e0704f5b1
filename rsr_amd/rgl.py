"""Python binding of the gfx950 rgl pipeline (librgl_hip.so) — the
rglv::GPU / rglv::GL surface over the C ABI in include/rgl_hip.h.

Device memory and streams come from PyTorch (plumbing only); every
pipeline stage runs in the HIP library.  There is no CPU fallback: loading
fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# RGL_HIP_LIB names another build of the library (for comparing builds)
LIB_PATH = os.environ.get("RGL_HIP_LIB") or os.path.join(PKG, "librgl_hip.so")

RGL_E_CAPACITY = -5


class RglError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "tiles", "commands", "primitives_submitted", "primitives_clipped", "primitives_binned",
        "tile_pairs", "clipped_tile_pairs", "fragments", "overflow", "fallback_quads")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RglError(f"{LIB_PATH} is missing: build it with `python -m rsr_amd.build` "
                           "(the HIP pipeline has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, f, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint32
        L.rgl_gpu_create.argtypes = [i, i, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.rgl_gpu_destroy.argtypes = [vp]
        L.rgl_gpu_install.argtypes = [vp, i, u32, i]
        L.rgl_gpu_reset.argtypes = [vp, i, i, i, i]
        L.rgl_gpu_run.argtypes = [vp, vp]
        L.rgl_gpu_wait.argtypes = [vp]
        L.rgl_gpu_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.rgl_gpu_set_timing.argtypes = [vp, i]
        L.rgl_gpu_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i, ctypes.c_char_p, i]
        L.rgl_gpu_last_error.argtypes = [vp]
        L.rgl_gpu_last_error.restype = ctypes.c_char_p
        for n in ("enable", "disable", "depth_func", "depth_write_mask", "color_write_mask",
                  "cull_face", "use_program", "use_uniforms", "clear"):
            getattr(L, "rgl_gl_" + n).argtypes = [vp, i]
        L.rgl_gl_scissor.argtypes = [vp, i, i, i, i]
        L.rgl_gl_viewport.argtypes = [vp, i, i, i, i]
        L.rgl_gl_alloc_uniform_buffer.argtypes = [vp, vp, i]
        L.rgl_gl_uniform_buffer_addr.argtypes = [vp, i]
        L.rgl_gl_uniform_buffer_addr.restype = vp
        L.rgl_gl_renderbuffer_type.argtypes = [vp, i, i]
        L.rgl_gl_bind_texture.argtypes = [vp, i, vp, i, i, i, i]
        L.rgl_gl_use_buffer.argtypes = [vp, i, vp, i]
        L.rgl_gl_bind_texture3.argtypes = [vp, vp, i]
        L.rgl_gl_clear_color.argtypes = [vp, f, f, f]
        L.rgl_gl_clear_depth.argtypes = [vp, f]
        for n in ("view_matrix", "projection_matrix", "normal_matrix"):
            getattr(L, "rgl_gl_" + n).argtypes = [vp, vp]
        L.rgl_gl_draw_elements.argtypes = [vp, i, i, i, vp, i]
        L.rgl_gl_draw_arrays.argtypes = [vp, i, i, i]
        L.rgl_gl_draw_elements_instanced.argtypes = [vp, i, i, i, vp, i]
        L.rgl_gl_draw_arrays_instanced.argtypes = [vp, i, i, i, i]
        L.rgl_gl_store_color_tc.argtypes = [vp, vp, i]
        L.rgl_gl_store_depth.argtypes = [vp, vp]
        L.rgl_gl_store_primitive_id.argtypes = [vp, vp]
        L.rgl_gl_store_color_fp.argtypes = [vp, vp, i, i]
        L.rgl_gl_store_color_quads.argtypes = [vp, vp, i]
        L.rgl_kawase_blur.argtypes = [vp, vp, i, i, i, i, i, vp]
        L.rgl_gl_reset.argtypes = [vp]
        L.rgl_texture_make_mipmap.argtypes = [vp, i]
        L.rgl_texture_from_rgba8.argtypes = [vp, i, i, i, vp, vp]
        L.rgl_texture_pow2.argtypes = [vp, i, i, vp, i, vp]
        L.rgl_texture_mipmap.argtypes = [vp, i, vp]
        L.rgl_gpu_run_batch.argtypes = [vp, ctypes.POINTER(vp), i, vp]
        L.rgl_gpu_set_graph.argtypes = [vp, i]
        L.rgl_gpu_set_double_buffer.argtypes = [vp, i]
        L.rgl_gpu_wait_previous.argtypes = [vp]
        L.rgl_gpu_graph_captures.argtypes = [vp]
        L.rgl_gpu_set_raster_params.argtypes = [vp, i, i]
        L.rgl_gpu_set_guard_band.argtypes = [vp, i]
        L.rgl_mlp_param_count.argtypes = []
        L.rgl_mlp_trainable_count.argtypes = []
        L.rgl_gpu_set_learned_shading.argtypes = [vp, vp, f]
        L.rgl_gl_store_training_samples.argtypes = [vp, vp, i, vp, f, u32]
        L.rgl_mlp_forward.argtypes = [vp, vp, i, vp, vp, vp]
        L.rgl_mlp_gradient.argtypes = [vp, vp, i, vp, vp, ctypes.POINTER(f), vp]
        L.rgl_mlp_train.argtypes = [vp, vp, ctypes.c_longlong, vp, vp, i, i, f, u32, i, ctypes.POINTER(f), vp]
        L.rgl_mlp_train_noalloc.argtypes = [vp, vp, ctypes.c_longlong, vp, vp, i, i, f, u32, i, vp, vp]
        _lib = L
    return _lib


# every symbol include/rgl_hip.h declares
EXPORTS = [
    "rgl_gpu_create", "rgl_gpu_destroy", "rgl_gpu_install", "rgl_gpu_reset", "rgl_gpu_run",
    "rgl_gpu_wait", "rgl_gpu_stats", "rgl_gpu_set_timing", "rgl_gpu_kernel_times", "rgl_gpu_last_error",
    "rgl_gl_enable", "rgl_gl_disable", "rgl_gl_depth_func", "rgl_gl_depth_write_mask",
    "rgl_gl_color_write_mask", "rgl_gl_cull_face", "rgl_gl_scissor", "rgl_gl_viewport",
    "rgl_gl_use_program", "rgl_gl_use_uniforms", "rgl_gl_alloc_uniform_buffer",
    "rgl_gl_uniform_buffer_addr", "rgl_gl_renderbuffer_type", "rgl_gl_bind_texture", "rgl_gl_bind_texture3",
    "rgl_gl_use_buffer", "rgl_gl_clear_color", "rgl_gl_clear_depth", "rgl_gl_view_matrix",
    "rgl_gl_projection_matrix", "rgl_gl_normal_matrix", "rgl_gl_draw_elements",
    "rgl_gl_draw_arrays", "rgl_gl_draw_elements_instanced", "rgl_gl_draw_arrays_instanced",
    "rgl_gl_clear", "rgl_gl_store_color_tc", "rgl_gl_store_depth", "rgl_gl_store_primitive_id",
    "rgl_gl_store_color_fp", "rgl_gl_store_color_quads", "rgl_kawase_blur",
    "rgl_gl_reset", "rgl_texture_make_mipmap", "rgl_texture_from_rgba8", "rgl_texture_pow2", "rgl_texture_mipmap",
    "rgl_gpu_run_batch", "rgl_gpu_set_graph", "rgl_gpu_set_double_buffer", "rgl_gpu_wait_previous", "rgl_gpu_graph_captures", "rgl_gpu_set_raster_params",
    "rgl_gpu_set_guard_band",
    "rgl_mlp_param_count", "rgl_mlp_trainable_count", "rgl_gpu_set_learned_shading",
    "rgl_gl_store_training_samples", "rgl_mlp_forward", "rgl_mlp_gradient", "rgl_mlp_train",
    "rgl_mlp_train_noalloc",
]


def texture_from_rgba8(rgba8, premultiply: bool = False, mip: bool = True):
    """A decoded RGBA8 image (H x W x 4 uint8, host or CUDA tensor or numpy)
    to the pipeline's texture on the device, as rglr::LoadPNG,
    ensurePowerOf2 and maybe_make_mipmap prepare it: returns (texels, dim)
    with texels a (2 dim) x dim x 4 float32 CUDA tensor (base level on top,
    the mip chain below) when mip, else the dim x dim level alone."""
    import torch
    src = torch.as_tensor(np.asarray(rgba8) if isinstance(rgba8, np.ndarray) else rgba8)
    if not src.is_cuda:
        src = src.cuda()
    src = src.contiguous()
    assert src.dtype == torch.uint8 and src.dim() == 3 and src.shape[2] == 4
    h, w = int(src.shape[0]), int(src.shape[1])
    s = torch.cuda.current_stream(src.device).cuda_stream
    L = lib()
    f = torch.empty((h, w, 4), dtype=torch.float32, device=src.device)
    if L.rgl_texture_from_rgba8(src.data_ptr(), w, h, int(premultiply), f.data_ptr(), s) != 0:
        raise RglError("rgl_texture_from_rgba8 failed")
    dim = 1 << max(1, (max(w, h) - 1).bit_length())
    if w == h and w == dim:
        sq = f
    else:
        sq = torch.empty((dim, dim, 4), dtype=torch.float32, device=src.device)
        if L.rgl_texture_pow2(f.data_ptr(), w, h, sq.data_ptr(), dim, s) != 0:
            raise RglError("rgl_texture_pow2 failed")
    if not mip:
        return sq, dim
    out = torch.zeros((2 * dim, dim, 4), dtype=torch.float32, device=src.device)
    out[:dim] = sq
    if L.rgl_texture_mipmap(out.data_ptr(), dim, s) != 0:
        raise RglError("rgl_texture_mipmap failed")
    return out, dim


def kawase_blur(src, dst, d: int, stream=None) -> None:
    """One Kawase pass (rglr_kawase.cxx:21-85) from src into dst: (H, W, 4)
    float32 CUDA tensors, rows contiguous."""
    import torch
    assert src.shape == dst.shape and src.dim() == 3 and src.shape[2] == 4 and src.dtype == torch.float32
    h, w = int(src.shape[0]), int(src.shape[1])
    s = stream if stream is not None else torch.cuda.current_stream(src.device).cuda_stream
    rc = lib().rgl_kawase_blur(src.data_ptr(), dst.data_ptr(), w, h, int(src.stride(0) // 4), int(dst.stride(0) // 4),
                               int(d), s)
    if rc != 0:
        raise RglError(f"rgl_kawase_blur failed ({rc})")


def make_mipmap(texels: np.ndarray, dim: int) -> np.ndarray:
    """Texture::maybe_make_mipmap on a host float RGBA N x 2N buffer (in place)."""
    assert texels.dtype == np.float32 and texels.flags["C_CONTIGUOUS"] and texels.size >= dim * dim * 8
    rc = lib().rgl_texture_make_mipmap(texels.ctypes.data, dim)
    if rc != 0:
        raise RglError("rgl_texture_make_mipmap failed")
    return texels


class GPU:
    """rglv::GPU on one HIP device (rglv_gpu.hxx:97-276)."""

    def __init__(self, device: int = 0, concurrency: int = 1, guid: str = "unnamed"):
        self.L = lib()
        h = ctypes.c_void_p()
        rc = self.L.rgl_gpu_create(device, concurrency, guid.encode(), ctypes.byref(h))
        if rc != 0:
            raise RglError(f"rgl_gpu_create failed ({rc})")
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.L.rgl_gpu_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc < 0:
            raise RglError(self.L.rgl_gpu_last_error(self.h).decode() + f" ({rc})")
        return rc

    def install(self, program_id, state_key, kind):
        self._check(self.L.rgl_gpu_install(self.h, program_id, state_key, kind))

    def reset(self, w, h, tbx=8, tby=8):
        self._keep = []
        self._check(self.L.rgl_gpu_reset(self.h, w, h, tbx, tby))

    def run(self, stream=None):
        self._check(self.L.rgl_gpu_run(self.h, stream))

    def wait(self):
        return self._check(self.L.rgl_gpu_wait(self.h))

    def run_and_wait(self, stream=None, retries: int = 3):
        """Run; when a per-frame capacity overflowed the library has grown it,
        so the frame is drawn again."""
        for _ in range(retries + 1):
            self.run(stream)
            rc = self.L.rgl_gpu_wait(self.h)
            if rc != RGL_E_CAPACITY:
                return self._check(rc)
        raise RglError("capacity overflow persisted")

    def stats(self) -> dict:
        s = Stats()
        self._check(self.L.rgl_gpu_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def set_timing(self, enable: bool):
        self._check(self.L.rgl_gpu_set_timing(self.h, int(enable)))

    def kernel_times(self) -> dict:
        ms = (ctypes.c_double * 16)()
        names = ctypes.create_string_buffer(512)
        n = self._check(self.L.rgl_gpu_kernel_times(self.h, ms, 16, names, 512))
        return dict(zip(names.value.decode().split(","), [ms[i] for i in range(n)]))

    # --- batches (no reference counterpart) -----------------------------------
    def run_batch(self, frames, stream=None):
        """Draws the frames recorded in `frames` (GPU recorders, may include
        self) with one launch sequence on this GPU's buffers."""
        arr = (ctypes.c_void_p * len(frames))(*[f.h for f in frames])
        self._check(self.L.rgl_gpu_run_batch(self.h, arr, len(frames), stream))

    def run_batch_and_wait(self, frames, stream=None, retries: int = 3):
        for _ in range(retries + 1):
            self.run_batch(frames, stream)
            rc = self.L.rgl_gpu_wait(self.h)
            if rc != RGL_E_CAPACITY:
                return self._check(rc)
        raise RglError("capacity overflow persisted")

    def set_graph(self, enable: bool):
        self._check(self.L.rgl_gpu_set_graph(self.h, int(enable)))

    def set_double_buffer(self, enable: bool):
        """rglv::doubleBuffer (rglv_gpu.cxx:16): frame N's binning beside
        frame N-1's drawing; read the outputs after wait()."""
        self._check(self.L.rgl_gpu_set_double_buffer(self.h, int(enable)))

    def wait_previous(self):
        """Double-buffered: wait for the frame before the last Run."""
        return self._check(self.L.rgl_gpu_wait_previous(self.h))

    def graph_captures(self) -> int:
        return self._check(self.L.rgl_gpu_graph_captures(self.h))

    def set_guard_band(self, fixed: bool):
        self._check(self.L.rgl_gpu_set_guard_band(self.h, int(fixed)))

    def set_raster_params(self, seg_pairs: int, split_quads: int = 0):
        self._check(self.L.rgl_gpu_set_raster_params(self.h, seg_pairs, split_quads))

    # --- learned shading (no reference counterpart) ---------------------------
    def set_learned_shading(self, dev_params, max_rms_err: float = 2.0 / 255.0):
        self._check(self.L.rgl_gpu_set_learned_shading(self.h, dev_params, max_rms_err))

    def gl_store_training_samples(self, dev_out, capacity, dev_count, keep, seed):
        self._check(self.L.rgl_gl_store_training_samples(self.h, dev_out, capacity, dev_count, keep, seed))

    def mlp_forward(self, dev_samples, n, dev_params, dev_out, stream=None):
        self._check(self.L.rgl_mlp_forward(self.h, dev_samples, n, dev_params, dev_out, stream))

    def mlp_gradient(self, dev_samples, n, dev_params, dev_grad, stream=None) -> float:
        loss = ctypes.c_float()
        self._check(self.L.rgl_mlp_gradient(self.h, dev_samples, n, dev_params, dev_grad, ctypes.byref(loss), stream))
        return float(loss.value)

    def mlp_train(self, dev_samples, n, dev_params, dev_opt, steps, batch, lr, seed, step0, stream=None) -> float:
        loss = ctypes.c_float()
        self._check(self.L.rgl_mlp_train(self.h, dev_samples, n, dev_params, dev_opt, steps, batch, lr, seed, step0,
                                         ctypes.byref(loss), stream))
        return float(loss.value)

    def mlp_train_noalloc(self, dev_samples, n, dev_params, dev_opt, steps, batch, lr, seed, step0, dev_loss=None,
                          stream=None):
        """rgl_mlp_train without allocation: the loss goes to a device float."""
        self._check(self.L.rgl_mlp_train_noalloc(self.h, dev_samples, n, dev_params, dev_opt, steps, batch, lr, seed, step0,
                                                 dev_loss, stream))

    # --- rglv::GL -----------------------------------------------------------
    def gl_alloc_uniform_buffer(self, values=None):
        v = np.zeros(32, np.float32)
        if values is not None:
            v[:len(values)] = values
        return self._check(self.L.rgl_gl_alloc_uniform_buffer(self.h, v.ctypes.data, 32))

    def gl_use_buffer(self, slot, dev_ptr, count=0):
        return self._check(self.L.rgl_gl_use_buffer(self.h, slot, dev_ptr, count if dev_ptr else 0))

    def __getattr__(self, name):
        if not name.startswith("gl_"):
            raise AttributeError(name)
        fn = getattr(self.L, "rgl_" + name)

        def call(*args):
            conv = []
            for a in args:
                if isinstance(a, np.ndarray):  # host data copied by the call (matrices)
                    a = np.ascontiguousarray(a, dtype=np.float32)
                    self._keep.append(a)
                    conv.append(a.ctypes.data)
                elif isinstance(a, bool):
                    conv.append(int(a))
                else:
                    conv.append(a)
            return self._check(fn(self.h, *conv))
        return call


class DeviceScene:
    """A Scene's arrays resident in HBM (torch tensors) plus output targets."""

    def __init__(self, scene, device="cuda:0", prim: bool = True):
        import torch
        from .scenes import scene_arrays
        self.scene = scene
        self.tensors = {}
        for a in scene_arrays(scene):
            t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
            self.tensors[id(a)] = t
        self.ptrs = {k: t.data_ptr() for k, t in self.tensors.items()}
        h, w = scene.height, scene.width
        self.color = torch.zeros((h, w), dtype=torch.int32, device=device)
        self.depth = torch.zeros((h, w), dtype=torch.float32, device=device)
        self.prim = torch.zeros((h, w), dtype=torch.int32, device=device) if prim else None

    def record(self, gpu: GPU):
        from .scenes import render_scene
        render_scene(gpu, self.scene, self.color.data_ptr(), self.depth.data_ptr(),
                     self.prim.data_ptr() if self.prim is not None else None, buffers=self.ptrs)

    def outputs(self):
        c = self.color.cpu().numpy().view(np.uint32)
        d = self.depth.cpu().numpy()
        p = self.prim.cpu().numpy().view(np.uint32) if self.prim is not None else None
        return c, d, p


def render(scene, gpu: GPU | None = None, device="cuda:0", prim: bool = True):
    """Draw a Scene on the GPU; returns (color u32, depth f32, prim u32, stats)."""
    import torch
    own = gpu is None
    if own:
        gpu = GPU(int(str(device).split(":")[-1]) if ":" in str(device) else 0)
    ds = DeviceScene(scene, device, prim)
    ds.record(gpu)
    stream = torch.cuda.current_stream(device).cuda_stream
    gpu.run_and_wait(stream)
    out = ds.outputs()
    st = gpu.stats()
    if own:
        gpu.close()
    return (*out, st)


def render_calls(sc, gpu: GPU | None = None, device="cuda:0"):
    """Draw a CallScene on the GPU; returns (outputs as numpy arrays, stats)."""
    import torch
    from .scenes import render_calls as issue
    own = gpu is None
    if own:
        gpu = GPU(int(str(device).split(":")[-1]) if ":" in str(device) else 0)
    dev = {k: torch.from_numpy(np.ascontiguousarray(a)).to(device) for k, a in sc.arrays.items()}
    ptr_of = {id(sc.arrays[k]): t.data_ptr() for k, t in dev.items()}
    from .scenes import output_array
    # outputs start as a pattern the frame must overwrite
    host = [output_array(kind, sc.width, sc.height) for kind in sc.outputs]
    outs = [torch.from_numpy(h.view(np.int32) if h.dtype == np.uint32 else h).to(device) for h in host]
    issue(gpu, sc, [o.data_ptr() for o in outs], ptr=lambda a: ptr_of[id(a)])
    stream = torch.cuda.current_stream(device).cuda_stream
    gpu.run_and_wait(stream)
    res = [o.cpu().numpy().view(h.dtype) for o, h in zip(outs, host)]
    st = gpu.stats()
    if own:
        gpu.close()
    return res, st
