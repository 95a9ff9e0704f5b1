/*
 * rgl_hip.h — C ABI of the MI355X (gfx950) rgl rasteriser pipeline.
 *
 * Each entry point replaces one member of the reference's rglv::GPU /
 * rglv::GL surface (usrlocalben/rsr, src/rgl/rglv); the reference
 * declaration it stands for is cited beside it.  Plain pointers and sizes
 * only.  Pointers named dev_* are device (HBM) addresses the caller owns;
 * they must stay valid until the frame that uses them has completed
 * (rgl_gpu_wait), as the reference requires of its borrowed buffers until
 * the next Run() completes (rglv_gpu.hxx:176-192).
 *
 * Every call returns 0 on success and a negative RGL_E_* code on failure;
 * rgl_gpu_last_error() then describes it.  The reference calls
 * std::exit(1) in the same situations (rglv_gpu.cxx:199-202, 300-309).
 */
#ifndef RGL_HIP_H
#define RGL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rgl_gpu rgl_gpu;

enum {
	RGL_OK = 0,
	RGL_E_ARG = -1,          /* invalid argument */
	RGL_E_HIP = -2,          /* a HIP runtime call failed */
	RGL_E_DISPATCH = -3,     /* no installed program for a state key */
	RGL_E_UNSUPPORTED = -4,  /* state this pipeline does not implement */
	RGL_E_CAPACITY = -5,     /* a per-frame limit was exceeded */
	RGL_E_STATE = -6         /* call out of order */
};

/* install kinds (the three Install overloads, rglv_gpu.hxx:154-156) */
enum { RGL_BIN_PROGRAM = 0, RGL_DRAW_PROGRAM = 1, RGL_BLT_PROGRAM = 2 };

/* pipeline counters (GPUStats, rglv_gpu.hxx:51-62, plus fragment and bin counts) */
typedef struct rgl_stats {
	int64_t tiles;
	int64_t commands;
	int64_t primitives_submitted;
	int64_t primitives_clipped;
	int64_t primitives_binned;   /* triangles accepted into at least one tile */
	int64_t tile_pairs;          /* (triangle, tile) pairs in the tile lists */
	int64_t clipped_tile_pairs;  /* (clipped triangle, tile) pairs */
	int64_t fragments;           /* covered pixels that reached the depth test */
	int64_t overflow;            /* nonzero: a capacity was exceeded, frame invalid */
	int64_t fallback_quads;      /* learned shading: tile quadrants re-shaded by the plain shader (4 per tile) */
} rgl_stats;

/* --- rglv::GPU ---------------------------------------------------------- */

/* GPU(int concurrency, std::string guid)   rglv_gpu.hxx:153.
   device_ordinal selects the HIP device; concurrency is accepted for API
   compatibility (the device schedules its own tiles). */
int rgl_gpu_create(int device_ordinal, int concurrency, const char* guid, rgl_gpu** out);
int rgl_gpu_destroy(rgl_gpu* gpu);

/* Install(programId, stateKey, Bin/Draw/BltProgramPtrs)   rglv_gpu.hxx:154-156,
   keyed (programId << 24) | stateKey as in rglv_gpu.cxx:26-47. */
int rgl_gpu_install(rgl_gpu* gpu, int program_id, uint32_t state_key, int kind);

/* Reset(ivec2 bufferDimensionsInPixels, ivec2 tileDimensionsInBlocks)
   rglv_gpu.hxx:159; tiles are tile_blocks * 8 pixels (rglv_gpu.hxx:27).
   Also resets the recording context (GL::Reset, rglv_gl.cxx:440-445). */
int rgl_gpu_reset(rgl_gpu* gpu, int width, int height, int tile_blocks_x, int tile_blocks_y);

/* Run()   rglv_gpu.hxx:167-168: bin and draw everything recorded since
   Reset, asynchronously on hip_stream (a hipStream_t; NULL = the default
   stream).  Completion is observed with rgl_gpu_wait. */
int rgl_gpu_run(rgl_gpu* gpu, void* hip_stream);

/* Blocks until the last Run has completed and checks its device counters
   (double-buffered: every Run in flight; the stats are the last one's). */
int rgl_gpu_wait(rgl_gpu* gpu);

/* rglv::doubleBuffer   rglv_gpu.cxx:16,96-116, rglv_gpu.hxx:176-192: with
   enable = 1, frame N's binning runs beside frame N-1's drawing.  The
   reference flag is global to every GPU and lags the image by one Run;
   here it is per GPU and Run draws its own frame: consecutive Runs go
   round two executors (enable = 1; 2 to 4: that many frames in flight),
   each with its own per-frame buffers and a stream of the GPU's own.  A
   frame starts after the work queued on hip_stream before its Run; its
   first kernel that writes or samples caller memory waits for the previous
   frame's end, so frames store in submission order.  The caller's stream
   does not wait for the frame: read the outputs after rgl_gpu_wait.  What a
   frame's vertex stage reads must not be written by the frame before it.
   Reset does not wait for frames in flight.  A per-frame capacity overflow
   is reported by the next rgl_gpu_wait (the capacities are grown; the
   frame must be recorded and run again).  Learned shading, training
   samples, timing, graph replay and batches run single-buffered
   (RGL_E_UNSUPPORTED).  Switching waits for every frame in flight.
   enable = 0 returns to single-buffered runs. */
int rgl_gpu_set_double_buffer(rgl_gpu* gpu, int enable);

/* Double-buffered: blocks until the frame before the last Run has
   completed (and with it every earlier frame: each frame's stores waited
   for the previous frame's end), while the last frame may still run; the
   stats are that frame's.  The reference's Run returns with the previous
   frame drawn (rglv_gpu.cxx:96-116); a viewer that renders into two
   canvases in turn displays the previous one after this call.
   Single-buffered: rgl_gpu_wait. */
int rgl_gpu_wait_previous(rgl_gpu* gpu);

/* GPUStats of the last completed Run. */
int rgl_gpu_stats(rgl_gpu* gpu, rgl_stats* out);

/* Average duration in milliseconds of each pipeline stage over the runs
   since the last call, when timing is enabled: k_vis and k_resolve by
   start and stop events recorded by their own dispatch (the kernel's
   execution alone), the other stages between events on the run's stream;
   names_out receives a comma-separated stage list.  No counterpart in the
   reference (measurement only). */
int rgl_gpu_set_timing(rgl_gpu* gpu, int enable);
int rgl_gpu_kernel_times(rgl_gpu* gpu, double* ms_out, int max_kernels, char* names_out, int names_len);

const char* rgl_gpu_last_error(rgl_gpu* gpu);

/* --- rglv::GL (rglv_gl.hxx:186-331), recording into GPU::IC() ----------- */
int rgl_gl_enable(rgl_gpu* gpu, int cap);                        /* GL::Enable          :186-197 */
int rgl_gl_disable(rgl_gpu* gpu, int cap);                       /* GL::Disable         :199-210 */
int rgl_gl_depth_func(rgl_gpu* gpu, int func);                   /* GL::DepthFunc       :212-214 */
int rgl_gl_depth_write_mask(rgl_gpu* gpu, int enable);           /* GL::DepthWriteMask  :216-218 */
int rgl_gl_color_write_mask(rgl_gpu* gpu, int enable);           /* GL::ColorWriteMask  :220-222 */
int rgl_gl_cull_face(rgl_gpu* gpu, int face);                    /* GL::CullFace        :224-226 */
int rgl_gl_scissor(rgl_gpu* gpu, int x, int y, int w, int h);    /* GL::Scissor         :228-231 */
int rgl_gl_viewport(rgl_gpu* gpu, int x, int y, int w, int h);   /* GL::Viewport        :233-236 */
int rgl_gl_use_program(rgl_gpu* gpu, int program_id);           /* GL::UseProgram      :238-240 */
int rgl_gl_use_uniforms(rgl_gpu* gpu, int ofs);                  /* GL::UseUniforms     :242-244 */
/* GL::AllocUniformBuffer :253-258 — returns the offset; the 32 floats are
   copied from values (n <= 32), the caller may also rewrite them before Run
   through rgl_gl_uniform_buffer_addr. */
int rgl_gl_alloc_uniform_buffer(rgl_gpu* gpu, const float* values, int n);
float* rgl_gl_uniform_buffer_addr(rgl_gpu* gpu, int ofs);        /* GL::GetUniformBufferAddr :260-263 */
int rgl_gl_renderbuffer_type(rgl_gpu* gpu, int attachment, int type); /* GL::RenderbufferType :246-251 */
/* GL::BindTexture :265-273; dev_texels = float RGBA, mip chain stacked
   below the base level (N x 2N texels for power-of-two N). */
int rgl_gl_bind_texture(rgl_gpu* gpu, int unit, const float* dev_texels, int width, int height, int stride, int mode);
/* GL::BindTexture3 :275-278 — the depth texture unit (DepthTextureUnit,
   rglr_texture_sampler.hxx:33-79): dev_depth = dim x dim floats, dim a
   power of two <= 8192, nearest, border -1 outside [0,1)^2; null unbinds (lookups read 0).  The
   array must not be a store target of the same run. */
int rgl_gl_bind_texture3(rgl_gpu* gpu, const float* dev_depth, int dim);
/* GL::UseBuffer :280-288; count = number of floats in the array (slot 0's
   count is the vertex count of the draws that use it). */
int rgl_gl_use_buffer(rgl_gpu* gpu, int slot, const float* dev_ptr, int count);
int rgl_gl_clear_color(rgl_gpu* gpu, float r, float g, float b);  /* GL::ClearColor  :290-292 */
int rgl_gl_clear_depth(rgl_gpu* gpu, float value);                /* GL::ClearDepth  :294-296 */
int rgl_gl_view_matrix(rgl_gpu* gpu, const float* m16);           /* GL::ViewMatrix       :298-300, column-major */
int rgl_gl_projection_matrix(rgl_gpu* gpu, const float* m16);     /* GL::ProjectionMatrix :302-304 */
int rgl_gl_normal_matrix(rgl_gpu* gpu, const float* m16);         /* GL::NormalMatrix     :306-308 */
/* GL::DrawElements rglv_gl.cxx:34-43; dev_indices = uint16 triangle list */
int rgl_gl_draw_elements(rgl_gpu* gpu, int mode, int count, int type, const uint16_t* dev_indices, int hint);
int rgl_gl_draw_arrays(rgl_gpu* gpu, int mode, int start, int count);                      /* rglv_gl.cxx:13-20 */
int rgl_gl_draw_elements_instanced(rgl_gpu* gpu, int mode, int count, int type,
                                   const uint16_t* dev_indices, int instance_count);       /* rglv_gl.cxx:46-56 */
int rgl_gl_draw_arrays_instanced(rgl_gpu* gpu, int mode, int start, int count, int instance_count); /* :23-31 */
int rgl_gl_clear(rgl_gpu* gpu, int bits);                                                  /* rglv_gl.cxx:58-61 */
/* GL::StoreColor(TrueColorCanvas*, bool) rglv_gl.cxx:79-83: dev_pixels =
   width*height uint32 0x00RRGGBB, row stride = width. */
int rgl_gl_store_color_tc(rgl_gpu* gpu, uint32_t* dev_pixels, int enable_gamma);
/* GL::StoreDepth rglv_gl.cxx:86-89: width*height float, row stride = width. */
int rgl_gl_store_depth(rgl_gpu* gpu, float* dev_depth);
/* GL::StoreColor(FloatingPointCanvas*, bool downsample) rglv_gl.cxx:64-70
   (CMD_STORE_COLOR_FULL_LINEAR_FP / _HALF_LINEAR_FP, rglv_gpu.cxx:345-370):
   dev_rgba = RGBA float pixels (alpha 0), row stride stride_px pixels (0:
   width, or (width + 1) / 2 downsampled).  downsample = the 2 x 2 box mean
   into pixel (x / 2, y / 2) (rglr_algorithm.cxx:118-165). */
int rgl_gl_store_color_fp(rgl_gpu* gpu, float* dev_rgba, int stride_px, int downsample);
/* GL::StoreColor(QFloat4Canvas*) rglv_gl.cxx:73-77 (CMD_STORE_COLOR_FULL_QUADS_FP,
   rglv_gpu.cxx:371-384): the quad-swizzled canvas, quad (x / 2, y / 2) at
   index (y / 2) * stride_quads + x / 2 (0: (width + 1) / 2), 16 floats
   r[4] g[4] b[4] a[4], lane = (y & 1) * 2 + (x & 1); alpha 1 from
   RB_RGBF32, the fourth channel otherwise (RB_COLOR_DEPTH: the depth). */
int rgl_gl_store_color_quads(rgl_gpu* gpu, float* dev_quads, int stride_quads);
/* No counterpart in the reference: stores, per pixel, the id of the
   triangle that last passed the depth test (instance * triangles per
   instance + triangle index, counted across the frame's draws in
   submission order), 0xffffffff where nothing was drawn. */
int rgl_gl_store_primitive_id(rgl_gpu* gpu, uint32_t* dev_ids);
int rgl_gl_reset(rgl_gpu* gpu);                                                            /* GL::Reset rglv_gl.cxx:92-97 */

/* --- post effects -------------------------------------------------------- */
/* rglr::KawaseBlurFilter rglr_kawase.cxx:21-85 over a whole canvas: one
   blur pass of distance d from dev_src into dev_dst (RGBA float pixels,
   width x height, row strides in pixels; src and dst must differ), queued
   on hip_stream. */
int rgl_kawase_blur(const float* dev_src, float* dev_dst, int width, int height, int src_stride, int dst_stride,
                    int d, void* hip_stream);

/* --- texture preparation ------------------------------------------------ */
/* Texture::maybe_make_mipmap (rglr_texture.cxx:33-81) on host memory:
   texels = dim*dim*2 RGBA floats with the base level in the first dim rows;
   fills the box-filtered mip chain below it. */
int rgl_texture_make_mipmap(float* host_texels, int dim);
/* PNGToTexture's texel loop (rglr_texture_load.cxx:39-66) on the device:
   dev_rgba8 = width*height decoded RGBA8 texels -> dev_rgba = RGBA floats,
   R, G, B through ryg::srgb8_to_float (ryg-srgb.h:33-56), A = a / 255,
   RGB times A when premultiply; queued on hip_stream. */
int rgl_texture_from_rgba8(const uint8_t* dev_rgba8, int width, int height, int premultiply, float* dev_rgba,
                           void* hip_stream);
/* rglr::ensurePowerOf2 (rglr_texture.cxx:13-31) on the device: the dim x
   dim nearest resample (dim = the power of two at or above the long edge)
   of a width x height RGBA float texture. */
int rgl_texture_pow2(const float* dev_src, int width, int height, float* dev_dst, int dim, void* hip_stream);
/* Texture::maybe_make_mipmap on the device: dev_texels = dim*dim*2 RGBA
   floats, base level in the first dim rows; one launch per level. */
int rgl_texture_mipmap(float* dev_texels, int dim, void* hip_stream);

/* --- batches ----------------------------------------------------------- */
/* No counterpart in the reference, which runs one GPU object per canvas
   (several at once for shadow passes, node/gllayer.cxx:152-181).  Draws
   the frames recorded in frames[0..n) (each its own rgl_gpu recorder with
   its own Reset size, tile size, state and stores; 1 <= n <= 64) with ONE
   launch sequence on gpu's buffers and hip_stream: triangles, tiles and
   resolve blocks are numbered across the batch.  Every frame's outputs
   are the ones it would get alone.  gpu may be one of the frames.  Wait
   and stats on gpu cover the whole batch. */
int rgl_gpu_run_batch(rgl_gpu* gpu, rgl_gpu* const* frames, int n, void* hip_stream);

/* enable != 0: the launch sequence of a run is captured once into a HIP
   graph and replayed while the batch's plan (draws, frames, buffers,
   capacities) is unchanged; runs with kernel timing on are never captured. */
int rgl_gpu_set_graph(rgl_gpu* gpu, int enable);
int rgl_gpu_graph_captures(rgl_gpu* gpu);   /* graphs captured so far */

/* Guard band: mode 0 (default) = the reference's CalcGuardBandFactor
   (rglv_view_frustum.hxx:36-39), which drops below 1 for targets wider
   or taller than 2048 pixels, so that on-screen triangles beyond +-factor
   NDC are culled or clipped as in the reference; mode 1 = the factor raised
   to at least 1 (a documented deviation for large targets). */
int rgl_gpu_set_guard_band(rgl_gpu* gpu, int mode);

/* The autotuner's knobs for how many workgroups share a dense tile:
   seg_pairs = (triangle, tile) pairs per visibility work item (default
   512); split_quads = quads dealt out per 32 x 32 quadrant (measured on
   the previous frame) above which the quadrant is cut in two, and above
   twice that in four (default 0 = never).  Every setting draws identical
   frames. */
int rgl_gpu_set_raster_params(rgl_gpu* gpu, int seg_pairs, int split_quads);

/* --- learned shading --------------------------------------------------- */
/* No counterpart in the reference (its shaders are analytic).  A small MLP
   (13 shading inputs, 4-band frequency encoding to 117 values, 3 hidden
   layers of 64, ReLU, sRGB colour out; BF16 weights and activations, FP32
   accumulation on MFMA) stands in for the heavy GGX + environment program
   (id 82).  Parameters are an fp32 block of rgl_mlp_param_count() floats
   on the device: W1[64][128] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
   W4[3][64] b4[3], then 4 normalisation floats (position centre xyz,
   1 / half-extent) that training leaves alone. */
int rgl_mlp_param_count(void);
int rgl_mlp_trainable_count(void);

/* Frames drawn with program 82 shade with the network from the next Run on:
   the plain shader checks 64 probe pixels per tile and re-shades every
   tile whose RMS error (8-bit, over its covered probes) exceeds
   max_rms_err (in
   [0,1] units, e.g. 2/255); max_rms_err >= 1 checks nothing and keeps the
   network's colours everywhere.  Coverage, depth and primitive ids are
   never affected.  dev_params = NULL switches back to the plain shader. */
int rgl_gpu_set_learned_shading(rgl_gpu* gpu, const float* dev_params, float max_rms_err);

/* Training data: covered pixels of the frame, each kept with probability
   keep (a hash of pixel and seed), appended at dev_out[atomic(*dev_count)]
   as 16 floats: 13 shading inputs (normal, view vector, world position,
   albedo, roughness) and the plain shader's sRGB colour / 255.  Frames must
   use program 82. */
int rgl_gl_store_training_samples(rgl_gpu* gpu, float* dev_out, int capacity, uint32_t* dev_count, float keep, uint32_t seed);

/* Network output (sRGB-space RGB, before clamping) for n samples in the
   16-float layout above: dev_out = n x 3 floats. */
int rgl_mlp_forward(rgl_gpu* gpu, const float* dev_samples, int n, const float* dev_params, float* dev_out, void* hip_stream);

/* Gradient of the mean squared error over the first n samples (all 3
   channels) with respect to the trainable parameters -> dev_grad
   (rgl_mlp_trainable_count() floats); *loss_out (host) = the loss. */
int rgl_mlp_gradient(rgl_gpu* gpu, const float* dev_samples, int n, float* dev_params, float* dev_grad, float* loss_out,
                     void* hip_stream);

/* steps Adam steps (beta 0.9 / 0.99) on batches of `batch` samples drawn by
   hash from the n samples; dev_opt = 2 x rgl_mlp_trainable_count() floats of
   optimiser state (zero before the first step); step0 = steps already taken.
   *loss_out (host, may be NULL) = the last batch's loss. */
int rgl_mlp_train(rgl_gpu* gpu, const float* dev_samples, long long n, float* dev_params, float* dev_opt, int steps, int batch,
                  float lr, uint32_t seed, int step0, float* loss_out, void* hip_stream);

/* --- no allocation ------------------------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving or training loops,
   harnesses that forbid hipMalloc): rgl_mlp_train with every training buffer
   in one file-scope device array of 8.6 MB that each GPU holds until the
   process ends, so the call makes no hipMalloc, hipFree or other allocation.
   batch <= 4096 (RGL_E_CAPACITY beyond).  The last batch's loss goes to the
   device float *dev_loss_out (may be NULL) instead of host memory.  All
   calls on one GPU share that array: calls that overlap on one GPU (other
   streams, other rgl_gpu objects) race. */
int rgl_mlp_train_noalloc(rgl_gpu* gpu, const float* dev_samples, long long n, float* dev_params, float* dev_opt, int steps,
                          int batch, float lr, uint32_t seed, int step0, float* dev_loss_out, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* RGL_HIP_H */
