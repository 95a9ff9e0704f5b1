// rgl_assets.h — texture preparation of the asset path on gfx950: decoded
// RGBA8 images to float texels (rglr_texture_load.cxx:39-66), the
// power-of-two resample (rglr_texture.cxx:13-31) and the mip chain
// (rglr_texture.cxx:34-81).
#pragma once

#include "rgl_device.h"

namespace rgl {

// ryg::srgb8_to_float's 256 values (ryg-srgb.h:33-56), made once on the host
struct Srgb8Table { float v[256]; };

// PNGToTexture's texel loop: R, G, B through the sRGB table, A = a / 255,
// RGB times A when premultiplying.  One thread per texel.
__global__ __launch_bounds__(256) void k_texture_from_rgba8(const uchar4* __restrict__ src, float4* __restrict__ dst,
                                                             long long n, int premultiply, Srgb8Table tab) {
	const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uchar4 p = src[i];
	float4 o = make_float4(tab.v[p.x], tab.v[p.y], tab.v[p.z], static_cast<float>(p.w) / 255.0F);
	if (premultiply) { o.x = o.x * o.w; o.y = o.y * o.w; o.z = o.z * o.w; }
	dst[i] = o; }

// ensurePowerOf2: a dim x dim nearest resample of a w x h texture, source
// row / column int(i / (dim - 1) * (size - 1)) in double precision.
__global__ __launch_bounds__(256) void k_texture_pow2(const float4* __restrict__ src, int w, int h, float4* __restrict__ dst,
                                                       int dim) {
	const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= static_cast<long long>(dim) * dim) return;
	const int iy = static_cast<int>(i / dim), ix = static_cast<int>(i - static_cast<long long>(iy) * dim);
	const int sy = static_cast<int>(static_cast<double>(iy) / static_cast<double>(dim - 1) * static_cast<double>(h - 1));
	const int sx = static_cast<int>(static_cast<double>(ix) / static_cast<double>(dim - 1) * static_cast<double>(w - 1));
	dst[i] = src[static_cast<long long>(sy) * w + sx]; }

// one level of Texture::maybe_make_mipmap: texel (x, y) of the level of
// size `size` (rows from `dstRow`) is the 2 x 2 mean of the level above
// (rows from `srcRow`), summed ((a + b) + c) + d and divided by 4; rows of
// `stride` texels
__global__ __launch_bounds__(256) void k_mip_level(float4* __restrict__ buf, int stride, int srcRow, int dstRow, int size) {
	const int i = static_cast<int>(blockIdx.x) * blockDim.x + static_cast<int>(threadIdx.x);
	if (i >= size * size) return;
	const int y = i / size, x = i - y * size;
	const float4* r1 = buf + static_cast<long long>(srcRow + 2 * y) * stride + 2 * x;
	const float4* r2 = r1 + stride;
	const float4 a = r1[0], b = r1[1], c = r2[0], d = r2[1];
	buf[static_cast<long long>(dstRow + y) * stride + x] = make_float4(
		(((a.x + b.x) + c.x) + d.x) / 4.0F, (((a.y + b.y) + c.y) + d.y) / 4.0F,
		(((a.z + b.z) + c.z) + d.z) / 4.0F, (((a.w + b.w) + c.w) + d.w) / 4.0F); }

}  // namespace rgl
