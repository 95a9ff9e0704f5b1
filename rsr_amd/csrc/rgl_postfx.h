// rgl_postfx.h — the float-canvas post effect of the rgl pipeline on
// gfx950: the Kawase blur pass (rglr_kawase.cxx:21-85).
#pragma once

#include "rgl_device.h"

namespace rgl {

// One Kawase pass over a W x H RGBA float canvas (row strides in pixels):
// each output pixel is the mean of four 2 x 2 boxes whose corners sit
// (d + 1) pixels up-left, up-right, down-left and down-right of it, with
// coordinates clamped to the canvas (KawaseBlurFilter's sample_safe; its
// unclamped interior loop reads the same texels).  Sums in the reference's
// order: the box ((a + b) + c) + d, the boxes accumulated from zero, then
// times 1/16.  One thread per output pixel, 64 x 4 per workgroup so that a
// wave reads whole rows.
__global__ __launch_bounds__(256) void k_kawase(const float4* __restrict__ src, float4* __restrict__ dst, int W, int H,
                                                int srcStride, int dstStride, int d) {
	const int x = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x & 63);
	const int y = static_cast<int>(blockIdx.y) * 4 + static_cast<int>(threadIdx.x >> 6);
	if (x >= W || y >= H) return;
	auto at = [&](int px, int py) {
		px = px < 0 ? 0 : px > W - 1 ? W - 1 : px;
		py = py < 0 ? 0 : py > H - 1 ? H - 1 : py;
		return src[static_cast<long long>(py) * srcStride + px]; };
	const int ox[4] = { -d - 1, d, -d - 1, d }, oy[4] = { -d - 1, -d - 1, d, d };
	float4 ax = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		const float4 a = at(x + ox[k], y + oy[k]), b = at(x + ox[k] + 1, y + oy[k]);
		const float4 c = at(x + ox[k], y + oy[k] + 1), e = at(x + ox[k] + 1, y + oy[k] + 1);
		ax.x = ax.x + (((a.x + b.x) + c.x) + e.x);
		ax.y = ax.y + (((a.y + b.y) + c.y) + e.y);
		ax.z = ax.z + (((a.z + b.z) + c.z) + e.z);
		ax.w = ax.w + (((a.w + b.w) + c.w) + e.w); }
	const float k16 = 1.0F / 16.0F;
	dst[static_cast<long long>(y) * dstStride + x] = make_float4(ax.x * k16, ax.y * k16, ax.z * k16, ax.w * k16); }

}  // namespace rgl
