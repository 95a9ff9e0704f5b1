// rgl_kernels.h — the gfx950 kernels of one rgl frame.
//
// Frame = vertex -> setup/count -> scan -> ordered scatter -> clip ->
// clip count/scan/scatter -> per-tile raster + resolve.
//
// Visibility is computed order-independently: for the states this pipeline
// accepts (depth test LESS or LEQUAL with depth writes, no blending,
// early-Z programs) the reference's in-order per-pixel result
// (TriangleProgram::Render, rglv_gpu_impl.hxx:166-222) is the fragment with
// the smallest (depth, submission slot) key — for LESS the first fragment
// reaching the minimum depth, for LEQUAL the last.  Each tile keeps one
// 64-bit key per pixel in LDS and takes ds_min_u64 of every covered
// fragment's key; the resolve then shades only the winning fragment of each
// pixel, recomputing its quad exactly as the reference rasteriser would
// have (same tile-clamped edge start values, same barycentric scale).
#pragma once

#include "rgl_device.h"

namespace rgl {

constexpr int kUnitTris = 256;      // triangles per binning unit (one wave)
constexpr int kRasterThreads = 256;
constexpr int kMaxTileDim = 64;     // LDS key array is kMaxTileDim^2

struct Counters {
	unsigned long long fragments;
	uint32_t clipCount;
	uint32_t pairs;          // inline (triangle, tile) pairs, from the scan
	uint32_t clipPairs;
	uint32_t binned;         // triangles accepted into >= 1 tile
	uint32_t overflow;       // bit 0 list, 1 clip queue, 2 clipped list, 3 bad index
	uint32_t pad[3];
};

// compact per-draw view used by the multi-draw kernels
struct DrawLite {
	long long triBase;
	long long vBase;
	const uint16_t* indices;
	int tris, instances, verts, elements;
	int keepBacks, keepFronts;
	int scissor[4];
	uint32_t slotBase;
	uint32_t primBase;
};

__device__ inline int find_draw_ll(const DrawLite* dl, int nDraws, long long s) {
	int d = 0;
	for (int k = 1; k < nDraws; ++k) if (dl[k].triBase <= s) d = k;
	return d; }

__device__ inline int find_draw_slot(const DrawLite* dl, int nDraws, uint32_t slot) {
	int d = 0;
	for (int k = 1; k < nDraws; ++k) if (dl[k].slotBase <= slot) d = k;
	return d; }

__device__ inline void load_lite(DrawLite* dl, const DrawDesc* draws, int nDraws) {
	for (int k = threadIdx.x; k < nDraws; k += blockDim.x) {
		const DrawDesc& d = draws[k];
		DrawLite l;
		l.triBase = d.triBase; l.vBase = d.vBase; l.indices = d.indices;
		l.tris = d.tris; l.instances = d.instances; l.verts = d.verts; l.elements = d.elements;
		l.keepBacks = d.keepBacks; l.keepFronts = d.keepFronts;
		for (int i = 0; i < 4; ++i) l.scissor[i] = d.scissor[i];
		l.slotBase = d.slotBase; l.primBase = d.primBase;
		dl[k] = l; }}

// ---------------------------------------------------------------------------
// vertex stage: GPUBinImpl::BinTriangles2P vertex loop (rglv_gpu_impl.hxx:364-385)
// once per (instance, vertex) instead of once per tile (DrawTriangles, :880-886)
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(256) void k_vertex(DrawDesc d, float4* __restrict__ vpos,
                                                 uint8_t* __restrict__ vflags, float* __restrict__ vvary) {
	const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	const long long total = static_cast<long long>(d.instances) * d.verts;
	if (i >= total) return;
	const int iid = static_cast<int>(i / d.verts);
	const int v = static_cast<int>(i - static_cast<long long>(iid) * d.verts);
	float pos[4], vo[P::kVary];
	P::vertex(d, iid, v, pos, vo);
	// ViewFrustum::Test, SIMD form with cmple (rglv_view_frustum.hxx:281-294)
	const float w = pos[3] * d.gbFactor;
	uint32_t f = 0;
	if (w + pos[0] <= 0.0F) f |= 1u;
	if (w + pos[1] <= 0.0F) f |= 2u;
	if (pos[3] + pos[2] <= 0.0F) f |= 4u;
	if (w - pos[0] <= 0.0F) f |= 8u;
	if (w - pos[1] <= 0.0F) f |= 16u;
	// pdiv (rglv_math.hxx:19-23), then x * DS + DO
	const float iw = oneover(pos[3]);
	const float dcx = (pos[0] * iw) * d.DS[0] + d.DO[0];
	const float dcy = (pos[1] * iw) * d.DS[1] + d.DO[1];
	const long long g = d.vBase + i;
	vpos[g] = make_float4(dcx, dcy, pos[2] * iw, iw);
	vflags[g] = static_cast<uint8_t>(f);
	float* out = vvary + g * kMaxVary;
	for (int k = 0; k < P::kVary; ++k) out[k] = vo[k]; }

// ---------------------------------------------------------------------------
// triangle setup + per-unit tile counts: the primitive loop of
// BinTriangles1P/2P (rglv_gpu_impl.hxx:566-661) — cull, clip test,
// backface reorder, integer bbox, tile range.
// ---------------------------------------------------------------------------
__device__ inline uint2 pack_rect(int tx0, int ty0, int tx1, int ty1) {
	return make_uint2(static_cast<uint32_t>(tx0) | (static_cast<uint32_t>(tx1) << 16),
	                  static_cast<uint32_t>(ty0) | (static_cast<uint32_t>(ty1) << 16)); }
__device__ inline const uint2 empty_rect() { return make_uint2(0x0000ffffu, 0x0000ffffu); }

__global__ __launch_bounds__(64) void k_setup_count(const DrawDesc* __restrict__ draws, FrameDesc f,
		const float4* __restrict__ vpos, const uint8_t* __restrict__ vflags,
		uint2* __restrict__ trect, uint4* __restrict__ tvert, uint32_t* __restrict__ hist,
		uint32_t* __restrict__ clipq, uint32_t* __restrict__ clipqOf, uint32_t clipCap,
		Counters* __restrict__ ctr) {
	extern __shared__ uint32_t lhist[];  // nTiles counters
	__shared__ DrawLite dl[kMaxDraws];
	const int u = blockIdx.x;
	load_lite(dl, draws, f.nDraws);
	for (int t = threadIdx.x; t < f.nTiles; t += 64) lhist[t] = 0;
	__syncthreads();
	uint32_t binned = 0;
	const long long s0 = static_cast<long long>(u) * kUnitTris;
	for (int r = 0; r < kUnitTris / 64; ++r) {
		const long long s = s0 + r * 64 + threadIdx.x;
		if (s >= f.totalTris) break;
		const int di = find_draw_ll(dl, f.nDraws, s);
		const DrawLite& d = dl[di];
		const long long n = s - d.triBase;
		const int iid = static_cast<int>(n / d.tris);
		const int t = static_cast<int>(n - static_cast<long long>(iid) * d.tris);
		uint32_t ii[3];
		bool bad = false;
		for (int k = 0; k < 3; ++k) {
			ii[k] = d.elements ? d.indices[3 * t + k] : static_cast<uint32_t>(3 * t + k);
			bad |= ii[k] >= static_cast<uint32_t>(d.verts); }
		uint2 rect = empty_rect();
		uint4 tv = make_uint4(0, 0, 0, 0);
		if (bad) {
			atomicOr(&ctr->overflow, 8u); }
		else {
			const long long vb = d.vBase + static_cast<long long>(iid) * d.verts;
			uint32_t cf[3]; float dcx[3], dcy[3];
			for (int k = 0; k < 3; ++k) {
				const float4 p = vpos[vb + ii[k]];
				dcx[k] = p.x; dcy[k] = p.y;
				cf[k] = vflags[vb + ii[k]]; }
			const bool pointsOutside = (cf[0] | cf[1] | cf[2]) != 0u;
			const bool primsOutside = (cf[0] & cf[1] & cf[2]) != 0u;
			if (!primsOutside && pointsOutside) {
				const uint32_t q = atomicAdd(&ctr->clipCount, 1u);
				if (q < clipCap) { clipq[q] = static_cast<uint32_t>(s); clipqOf[s] = q; }
				else atomicOr(&ctr->overflow, 2u); }
			if (!pointsOutside) {
				const float d31x = dcx[2] - dcx[0], d31y = dcy[2] - dcy[0];
				const float d21x = dcx[1] - dcx[0], d21y = dcy[1] - dcy[0];
				const float area = d31x*d21y - d31y*d21x;
				const bool front = area > 0.0F;
				const int32_t ix0 = cvtt(dcx[0]), iy0 = cvtt(dcy[0]);
				const int32_t ix1 = cvtt(dcx[1]), iy1 = cvtt(dcy[1]);
				const int32_t ix2 = cvtt(dcx[2]), iy2 = cvtt(dcy[2]);
				const int32_t vminx = imax(imin(ix0, imin(ix1, ix2)), d.scissor[0]);
				const int32_t vminy = imax(imin(iy0, imin(iy1, iy2)), d.scissor[1]);
				const int32_t vmaxx = imin(wadd(imax(ix0, imax(ix1, ix2)), 1), d.scissor[2] - 1);
				const int32_t vmaxy = imin(wadd(imax(iy0, imax(iy1, iy2)), 1), d.scissor[3] - 1);
				const bool notCulled = front ? d.keepFronts != 0 : d.keepBacks != 0;
				if (notCulled && vmaxx > vminx && vmaxy > vminy) {
					const int tx0 = vminx / f.tileW, ty0 = vminy / f.tileH;
					const int tx1 = vmaxx / f.tileW, ty1 = vmaxy / f.tileH;
					rect = pack_rect(tx0, ty0, tx1, ty1);
					// backface reorder (:634-637); the 0x8000 flag is not needed
					const uint32_t g0 = static_cast<uint32_t>(vb + (front ? ii[0] : ii[2]));
					const uint32_t g2 = static_cast<uint32_t>(vb + (front ? ii[2] : ii[0]));
					tv = make_uint4(g0, static_cast<uint32_t>(vb + ii[1]), g2, 0u);
					++binned;
					for (int ty = ty0; ty <= ty1; ++ty)
						for (int tx = tx0; tx <= tx1; ++tx)
							atomicAdd(&lhist[ty * f.tilesX + tx], 1u); }}}
		trect[s] = rect;
		tvert[s] = tv; }
	__syncthreads();
	for (int t = threadIdx.x; t < f.nTiles; t += 64) hist[static_cast<long long>(t) * f.units + u] = lhist[t];
	// wave reduce of the binned count
	for (int o = 32; o > 0; o >>= 1) binned += __shfl_xor(binned, o, 64);
	if (threadIdx.x == 0 && binned) atomicAdd(&ctr->binned, binned); }

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (three passes); out may alias in.
// n+1 entries are written: out[n] = total.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256, kScanPer = 4, kScanChunk = kScanThreads * kScanPer;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
	// sh: blockDim.x + 32 entries; Hillis-Steele over waves then lanes
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
	uint32_t x = v;
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t y = __shfl_up(x, o, 64);
		if (lane >= o) x += y; }
	if (lane == 63) sh[wave] = x;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t acc = 0;
		for (int w = 0; w < nw; ++w) { const uint32_t t = sh[w]; sh[w] = acc; acc += t; }
		sh[nw] = acc; }
	__syncthreads();
	const uint32_t r = sh[wave] + x - v;
	total = sh[nw];
	__syncthreads();
	return r; }

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in, long long n, uint32_t* __restrict__ bsum) {
	__shared__ uint32_t sh[kScanThreads / 64 + 1];
	const long long base = static_cast<long long>(blockIdx.x) * kScanChunk;
	uint32_t acc = 0;
	for (int k = 0; k < kScanPer; ++k) {
		const long long i = base + k * kScanThreads + threadIdx.x;
		if (i < n) acc += in[i]; }
	for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t t = 0;
		for (int w = 0; w < kScanThreads / 64; ++w) t += sh[w];
		bsum[blockIdx.x] = t; }}

__global__ __launch_bounds__(1024) void k_scan_bsums(uint32_t* __restrict__ bsum, int nb, uint32_t* __restrict__ total_out) {
	__shared__ uint32_t sh[1024 / 64 + 1];
	uint32_t carry = 0;
	for (int base = 0; base < nb; base += 1024) {
		const int i = base + threadIdx.x;
		const uint32_t v = i < nb ? bsum[i] : 0u;
		uint32_t tot;
		const uint32_t ex = block_exclusive_scan(v, sh, tot);
		if (i < nb) bsum[i] = carry + ex;
		carry += tot; }
	if (threadIdx.x == 0) *total_out = carry; }

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t* in, long long n, const uint32_t* __restrict__ bsum,
                                                          uint32_t* out, const uint32_t* __restrict__ total) {
	__shared__ uint32_t sh[kScanThreads / 64 + 1];
	const long long base = static_cast<long long>(blockIdx.x) * kScanChunk;
	// each thread owns kScanPer consecutive elements
	uint32_t v[kScanPer];
	uint32_t local = 0;
	for (int k = 0; k < kScanPer; ++k) {
		const long long i = base + static_cast<long long>(threadIdx.x) * kScanPer + k;
		v[k] = i < n ? in[i] : 0u;
		local += v[k]; }
	uint32_t tot;
	uint32_t run = bsum[blockIdx.x] + block_exclusive_scan(local, sh, tot);
	for (int k = 0; k < kScanPer; ++k) {
		const long long i = base + static_cast<long long>(threadIdx.x) * kScanPer + k;
		if (i < n) out[i] = run;
		run += v[k]; }
	if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = *total; }

// ---------------------------------------------------------------------------
// ordered scatter: the per-tile append loop of BinTriangles (:651-661),
// keeping submission order: units in order (scan offsets), rounds in
// order, lanes in order (ballot rank), one wave per unit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_scatter(FrameDesc f, const uint2* __restrict__ trect,
		const uint32_t* __restrict__ off, uint32_t* __restrict__ list, uint32_t listCap, Counters* __restrict__ ctr) {
	extern __shared__ uint32_t base[];  // nTiles write cursors
	const int u = blockIdx.x;
	for (int t = threadIdx.x; t < f.nTiles; t += 64) base[t] = off[static_cast<long long>(t) * f.units + u];
	__syncthreads();
	const int lane = threadIdx.x;
	const unsigned long long lt = (1ull << lane) - 1ull;
	bool ovf = false;
	const long long s0 = static_cast<long long>(u) * kUnitTris;
	for (int r = 0; r < kUnitTris / 64; ++r) {
		const long long s = s0 + r * 64 + lane;
		int tx0 = 1, tx1 = 0, ty0 = 1, ty1 = 0;
		if (s < f.totalTris) {
			const uint2 rc = trect[s];
			tx0 = rc.x & 0xffffu; tx1 = rc.x >> 16; ty0 = rc.y & 0xffffu; ty1 = rc.y >> 16; }
		const bool has = tx0 <= tx1 && ty0 <= ty1;
		int cx = tx0, cy = ty0;
		uint32_t cur = has ? static_cast<uint32_t>(cy * f.tilesX + cx) : 0xffffffffu;
		for (;;) {
			uint32_t m = cur;
			for (int o = 32; o > 0; o >>= 1) { const uint32_t y = __shfl_xor(m, o, 64); m = y < m ? y : m; }
			if (m == 0xffffffffu) break;
			const unsigned long long mask = __ballot(cur == m);
			const uint32_t b = base[m];
			if (cur == m) {
				const uint32_t pos = b + static_cast<uint32_t>(__popcll(mask & lt));
				if (pos < listCap) list[pos] = static_cast<uint32_t>(s); else ovf = true;
				if (++cx > tx1) { cx = tx0; ++cy; }
				cur = cy <= ty1 ? static_cast<uint32_t>(cy * f.tilesX + cx) : 0xffffffffu; }
			__syncthreads();
			if (lane == 0) base[m] = b + static_cast<uint32_t>(__popcll(mask));
			__syncthreads(); }}
	if (ovf) atomicOr(&ctr->overflow, 1u); }

// ---------------------------------------------------------------------------
// clipping: GPUBinImpl::ClipTriangles (rglv_gpu_impl.hxx:677-793),
// Sutherland-Hodgman against L, B, N, R, T, fan triangulation, per-tile
// counts (ForEachCoveredTile, :795-817).
// ---------------------------------------------------------------------------
__device__ inline float plane_dist(int plane, const float4& p) {
	switch (plane) {
	case 0: return p.w + p.x;
	case 1: return p.w + p.y;
	case 2: return p.w + p.z;
	case 3: return p.w - p.x;
	default: return p.w - p.y; }}

template <class P>
__global__ __launch_bounds__(64) void k_clip(const DrawDesc* __restrict__ draws, int drawIdx, FrameDesc f,
		const uint32_t* __restrict__ clipq, uint32_t clipCap, float4* __restrict__ cvpos, float* __restrict__ cvvary,
		uint8_t* __restrict__ fcount, uint2* __restrict__ frect, uint32_t* __restrict__ clipHist, Counters* __restrict__ ctr) {
	const DrawDesc& d = draws[drawIdx];
	const uint32_t nq = min(ctr->clipCount, clipCap);
	for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
		const long long s = clipq[q];
		if (s < d.triBase || s >= d.triBase + static_cast<long long>(d.tris) * d.instances) continue;
		const long long n = s - d.triBase;
		const int iid = static_cast<int>(n / d.tris);
		const int t = static_cast<int>(n - static_cast<long long>(iid) * d.tris);
		struct CV { float4 c; float v[P::kVary]; };
		CV A[kClipVerts], B[kClipVerts];
		int na = 3;
		for (int k = 0; k < 3; ++k) {
			const int idx = d.elements ? d.indices[3 * t + k] : 3 * t + k;
			float pos[4];
			P::vertex(d, iid, idx, pos, A[k].v);
			A[k].c = make_float4(pos[0], pos[1], pos[2], pos[3]); }
		// all three vertices outside one clipping plane: Sutherland-Hodgman
		// would emit an empty polygon for that plane
		bool rejected = false;
		for (int plane = 0; plane < 5; ++plane)
			rejected |= plane_dist(plane, A[0].c) < 0.0F && plane_dist(plane, A[1].c) < 0.0F && plane_dist(plane, A[2].c) < 0.0F;
		if (rejected) { fcount[q] = 0; continue; }
		for (int plane = 0; plane < 5 && na > 0; ++plane) {
			bool hereIsInside = plane_dist(plane, A[0].c) >= 0.0F;
			int nb = 0;
			for (int h = 0; h < na; ++h) {
				const int nx = (h + 1) % na;
				const bool nextIsInside = plane_dist(plane, A[nx].c) >= 0.0F;
				if (hereIsInside) B[nb++] = A[h];
				if (hereIsInside != nextIsInside) {
					const CV& from = hereIsInside ? A[h] : A[nx];
					const CV& to = hereIsInside ? A[nx] : A[h];
					const float fa = plane_dist(plane, from.c), fb = plane_dist(plane, to.c);
					const float dd = fa / (fa - fb);
					const float om = 1.0F - dd;
					CV& o = B[nb++];
					o.c = make_float4(om * from.c.x + dd * to.c.x, om * from.c.y + dd * to.c.y,
					                  om * from.c.z + dd * to.c.z, om * from.c.w + dd * to.c.w);
					for (int k = 0; k < P::kVary; ++k) o.v[k] = om * from.v[k] + dd * to.v[k];
					hereIsInside = !hereIsInside; }}
			na = nb;
			for (int k = 0; k < na; ++k) A[k] = B[k]; }
		if (na == 0) { fcount[q] = 0; continue; }
		for (int k = 0; k < na; ++k) {
			const float iw = oneover(A[k].c.w);
			float x = A[k].c.x * iw, y = A[k].c.y * iw;
			const float z = A[k].c.z * iw;
			x = x * d.DS[0] + d.DO[0];
			y = y * d.DS[1] + d.DO[1];
			A[k].c = make_float4(x, y, z, iw); }
		const float d31x = A[2].c.x - A[0].c.x, d31y = A[2].c.y - A[0].c.y;
		const float d21x = A[1].c.x - A[0].c.x, d21y = A[1].c.y - A[0].c.y;
		const bool backfacing = (d31x*d21y - d31y*d21x) < 0.0F;
		bool willCull = true;
		if (backfacing) { if (!d.cullingEnabled || (d.cullFace & 2) == 0) willCull = false; }
		else { if (!d.cullingEnabled || (d.cullFace & 1) == 0) willCull = false; }
		if (willCull) { fcount[q] = 0; continue; }
		for (int k = 0; k < na; ++k) {
			const CV& src = backfacing ? A[na - 1 - k] : A[k];
			cvpos[static_cast<long long>(q) * kClipVerts + k] = src.c;
			float* o = cvvary + (static_cast<long long>(q) * kClipVerts + k) * kMaxVary;
			for (int j = 0; j < P::kVary; ++j) o[j] = src.v[j]; }
		const int nf = na - 2;
		fcount[q] = static_cast<uint8_t>(nf);
		for (int j = 0; j < nf; ++j) {
			const float4 p0 = cvpos[static_cast<long long>(q) * kClipVerts + 0];
			const float4 p1 = cvpos[static_cast<long long>(q) * kClipVerts + j + 1];
			const float4 p2 = cvpos[static_cast<long long>(q) * kClipVerts + j + 2];
			const int ix0 = cvtt(p0.x), ix1 = cvtt(p1.x), ix2 = cvtt(p2.x);
			const int iy0 = cvtt(p0.y), iy1 = cvtt(p1.y), iy2 = cvtt(p2.y);
			const int vminx = imax(imin(ix0, imin(ix1, ix2)), d.scissor[0]);
			const int vminy = imax(imin(iy0, imin(iy1, iy2)), d.scissor[1]);
			const int vmaxx = imin(wadd(imax(ix0, imax(ix1, ix2)), 1), d.scissor[2] - 1);
			const int vmaxy = imin(wadd(imax(iy0, imax(iy1, iy2)), 1), d.scissor[3] - 1);
			int tx0 = vminx / f.tileW, ty0 = vminy / f.tileH;
			int tx1 = vmaxx / f.tileW, ty1 = vmaxy / f.tileH;
			tx1 = imin(tx1, f.tilesX - 1); ty1 = imin(ty1, f.tilesY - 1);
			tx0 = imax(tx0, 0); ty0 = imax(ty0, 0);
			if (tx0 > tx1 || ty0 > ty1) { frect[static_cast<long long>(q) * kFanMax + j] = empty_rect(); continue; }
			frect[static_cast<long long>(q) * kFanMax + j] = pack_rect(tx0, ty0, tx1, ty1);
			for (int ty = ty0; ty <= ty1; ++ty)
				for (int tx = tx0; tx <= tx1; ++tx)
					atomicAdd(&clipHist[ty * f.tilesX + tx], 1u); }}}

__global__ __launch_bounds__(1024) void k_clip_scan(FrameDesc f, const uint32_t* __restrict__ clipHist,
		uint32_t* __restrict__ clipOff, uint32_t* __restrict__ clipCursor, Counters* __restrict__ ctr) {
	__shared__ uint32_t sh[1024 / 64 + 1];
	uint32_t carry = 0;
	for (int base = 0; base < f.nTiles; base += 1024) {
		const int i = base + threadIdx.x;
		const uint32_t v = i < f.nTiles ? clipHist[i] : 0u;
		uint32_t tot;
		const uint32_t ex = block_exclusive_scan(v, sh, tot);
		if (i < f.nTiles) { clipOff[i] = carry + ex; clipCursor[i] = carry + ex; }
		carry += tot; }
	if (threadIdx.x == 0) { clipOff[f.nTiles] = carry; ctr->clipPairs = carry; }}

__global__ __launch_bounds__(256) void k_clip_scatter(FrameDesc f, const uint8_t* __restrict__ fcount,
		const uint2* __restrict__ frect, uint32_t clipCap, uint32_t* __restrict__ clipCursor,
		uint32_t* __restrict__ clist, uint32_t clistCap, Counters* __restrict__ ctr) {
	const uint32_t nq = min(ctr->clipCount, clipCap);
	for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
		const int nf = fcount[q];
		for (int j = 0; j < nf; ++j) {
			const uint2 rc = frect[static_cast<long long>(q) * kFanMax + j];
			const int tx0 = rc.x & 0xffffu, tx1 = rc.x >> 16, ty0 = rc.y & 0xffffu, ty1 = rc.y >> 16;
			for (int ty = ty0; ty <= ty1; ++ty)
				for (int tx = tx0; tx <= tx1; ++tx) {
					const uint32_t pos = atomicAdd(&clipCursor[ty * f.tilesX + tx], 1u);
					if (pos < clistCap) clist[pos] = q * kFanMax + j;
					else atomicOr(&ctr->overflow, 4u); }}}}

// ---------------------------------------------------------------------------
// per-tile raster + resolve
// ---------------------------------------------------------------------------
struct TriSetup {
	int32_t c1, c2, c3;
	int32_t dx12, dy12, dx23, dy23, dx31, dy31;
	int32_t minx, miny, nqx, nqy;
	float scale;
	float z0, z1, z2;
	uint32_t slot;
};

struct Rect { int x0, y0, x1, y1; };

// VTriangleRasterizer setup (rglv_triangle.hxx:206-240), one lane, rect r
__device__ inline void setup_vector(const Rect& r, const int32_t X[3], const int32_t Y[3], TriSetup& t) {
	const int32_t x1 = X[0], x2 = X[1], x3 = X[2], y1 = Y[0], y2 = Y[1], y3 = Y[2];
	int32_t minx = imax(imin(imin(x1, x2), x3) >> 4, r.x0);
	const int32_t maxx = imin(wadd(imax(imax(x1, x2), x3), 15) >> 4, r.x1);
	int32_t miny = imax(imin(imin(y1, y2), y3) >> 4, r.y0);
	const int32_t maxy = imin(wadd(imax(imax(y1, y2), y3), 15) >> 4, r.y1);
	minx &= ~1; miny &= ~1;
	t.dx12 = wsub(x1, x2); t.dy12 = wsub(y2, y1);
	t.dx23 = wsub(x2, x3); t.dy23 = wsub(y3, y2);
	t.dx31 = wsub(x3, x1); t.dy31 = wsub(y1, y3);
	const int32_t sx = wadd(shl(minx, 4), 8), sy = wadd(shl(miny, 4), 8);
	int32_t c1 = wadd(wmul(t.dy12, wsub(sx, x1)), wmul(t.dx12, wsub(sy, y1)));
	int32_t c2 = wadd(wmul(t.dy23, wsub(sx, x2)), wmul(t.dx23, wsub(sy, y2)));
	int32_t c3 = wadd(wmul(t.dy31, wsub(sx, x3)), wmul(t.dx31, wsub(sy, y3)));
	const int32_t tl12 = (t.dy12 > 0 || (t.dy12 == 0 && t.dx12 > 0)) ? 1 : 0;
	const int32_t tl23 = (t.dy23 > 0 || (t.dy23 == 0 && t.dx23 > 0)) ? 1 : 0;
	const int32_t tl31 = (t.dy31 > 0 || (t.dy31 == 0 && t.dx31 > 0)) ? 1 : 0;
	t.c1 = wsub(wadd(c1, tl12), 1) >> 4;
	t.c2 = wsub(wadd(c2, tl23), 1) >> 4;
	t.c3 = wsub(wadd(c3, tl31), 1) >> 4;
	t.scale = 1.0F / itof(wadd(wadd(t.c1, t.c2), t.c3));
	t.minx = minx; t.miny = miny;
	t.nqx = maxx > minx ? (maxx - minx + 1) >> 1 : 0;
	t.nqy = maxy > miny ? (maxy - miny + 1) >> 1 : 0; }

// TriangleRasterizer setup (rglv_triangle.hxx:84-131), int64 edges, clipped triangles
__device__ inline void setup_scalar(const Rect& r, const int32_t X[3], const int32_t Y[3], TriSetup& t) {
	const int x1 = X[0], x2 = X[1], x3 = X[2], y1 = Y[0], y2 = Y[1], y3 = Y[2];
	int minx = imax(imin(imin(x1, x2), x3) >> 4, r.x0);
	const int endx = imin(wadd(imax(imax(x1, x2), x3), 15) >> 4, r.x1);
	int miny = imax(imin(imin(y1, y2), y3) >> 4, r.y0);
	const int endy = imin(wadd(imax(imax(y1, y2), y3), 15) >> 4, r.y1);
	minx &= ~1; miny &= ~1;
	const long long dx12 = wsub(x1, x2), dy12 = wsub(y2, y1);
	const long long dx23 = wsub(x2, x3), dy23 = wsub(y3, y2);
	const long long dx31 = wsub(x3, x1), dy31 = wsub(y1, y3);
	const long long sx = (static_cast<long long>(minx) << 4) + 8, sy = (static_cast<long long>(miny) << 4) + 8;
	long long c1 = dy12*(sx - x1) + dx12*(sy - y1);
	long long c2 = dy23*(sx - x2) + dx23*(sy - y2);
	long long c3 = dy31*(sx - x3) + dx31*(sy - y3);
	if (dy12 > 0 || (dy12 == 0 && dx12 > 0)) c1++;
	--c1;
	if (dy23 > 0 || (dy23 == 0 && dx23 > 0)) c2++;
	--c2;
	if (dy31 > 0 || (dy31 == 0 && dx31 > 0)) c3++;
	--c3;
	c1 >>= 4; c2 >>= 4; c3 >>= 4;
	t.c1 = static_cast<int32_t>(c1); t.c2 = static_cast<int32_t>(c2); t.c3 = static_cast<int32_t>(c3);
	t.dx12 = static_cast<int32_t>(dx12); t.dy12 = static_cast<int32_t>(dy12);
	t.dx23 = static_cast<int32_t>(dx23); t.dy23 = static_cast<int32_t>(dy23);
	t.dx31 = static_cast<int32_t>(dx31); t.dy31 = static_cast<int32_t>(dy31);
	t.scale = 1.0F / static_cast<float>(c1 + c2 + c3);
	t.minx = minx; t.miny = miny;
	t.nqx = endx > minx ? (endx - minx + 1) >> 1 : 0;
	t.nqy = endy > miny ? (endy - miny + 1) >> 1 : 0; }

// edge values of the 4 lanes of quad (kx, ky) of a setup: the incremental
// steps of the rasteriser loops (rglv_triangle.hxx:253-304), wrapping
__device__ inline void quad_edges(const TriSetup& t, int kx, int ky, int32_t e1[4], int32_t e2[4], int32_t e3[4]) {
	const int32_t q1 = wadd(wadd(t.c1, wmul(kx, wmul(t.dy12, 2))), wmul(ky, wmul(t.dx12, 2)));
	const int32_t q2 = wadd(wadd(t.c2, wmul(kx, wmul(t.dy23, 2))), wmul(ky, wmul(t.dx23, 2)));
	const int32_t q3 = wadd(wadd(t.c3, wmul(kx, wmul(t.dy31, 2))), wmul(ky, wmul(t.dx31, 2)));
	e1[0] = q1; e1[1] = wadd(q1, t.dy12); e1[2] = wadd(q1, t.dx12); e1[3] = wadd(wadd(q1, t.dx12), t.dy12);
	e2[0] = q2; e2[1] = wadd(q2, t.dy23); e2[2] = wadd(q2, t.dx23); e2[3] = wadd(wadd(q2, t.dx23), t.dy23);
	e3[0] = q3; e3[1] = wadd(q3, t.dy31); e3[2] = wadd(q3, t.dx31); e3[3] = wadd(wadd(q3, t.dx31), t.dy31); }

__device__ inline Rect tile_rect(const FrameDesc& f, int tile) {
	const int ty = tile / f.tilesX, tx = tile - ty * f.tilesX;
	Rect r;
	r.x0 = tx * f.tileW; r.y0 = ty * f.tileH;
	r.x1 = imin(r.x0 + f.tileW, f.W); r.y1 = imin(r.y0 + f.tileH, f.H);
	return r; }

// the three device-coordinate vertices of a visibility slot
struct TriVerts {
	int32_t X[3], Y[3];
	float z[3], w[3];
	long long vary[3];   // index of each vertex's varyings (units of kMaxVary floats)
	bool clipped;
	int draw;
	uint32_t prim;
};

__device__ inline void inline_verts(const uint4 tv, const float4* __restrict__ vpos, TriVerts& tr) {
	const uint32_t g[3] = { tv.x, tv.y, tv.z };
	for (int k = 0; k < 3; ++k) {
		const float4 p = vpos[g[k]];
		tr.X[k] = cvtt(16.0F * p.x);   // GPUTileImpl::DrawTriangles (:885-886)
		tr.Y[k] = cvtt(16.0F * p.y);
		tr.z[k] = p.z; tr.w[k] = p.w;
		tr.vary[k] = g[k]; }
	tr.clipped = false; }

__device__ inline void clipped_verts(uint32_t fan, const float4* __restrict__ cvpos, TriVerts& tr) {
	const uint32_t q = fan / kFanMax, j = fan - q * kFanMax;
	const long long v[3] = { static_cast<long long>(q) * kClipVerts, static_cast<long long>(q) * kClipVerts + j + 1,
	                         static_cast<long long>(q) * kClipVerts + j + 2 };
	for (int k = 0; k < 3; ++k) {
		const float4 p = cvpos[v[k]];
		tr.X[k] = cvtt(p.x * 16.0F);   // DrawClipped (:997-998)
		tr.Y[k] = cvtt(p.y * 16.0F);
		tr.z[k] = p.z; tr.w[k] = p.w;
		tr.vary[k] = v[k]; }
	tr.clipped = true; }

struct RasterArgs {
	const DrawDesc* draws;
	const uint32_t* off;        // inline list offsets, tile-major [nTiles * units + 1]
	const uint32_t* list;
	uint32_t listCap;
	const uint4* tvert;
	const float4* vpos;
	const float* vvary;
	const uint32_t* clipq;
	const uint32_t* clipqOf;
	const float4* cvpos;
	const float* cvvary;
	const uint32_t* clipOff;
	const uint32_t* clist;
	uint32_t clistCap;
	uint32_t clipCap;
	Counters* ctr;
	uint32_t* vis;              // per-pixel winning slot, 0 = none
};

template <class P>
__device__ inline void shade_lanes(const DrawDesc& d, const TriSetup& t, const TriVerts& tr,
                                   const float* __restrict__ vary, int kx, int ky, float color[3][4], float depth[4]) {
	int32_t e1[4], e2[4], e3[4];
	quad_edges(t, kx, ky, e1, e2, e3);
	float bs[3][4], bp[3][4];
	for (int l = 0; l < 4; ++l) {
		bs[0][l] = itof(e2[l]) * t.scale;
		bs[2][l] = itof(e1[l]) * t.scale;
		bs[1][l] = (1.0F - bs[0][l]) - bs[2][l];
		depth[l] = (bs[0][l]*tr.z[0] + bs[1][l]*tr.z[1]) + bs[2][l]*tr.z[2];
		const float fragW = oneover((bs[0][l]*tr.w[0] + bs[1][l]*tr.w[1]) + bs[2][l]*tr.w[2]);
		bp[0][l] = (tr.w[0] * bs[0][l]) * fragW;
		bp[2][l] = (tr.w[2] * bs[2][l]) * fragW;
		bp[1][l] = (1.0F - bp[0][l]) - bp[2][l]; }
	float vv[P::kVary][4];
	const float* v0 = vary + tr.vary[0] * kMaxVary;
	const float* v1 = vary + tr.vary[1] * kMaxVary;
	const float* v2 = vary + tr.vary[2] * kMaxVary;
	for (int k = 0; k < P::kVary; ++k) {
		const float a = v0[k], b = v1[k], c = v2[k];
		for (int l = 0; l < 4; ++l) vv[k][l] = (bp[0][l]*a + bp[1][l]*b) + bp[2][l]*c; }
	P::fragment(d, vv, color); }

__device__ inline void shade_dispatch(const DrawDesc& d, const TriSetup& t, const TriVerts& tr,
                                      const float* vary, int kx, int ky, float color[3][4], float depth[4]) {
	switch (d.program) {
	case ProgAmy::kId: shade_lanes<ProgAmy>(d, t, tr, vary, kx, ky, color, depth); break;
	case ProgMany::kId: shade_lanes<ProgMany>(d, t, tr, vary, kx, ky, color, depth); break;
	case ProgLambertTex::kId: shade_lanes<ProgLambertTex>(d, t, tr, vary, kx, ky, color, depth); break;
	default: shade_lanes<ProgManyBlinnPhong>(d, t, tr, vary, kx, ky, color, depth); break; }}

__global__ __launch_bounds__(kRasterThreads) void k_vis(FrameDesc f, RasterArgs a) {
	__shared__ unsigned long long keys[kMaxTileDim * kMaxTileDim];
	__shared__ TriSetup ts[kRasterThreads];
	__shared__ uint32_t qoff[kRasterThreads + 1];
	__shared__ uint32_t scan_sh[kRasterThreads / 64 + 1];
	__shared__ DrawLite dl[kMaxDraws];
	const int tile = blockIdx.x;
	const Rect R = tile_rect(f, tile);
	const bool leq = f.depthFunc == 1;
	const unsigned long long initKey = (static_cast<unsigned long long>(depth_order(f.clearDepth)) << 32) |
	                                   (leq ? 0xffffffffull : 0ull);
	const int tileArea = f.tileW * f.tileH;
	for (int i = threadIdx.x; i < tileArea; i += kRasterThreads) keys[i] = initKey;
	load_lite(dl, a.draws, f.nDraws);
	__syncthreads();
	unsigned long long frags = 0;

	for (int pass = 0; pass < 2 && f.drawEnabled; ++pass) {
		uint32_t beg, end;
		if (pass == 0) {
			beg = a.off[static_cast<long long>(tile) * f.units];
			end = a.off[static_cast<long long>(tile + 1) * f.units];
			end = min(end, a.listCap); }
		else {
			beg = a.clipOff[tile]; end = min(a.clipOff[tile + 1], a.clistCap); }
		for (uint32_t b0 = beg; b0 < end; b0 += kRasterThreads) {
			const uint32_t e = b0 + threadIdx.x;
			uint32_t nq = 0;
			if (e < end) {
				TriVerts tr;
				TriSetup t;
				if (pass == 0) {
					uint32_t s = a.list[e];
					if (s >= f.totalTris) { atomicOr(&a.ctr->overflow, 16u); s = 0; }
					const int di = find_draw_ll(dl, f.nDraws, s);
					inline_verts(a.tvert[s], a.vpos, tr);
					setup_vector(R, tr.X, tr.Y, t);
					t.slot = dl[di].slotBase + static_cast<uint32_t>(s - dl[di].triBase); }
				else {
					uint32_t fan = a.clist[e];
					if (fan / kFanMax >= a.clipCap) { atomicOr(&a.ctr->overflow, 16u); fan = 0; }
					const uint32_t q = fan / kFanMax, j = fan - q * kFanMax;
					const uint32_t s = a.clipq[q];
					const int di = find_draw_ll(dl, f.nDraws, s);
					clipped_verts(fan, a.cvpos, tr);
					setup_scalar(R, tr.X, tr.Y, t);
					const uint32_t nTri = static_cast<uint32_t>(dl[di].tris) * static_cast<uint32_t>(dl[di].instances);
					t.slot = dl[di].slotBase + nTri + static_cast<uint32_t>(s - dl[di].triBase) * kFanMax + j; }
				t.z0 = tr.z[0]; t.z1 = tr.z[1]; t.z2 = tr.z[2];
				ts[threadIdx.x] = t;
				nq = static_cast<uint32_t>(t.nqx) * static_cast<uint32_t>(t.nqy); }
			uint32_t total;
			const uint32_t ex = block_exclusive_scan(nq, scan_sh, total);
			qoff[threadIdx.x] = ex;
			if (threadIdx.x == 0) qoff[kRasterThreads] = total;
			__syncthreads();
			for (uint32_t it = threadIdx.x; it < total; it += kRasterThreads) {
				int lo = 0, hi = kRasterThreads - 1;  // last j with qoff[j] <= it
				while (lo < hi) {
					const int mid = (lo + hi + 1) >> 1;
					if (qoff[mid] <= it) lo = mid; else hi = mid - 1; }
				const TriSetup& t = ts[lo];
				const uint32_t k = it - qoff[lo];
				const int ky = static_cast<int>(k / static_cast<uint32_t>(t.nqx));
				const int kx = static_cast<int>(k - static_cast<uint32_t>(ky) * static_cast<uint32_t>(t.nqx));
				int32_t e1[4], e2[4], e3[4];
				quad_edges(t, kx, ky, e1, e2, e3);
				const int x = t.minx + 2 * kx, y = t.miny + 2 * ky;
				for (int l = 0; l < 4; ++l) {
					if ((e1[l] | e2[l] | e3[l]) < 0) continue;
					++frags;
					const float bx = itof(e2[l]) * t.scale;
					const float bz = itof(e1[l]) * t.scale;
					const float by = (1.0F - bx) - bz;
					const float dz = (bx*t.z0 + by*t.z1) + bz*t.z2;
					if (dz != dz) continue;
					const unsigned long long key = (static_cast<unsigned long long>(depth_order(dz)) << 32) |
					                               (leq ? static_cast<unsigned long long>(~t.slot) : static_cast<unsigned long long>(t.slot));
					const int li = (y + (l >> 1) - R.y0) * f.tileW + (x + (l & 1) - R.x0);
					if (key < keys[li]) atomicMin(&keys[li], key); }}
			__syncthreads(); }}

	// ---- write each pixel's winning slot (0 = nothing drawn) ----
	const int rw = R.x1 - R.x0, rh = R.y1 - R.y0;
	for (int i = threadIdx.x; i < rw * rh; i += kRasterThreads) {
		const int ly = i / rw, lx = i - ly * rw;
		const unsigned long long k = keys[ly * f.tileW + lx];
		const uint32_t lo32 = static_cast<uint32_t>(k);
		a.vis[static_cast<long long>(R.y0 + ly) * f.W + R.x0 + lx] = (k == initKey) ? 0u : (leq ? ~lo32 : lo32); }

	for (int o = 32; o > 0; o >>= 1) frags += __shfl_xor(frags, o, 64);
	if ((threadIdx.x & 63) == 0 && frags) atomicAdd(&a.ctr->fragments, frags); }


// ---------------------------------------------------------------------------
// resolve: one thread per 2x2 quad of the screen.  Each pixel's winning slot
// is shaded by recomputing its triangle's quad exactly as the reference's
// rasteriser produced it in that pixel's tile (tile-clamped edge start,
// barycentric scale, lanes 0/1 for the texture LOD), then post-shaded and
// sRGB-encoded (FilterTile, rglr_algorithm.hxx:31-104), and colour, depth
// and primitive id are written once.
// ---------------------------------------------------------------------------
constexpr int kResolveQx = 32, kResolveQy = 8;   // 256 threads = 32 x 8 quads

__global__ __launch_bounds__(256) void k_resolve(FrameDesc f, RasterArgs a) {
	__shared__ DrawLite dl[kMaxDraws];
	load_lite(dl, a.draws, f.nDraws);
	__syncthreads();
	const int qx = blockIdx.x * kResolveQx + (threadIdx.x % kResolveQx);
	const int qy = blockIdx.y * kResolveQy + (threadIdx.x / kResolveQx);
	const int x = 2 * qx, y = 2 * qy;
	if (x >= f.W || y >= f.H) return;
	const int tx = x / f.tileW, ty = y / f.tileH;
	Rect R;
	R.x0 = tx * f.tileW; R.y0 = ty * f.tileH;
	R.x1 = imin(R.x0 + f.tileW, f.W); R.y1 = imin(R.y0 + f.tileH, f.H);
	uint32_t win[4];
	for (int l = 0; l < 4; ++l) {
		const int px = x + (l & 1), py = y + (l >> 1);
		win[l] = (px < f.W && py < f.H) ? a.vis[static_cast<long long>(py) * f.W + px] : 0u; }
	float col[3][4], dep[4];
	uint32_t prim[4];
	for (int l = 0; l < 4; ++l) {
		col[0][l] = f.clearColor[0]; col[1][l] = f.clearColor[1]; col[2][l] = f.clearColor[2];
		dep[l] = f.clearDepth; prim[l] = 0xffffffffu; }
	uint32_t done = 0;
	for (int l = 0; l < 4; ++l) {
		if (win[l] == 0u || (done & (1u << l))) continue;
		const uint32_t slot = win[l];
		const int di = find_draw_slot(dl, f.nDraws, slot);
		const DrawLite& d = dl[di];
		const uint32_t rel = slot - d.slotBase;
		const uint32_t nTri = static_cast<uint32_t>(d.tris) * static_cast<uint32_t>(d.instances);
		TriVerts tr;
		TriSetup t;
		uint32_t srcTri;
		const float* vary;
		if (rel < nTri) {
			srcTri = rel;
			inline_verts(a.tvert[d.triBase + rel], a.vpos, tr);
			setup_vector(R, tr.X, tr.Y, t);
			vary = a.vvary; }
		else {
			const uint32_t fslot = rel - nTri;
			srcTri = fslot / kFanMax;
			const uint32_t j = fslot - srcTri * kFanMax;
			const uint32_t cq = a.clipqOf[d.triBase + srcTri];
			clipped_verts(cq * kFanMax + j, a.cvpos, tr);
			setup_scalar(R, tr.X, tr.Y, t);
			vary = a.cvvary; }
		const int kx = (x - t.minx) >> 1, ky = (y - t.miny) >> 1;
		float c4[3][4], d4[4];
		shade_dispatch(a.draws[di], t, tr, vary, kx, ky, c4, d4);
		for (int m = l; m < 4; ++m) {
			if (win[m] != slot) continue;
			done |= 1u << m;
			col[0][m] = c4[0][m]; col[1][m] = c4[1][m]; col[2][m] = c4[2][m];
			dep[m] = d4[m];
			prim[m] = d.primBase + srcTri; }}
	for (int l = 0; l < 4; ++l) {
		const int px = x + (l & 1), py = y + (l >> 1);
		if (px >= f.W || py >= f.H) continue;
		const long long o = static_cast<long long>(py) * f.W + px;
		if (f.outColor) {
			float r = col[0][l], g = col[1][l], b = col[2][l];
			if (f.postProgram == 2) { r = r * f.exposure; g = g * f.exposure; b = b * f.exposure; }
			f.outColor[o] = f.gamma ? ((to_srgb8(r) << 16) | (to_srgb8(g) << 8) | to_srgb8(b))
			                        : ((to_linear8(r) << 16) | (to_linear8(g) << 8) | to_linear8(b)); }
		if (f.outDepth) f.outDepth[o] = dep[l];
		if (f.outPrim) f.outPrim[o] = prim[l]; }}

__global__ void k_zero_counters(Counters* ctr, uint32_t* clipHist, int nTiles) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0) { ctr->fragments = 0; ctr->clipCount = 0; ctr->pairs = 0; ctr->clipPairs = 0; ctr->binned = 0; ctr->overflow = 0; }
	for (int t = i; t < nTiles; t += gridDim.x * blockDim.x) clipHist[t] = 0u; }

}  // namespace rgl
