// rgl_kernels.h — the gfx950 kernels of one rgl frame.
//
// Frame = vertex -> setup/count -> clip -> scan -> scatter (pair setup
// records, clipped fans, key resets) -> visibility per tile -> resolve.
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

#include "rgl_programs.h"

namespace rgl {

constexpr int kUnitTris = 512;      // triangles per binning unit
constexpr int kUnitThreads = 256;   // threads per binning-unit workgroup (setup/count and scatter)
constexpr int kXcds = 8;             // workgroups are dealt round-robin over the eight XCDs

// Per-tile pair counts are kept in kCountGroups groups: unit u adds its
// counts into group u % kCountGroups and later takes its list positions
// from the same group's cursor, so a count address only takes atomics from
// the units of one XCD.  The scan lays a tile's groups out one after the
// other in its list.
constexpr int kCountGroups = kXcds;
constexpr int kRasterThreads = 256;
constexpr int kMaxTileDim = 64;     // key-path tiles are at most 64 x 64 pixels
constexpr int kMaxAnyTileDim = 128; // in-order tiles at most 128 x 128 (the reference's limit)
constexpr int kSegPairs = 512;       // default (triangle, tile) pairs per visibility work item
constexpr int kSplitQuads = 0;       // default quad work per quadrant from which visibility regions are halved
constexpr int kSubDim = kMaxTileDim / 2;   // visibility quadrant: 32 x 32 pixels
constexpr int kQuadrants = 4;
constexpr int kMaxRegions = 16;      // visibility regions per tile (quadrants cut in four)

// Tallies that many workgroups add to are split into kCtrSlots slots, one
// 128-byte line each, and a workgroup adds into slot blockIdx % kCtrSlots:
// workgroups are dealt round-robin over the eight XCDs, so each slot's
// atomics come from one XCD's L2.  One device-wide address taking atomics
// from every XCD costs tens of microseconds per kernel.
constexpr int kCtrSlots = 8;

struct alignas(128) CtrSlot {
	unsigned long long fragments;
	uint32_t binned;         // triangles accepted into >= 1 tile
	uint32_t pad[29];
};
static_assert(sizeof(CtrSlot) == 128, "one line per slot");

struct alignas(128) Counters {
	uint32_t clipCount;
	uint32_t pairs;          // inline (triangle, tile) pairs, from the scan
	uint32_t clipPairs;
	uint32_t overflow;       // bit 0 list, 1 clip queue, 2 clipped list, 3 bad index
	uint32_t workItems;      // visibility work items of the frame
	uint32_t workNext;       // next work item to hand out
	uint32_t fallbackQuads;  // learned shading: tile quadrants re-shaded by the plain shader (4 per tile)
	uint32_t pad[25];
	CtrSlot slot[kCtrSlots];
};
static_assert(sizeof(Counters) % 4 == 0, "zeroed as 32-bit words");

// compact per-draw view used by the multi-draw kernels, kept in dynamic
// LDS (nDraws entries, sized by the launch)
struct alignas(16) DrawLite {
	long long triBase;
	long long vBase;
	const uint16_t* indices;
	int tris, instances, verts, elements;
	int keepBacks, keepFronts;
	int scissor[4];
	uint32_t slotBase;
	uint32_t primBase;
	int tileW, tileH, tilesX, tilesY, tileBase;   // the draw's frame tile grid
	int fW, fH;                                  // and its size
	int scissorOn;                               // 1: the raster rect is the tile rect clipped by scissor[]
};

__device__ inline int find_draw_ll(const DrawLite* dl, int nDraws, long long s) {
	int lo = 0, hi = nDraws - 1;
	while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (dl[mid].triBase <= s) lo = mid; else hi = mid - 1; }
	return lo; }

// The tiles binning unit u can touch: triangles are numbered frame by
// frame, so they lie between the first tile of its first triangle's frame
// and the last tile of its last triangle's frame.  The unit's LDS tile
// table covers this window only (at most b.unitWindow entries, which the
// host sizes from the same rule).
struct TileWindow { int lo, n; };
__device__ inline TileWindow unit_window(const DrawLite* dl, const BatchDesc& b, int u) {
	const long long s0 = static_cast<long long>(u) * b.unitTris;
	const long long s1 = min(s0 + b.unitTris, b.totalTris) - 1;
	const DrawLite& d0 = dl[find_draw_ll(dl, b.nDraws, s0)];
	const DrawLite& d1 = dl[find_draw_ll(dl, b.nDraws, s1)];
	TileWindow w;
	w.lo = d0.tileBase;
	w.n = d1.tileBase + d1.tilesX * d1.tilesY - d0.tileBase;
	return w; }

__device__ inline int find_draw_slot(const DrawLite* dl, int nDraws, uint32_t slot) {
	int lo = 0, hi = nDraws - 1;
	while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (dl[mid].slotBase <= slot) lo = mid; else hi = mid - 1; }
	return lo; }

__host__ __device__ inline DrawLite make_lite(const DrawDesc& d) {
	DrawLite l;
	l.triBase = d.triBase; l.vBase = d.vBase; l.indices = d.indices;
	l.tris = d.tris; l.instances = d.instances; l.verts = d.verts; l.elements = d.elements;
	l.keepBacks = d.keepBacks; l.keepFronts = d.keepFronts;
	for (int i = 0; i < 4; ++i) l.scissor[i] = d.scissor[i];
	l.slotBase = d.slotBase; l.primBase = d.primBase;
	l.tileW = d.tileW; l.tileH = d.tileH; l.tilesX = d.tilesX; l.tilesY = d.tilesY; l.tileBase = d.tileBase;
	l.fW = d.fW; l.fH = d.fH;
	l.scissorOn = d.scissorOn;
	return l; }

__device__ inline void load_lite(DrawLite* dl, const DrawDesc* draws, int nDraws) {
	for (int k = threadIdx.x; k < nDraws; k += blockDim.x) dl[k] = make_lite(draws[k]); }

// LDS table of the frames' first tiles (n + 1 entries, the last = total)
__device__ inline void load_tile_bases(int* ftb, const BatchDesc& b) {
	for (int k = threadIdx.x; k <= b.nFrames; k += blockDim.x) ftb[k] = k < b.nFrames ? b.frames[k].tileBase : b.nTiles; }

__device__ inline void load_block_bases(int* fbb, const BatchDesc& b) {
	for (int k = threadIdx.x; k <= b.nFrames; k += blockDim.x) fbb[k] = k < b.nFrames ? b.frames[k].blockBase : b.nBlocks; }

// ---------------------------------------------------------------------------
// vertex stage: GPUBinImpl::BinTriangles2P vertex loop (rglv_gpu_impl.hxx:364-385)
// once per (instance, vertex) instead of once per tile (DrawTriangles, :880-886)
// ---------------------------------------------------------------------------
// one launch covers every draw of the batch that uses program P; a thread
// finds its draw by the draws' first vertex-cache entries
// the frame's start-of-frame clears: counters, clip histogram, per-tile
// pair counts and network queue counters (run by the first vertex launch,
// or by k_zero_counters when a frame has no vertices)
struct ZeroArgs {
	Counters* ctr; uint32_t* clipHist; int nTiles; uint32_t* tileCount; uint32_t* qcount;
	uint32_t* tileResv; uint32_t* clipResv;   // the fused scatter's reservation counters (null: none)
};

__device__ inline void zero_frame(const ZeroArgs& z) {
	const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	const long long n = static_cast<long long>(gridDim.x) * blockDim.x;
	for (long long c = i; c < static_cast<long long>(kCountGroups) * z.nTiles; c += n) z.tileCount[c] = 0u;
	for (long long c = i; c < static_cast<long long>(sizeof(Counters) / 4); c += n) reinterpret_cast<uint32_t*>(z.ctr)[c] = 0u;
	for (long long c = i; c < 8 * 32; c += n) z.qcount[c] = 0u;
	for (long long t = i; t < z.nTiles; t += n) z.clipHist[t] = 0u;
	if (z.tileResv) {
		for (long long c = i; c < static_cast<long long>(kCountGroups) * z.nTiles; c += n) z.tileResv[c] = 0u;
		for (long long t = i; t < z.nTiles; t += n) z.clipResv[t] = 0u; }}

// one transformed vertex: ShadeVertex, clip flags, pdiv, viewport
template <class P>
__device__ inline void vertex_one(const DrawDesc& d, long long gi, float4* __restrict__ vpos, uint8_t* __restrict__ vflags,
                                  float* __restrict__ vvary, int varyStride) {
	if (d.program != P::kId) return;
	const long long i = gi - d.vBase;
	const int iid = static_cast<int>(i / d.verts);
	const int v = static_cast<int>(i - static_cast<long long>(iid) * d.verts);
	float pos[4], vo[P::kVary];
	P::vertex(d, iid, v, pos, vo);
	// ViewFrustum::Test, SIMD form with cmple (rglv_view_frustum.hxx:60-73)
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
	vpos[gi] = make_float4(dcx, dcy, pos[2] * iw, iw);
	vflags[gi] = static_cast<uint8_t>(f);
	float* out = vvary + gi * varyStride;
	for (int k = 0; k < P::kVary; ++k) out[k] = vo[k]; }

template <class P>
__global__ __launch_bounds__(256) void k_vertex(const DrawDesc* __restrict__ draws, int nDraws, long long vtot,
                                                 float4* __restrict__ vpos, uint8_t* __restrict__ vflags, float* __restrict__ vvary,
                                                 ZeroArgs z, int zeroHere, int varyStride) {
	if (zeroHere) zero_frame(z);
	const long long gi = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (gi >= vtot) return;
	int lo = 0, hi = nDraws - 1;
	while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (draws[mid].vBase <= gi) lo = mid; else hi = mid - 1; }
	// a wave whose vertices all belong to one draw reads that draw's
	// descriptor through scalar loads
	const int d0 = __builtin_amdgcn_readfirstlane(lo);
	if (__all(lo == d0)) vertex_one<P>(draws[d0], gi, vpos, vflags, vvary, varyStride);
	else vertex_one<P>(draws[lo], gi, vpos, vflags, vvary, varyStride); }

// ---------------------------------------------------------------------------
// triangle setup + per-unit tile counts: the primitive loop of
// BinTriangles1P/2P (rglv_gpu_impl.hxx:566-661) — cull, clip test,
// backface reorder, integer bbox, tile range.
// ---------------------------------------------------------------------------
__device__ inline uint2 pack_rect(int tx0, int ty0, int tx1, int ty1) {
	return make_uint2(static_cast<uint32_t>(tx0) | (static_cast<uint32_t>(tx1) << 16),
	                  static_cast<uint32_t>(ty0) | (static_cast<uint32_t>(ty1) << 16)); }
__device__ inline const uint2 empty_rect() { return make_uint2(0x0000ffffu, 0x0000ffffu); }

__global__ __launch_bounds__(kUnitThreads) void k_setup_count(const DrawDesc* __restrict__ draws, BatchDesc b,
		const float4* __restrict__ vpos, const uint8_t* __restrict__ vflags,
		uint2* __restrict__ trect, uint4* __restrict__ tvert, uint32_t* __restrict__ tileCount,
		uint32_t* __restrict__ clipq, uint32_t* __restrict__ clipqOf, uint32_t clipCap,
		Counters* __restrict__ ctr) {
	// dynamic LDS: the batch's draws, then the unit's tile-window counters
	extern __shared__ uint4 dyn_lds[];
	DrawLite* dl = reinterpret_cast<DrawLite*>(dyn_lds);
	uint32_t* lhist = reinterpret_cast<uint32_t*>(dl + b.nDraws);
	const int u = blockIdx.x;
	load_lite(dl, draws, b.nDraws);
	__syncthreads();
	const TileWindow win = unit_window(dl, b, u);
	if (win.n > b.unitWindow) {   // the host's window bound disagrees: flag it, touch nothing
		if (threadIdx.x == 0) atomicOr(&ctr->overflow, 16u);
		return; }
	for (int t = threadIdx.x; t < win.n; t += blockDim.x) lhist[t] = 0;   // entry t: batch tile win.lo + t
	__syncthreads();
	uint32_t binned = 0;
	const long long s0 = static_cast<long long>(u) * b.unitTris;
	// A triangle in three stages: its vertex indices, then its vertices'
	// positions and clip flags, then the work.  A unit of kUnitTris
	// triangles (kPer per thread) runs each stage for all of a thread's
	// triangles before the next, so their loads are in flight together.
	struct Tri { long long s; long long vb; int di; bool live, bad; uint32_t ii[3]; uint32_t cf[3]; float dcx[3], dcy[3]; };
	auto indices = [&](Tri& q) {
		q.live = q.s < b.totalTris; q.bad = false; q.di = 0; q.vb = 0;
		if (!q.live) return;
		q.di = find_draw_ll(dl, b.nDraws, q.s);
		const DrawLite& d = dl[q.di];
		const long long n = q.s - d.triBase;
		const int iid = static_cast<int>(n / d.tris);
		const int t = static_cast<int>(n - static_cast<long long>(iid) * d.tris);
		for (int k = 0; k < 3; ++k) q.ii[k] = d.elements ? d.indices[3 * t + k] : static_cast<uint32_t>(3 * t + k);
		q.vb = d.vBase + static_cast<long long>(iid) * d.verts; };
	auto vertices = [&](Tri& q) {
		if (!q.live) return;
		const DrawLite& d = dl[q.di];
		for (int k = 0; k < 3; ++k) q.bad |= q.ii[k] >= static_cast<uint32_t>(d.verts);
		if (q.bad) return;
		for (int k = 0; k < 3; ++k) {
			const float4 p = vpos[q.vb + q.ii[k]];
			q.dcx[k] = p.x; q.dcy[k] = p.y;
			q.cf[k] = vflags[q.vb + q.ii[k]]; }};
	// every lane of the wave runs it (the clip queue takes a ballot)
	auto work = [&](const Tri& q) {
		uint2 rect = empty_rect();
		uint4 tv = make_uint4(0, 0, 0, 0);
		bool clips = false;
		if (q.live && q.bad) atomicOr(&ctr->overflow, 8u);
		if (q.live && !q.bad) {
			const DrawLite& d = dl[q.di];
			const uint32_t* ii = q.ii;
			const float* dcx = q.dcx;
			const float* dcy = q.dcy;
			const long long vb = q.vb;
			const bool pointsOutside = (q.cf[0] | q.cf[1] | q.cf[2]) != 0u;
			const bool primsOutside = (q.cf[0] & q.cf[1] & q.cf[2]) != 0u;
			clips = !primsOutside && pointsOutside;
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
					const int tx0 = vminx / d.tileW, ty0 = vminy / d.tileH;
					const int tx1 = vmaxx / d.tileW, ty1 = vmaxy / d.tileH;
					rect = pack_rect(tx0, ty0, tx1, ty1);
					// backface reorder (:634-637); the 0x8000 flag is not needed
					const uint32_t g0 = static_cast<uint32_t>(vb + (front ? ii[0] : ii[2]));
					const uint32_t g2 = static_cast<uint32_t>(vb + (front ? ii[2] : ii[0]));
					tv = make_uint4(g0, static_cast<uint32_t>(vb + ii[1]), g2, 0u);
					++binned;
					for (int ty = ty0; ty <= ty1; ++ty)
						for (int tx = tx0; tx <= tx1; ++tx)
							atomicAdd(&lhist[d.tileBase - win.lo + ty * d.tilesX + tx], 1u); }}}
		if (q.live) { trect[q.s] = rect; tvert[q.s] = tv; }
		// the wave's clipped triangles take consecutive clip-queue entries
		// from one atomic
		const unsigned long long cm = __ballot(clips);
		if (cm) {
			const unsigned lane = threadIdx.x & 63u;
			const int leader = __ffsll(static_cast<long long>(cm)) - 1;
			uint32_t q0 = 0;
			if (lane == static_cast<unsigned>(leader)) q0 = atomicAdd(&ctr->clipCount, static_cast<uint32_t>(__popcll(cm)));
			q0 = __shfl(q0, leader, 64);
			if (clips) {
				const uint32_t qi = q0 + static_cast<uint32_t>(__popcll(cm & ((1ull << lane) - 1ull)));
				if (qi < clipCap) { clipq[qi] = static_cast<uint32_t>(q.s); clipqOf[q.s] = qi; }
				else atomicOr(&ctr->overflow, 2u); }}};
	constexpr int kPer = kUnitTris / kUnitThreads;
	if (b.unitTris == kUnitTris) {
		Tri q[kPer];
#pragma unroll
		for (int k = 0; k < kPer; ++k) { q[k].s = s0 + threadIdx.x + k * kUnitThreads; indices(q[k]); }
#pragma unroll
		for (int k = 0; k < kPer; ++k) vertices(q[k]);
#pragma unroll
		for (int k = 0; k < kPer; ++k) work(q[k]); }
	else for (int r = threadIdx.x; r < b.unitTris; r += blockDim.x) {   // grown units
		Tri q;
		q.s = s0 + r;
		if (q.s >= b.totalTris) break;
		indices(q);
		vertices(q);
		work(q); }
	__syncthreads();
	uint32_t* tc = tileCount + static_cast<long long>(u % kCountGroups) * b.nTiles;
	for (int t = threadIdx.x; t < win.n; t += blockDim.x) {
		const uint32_t c = lhist[t];
		if (c) atomicAdd(&tc[win.lo + t], c); }
	// wave reduce of the binned count
	for (int o = 32; o > 0; o >>= 1) binned += __shfl_xor(binned, o, 64);
	if ((threadIdx.x & 63) == 0 && binned) atomicAdd(&ctr->slot[blockIdx.x % kCtrSlots].binned, binned); }

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (three passes); out may alias in.
// n+1 entries are written: out[n] = total.
// ---------------------------------------------------------------------------

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

// The list is tile-major: tile t's inline pairs take list positions
// [tileOff[t], tileOff[t + 1]), laid out by count group (the units u with
// u % kCountGroups == g take one contiguous block, tileCursor[g * nTiles + t]
// being its start).  Within a group, each unit reserves one block per tile
// with an atomic add on the group's cursor, so the order of units inside a
// tile's list is not fixed; the visibility result does not depend on it
// (every fragment competes by its depth and submission slot).
constexpr int kScanThreads = 512; // scan block: one tile per thread

// Visibility work items of a tile: its inline + clipped pairs are cut into
// segments of segPairs, and the tile into regions: the four 32 x 32
// quadrants, each cut in two (32 x 16) when its share of the tile's quad
// work exceeds splitQuads and in four (16 x 16) above twice that.  The quad
// work is the previous frame's (tileCost, accumulated by k_vis); every
// split draws the same frame.  A work item is (tile, segment, region); the
// visibility kernel recovers the region count from the tile's item count.
__device__ inline uint32_t vis_segments(const BatchDesc& b, uint32_t pairs) {
	const uint32_t sp = static_cast<uint32_t>(b.segPairs);
	return max(1u, (pairs + sp - 1) / sp); }

__device__ inline uint32_t vis_regions(const BatchDesc& b, uint32_t cost, uint32_t nseg) {
	if (b.splitQuads <= 0) return kQuadrants;
	const uint32_t per = cost / (nseg * kQuadrants), s = static_cast<uint32_t>(b.splitQuads);
	return per > 2 * s ? 4 * kQuadrants : per > s ? 2 * kQuadrants : kQuadrants; }

constexpr int kCostClasses = 33;   // cost 0, then [2^(c-1), 2^c) for c = 1 .. 32
constexpr int kScanSplitTiles = 2 * kScanThreads;   // tiles above which the scan takes three launches
__device__ inline int cost_class(uint32_t cost) { return cost ? 32 - __clz(cost) : 0; }

// A tile's visibility work as the scan lays it out, so that a work item
// reads its tile (itemTile) and then one 32-byte descriptor instead of a
// chain of table lookups.
struct alignas(16) VisTile {
	uint32_t item0;        // the tile's first work item
	uint32_t frame;
	uint32_t ib, cb;       // the tile's first inline and clipped records (clamped to the capacities)
	uint32_t ninl, ntot;   // the tile's inline records, and inline + clipped
	uint32_t nreg;         // regions per segment
	uint32_t multiSeg;     // 1: the tile has several segments (regions merge with atomic min)
};
static_assert(sizeof(VisTile) == 32, "two 16-byte loads");

// One block: per tile the list offset (exclusive scan of the tiles' pair
// counts) and the count groups' cursors, the clipped-list offset and cursor
// (scan of clipHist) and the tile's visibility work items.  kThreads
// threads, at most kTiles tiles; null cursors: the count groups' and
// clipped cursors are not written (the fused scatter reserves its list
// positions from zeroed counters and these offsets instead).
struct ScanOut {
	uint32_t* tileCursor; uint32_t* tileOff; uint32_t* clipOff; uint32_t* clipCursor;
	uint32_t* itemTile; uint32_t itemCap; VisTile* tileDesc; uint32_t listCap, clistCap; uint32_t* tileCost; Counters* ctr;
};

template <int kThreads, int kTiles>
__device__ inline void col_scan_block(const BatchDesc& b, const uint32_t* __restrict__ tileCount,
		const uint32_t* __restrict__ clipHist, const ScanOut& o) {
	__shared__ uint32_t sh[kThreads / 64 + 1];
	__shared__ int ftb[kMaxFrames + 1];
	load_tile_bases(ftb, b);
	__syncthreads();
	const int nT = b.nTiles;   // at most kTiles: each thread keeps its tiles' descriptors in registers
	constexpr int kPer = kTiles / kThreads;
	VisTile mine[kPer];
	int mcls[kPer];
	__shared__ uint32_t classItems[kCostClasses], classCursor[kCostClasses];
	for (int c = threadIdx.x; c < kCostClasses; c += kThreads) classItems[c] = 0u;
	uint32_t carry = 0, ccarry = 0;
#pragma unroll
	for (int it = 0; it < kPer; ++it) {
		if (it * kThreads >= nT) break;   // uniform: no tile left for any thread
		const int t = it * kThreads + threadIdx.x;
		uint32_t col = 0, v = 0, items = 0, nseg = 1, nreg = kQuadrants, cost = 0;
		uint32_t grp[kCountGroups];
#pragma unroll
		for (int g = 0; g < kCountGroups; ++g) {
			grp[g] = t < nT ? tileCount[static_cast<long long>(g) * nT + t] : 0u;
			col += grp[g]; }
		if (t < nT) {
			v = clipHist[t];
			cost = o.tileCost[t];
			nseg = vis_segments(b, col + v);
			nreg = vis_regions(b, cost, nseg);
			items = nseg * nreg;
			o.tileCost[t] = 0u; }
		uint32_t tot, ctot;
		const uint32_t ex = block_exclusive_scan(col, sh, tot);
		const uint32_t cex = block_exclusive_scan(v, sh, ctot);
		mcls[it] = cost_class(cost);
		if (t < nT) {
			o.tileOff[t] = carry + ex;
			if (o.tileCursor) {
				uint32_t run = carry + ex;
#pragma unroll
				for (int g = 0; g < kCountGroups; ++g) { o.tileCursor[static_cast<long long>(g) * nT + t] = run; run += grp[g]; }}
			o.clipOff[t] = ccarry + cex;
			if (o.clipCursor) o.clipCursor[t] = ccarry + cex;
			VisTile& d = mine[it];
			d.item0 = items;   // the item count until the ordering below
			d.frame = static_cast<uint32_t>(find_base(ftb, b.nFrames, t));
			d.ib = min(carry + ex, o.listCap);
			d.cb = min(ccarry + cex, o.clistCap);
			d.ninl = min(carry + ex + col, o.listCap) - d.ib;
			d.ntot = d.ninl + (min(ccarry + cex + v, o.clistCap) - d.cb);
			d.nreg = nreg;
			d.multiSeg = nseg > 1 ? 1u : 0u;
			atomicAdd(&classItems[mcls[it]], items); }
		carry += tot; ccarry += ctot; }
	// Work-item order: by the tiles' previous-frame quad work, heaviest
	// first in power-of-two classes (a counting sort), so that long items
	// start at once and short ones fill the end of the kernel.  Wave 0
	// scans the class counts from the heaviest class down.
	__syncthreads();
	if (threadIdx.x < 64) {
		const int lane = threadIdx.x, c = kCostClasses - 1 - lane;   // lane 0: the heaviest class
		const uint32_t n = c >= 0 ? classItems[c] : 0u;
		uint32_t x = n;
		for (int off = 1; off < 64; off <<= 1) {
			const uint32_t y = __shfl_up(x, off, 64);
			if (lane >= off) x += y; }
		if (c >= 0) classCursor[c] = x - n;
		if (lane == 63) classItems[0] = x; }   // the total, read after the barrier below
	__syncthreads();
	const uint32_t wcarry = classItems[0];
#pragma unroll
	for (int it = 0; it < kPer; ++it) {
		const int t = it * kThreads + threadIdx.x;
		if (t >= nT) continue;
		VisTile& d = mine[it];
		const uint32_t items = d.item0;
		const uint32_t item0 = atomicAdd(&classCursor[mcls[it]], items);
		for (uint32_t j = item0, je = min(j + items, o.itemCap); j < je; ++j) o.itemTile[j] = static_cast<uint32_t>(t);
		d.item0 = item0;
		o.tileDesc[t] = d; }
	if (threadIdx.x == 0) {
		o.tileOff[nT] = carry; o.clipOff[nT] = ccarry;
		o.ctr->pairs = carry; o.ctr->clipPairs = ccarry; o.ctr->workItems = min(wcarry, o.itemCap); o.ctr->workNext = 0; }}

__global__ __launch_bounds__(kScanThreads) void k_col_scan(BatchDesc b, const uint32_t* __restrict__ tileCount,
		const uint32_t* __restrict__ clipHist, ScanOut o) {
	col_scan_block<kScanThreads, kScanSplitTiles>(b, tileCount, clipHist, o); }

// Large batches scan in three launches instead of one block: per block of
// kScanThreads tiles the local prefixes and per-class item counts
// (k_scan_local), the blocks' bases, one wave per column (k_scan_parts),
// then every tile's offsets, cursors and items (k_scan_final).  Every
// choice of work-item order within a cost class is valid, so a tile's
// first item is taken from its block's LDS cursor of its class.

struct ScanPart {
	uint32_t col, clip;            // block totals, then the block's exclusive bases
	uint32_t cls[kCostClasses];    // items per cost class, then the block's offset within the class
	uint32_t pad[3];
};

__global__ __launch_bounds__(kScanThreads) void k_scan_local(BatchDesc b, const uint32_t* __restrict__ tileCount,
		const uint32_t* __restrict__ clipHist, uint32_t* __restrict__ tileCost, VisTile* __restrict__ tileDesc,
		ScanPart* __restrict__ part) {
	__shared__ uint32_t sh[kScanThreads / 64 + 1];
	__shared__ uint32_t classItems[kCostClasses];
	__shared__ int ftb[kMaxFrames + 1];
	load_tile_bases(ftb, b);
	for (int c = threadIdx.x; c < kCostClasses; c += kScanThreads) classItems[c] = 0u;
	__syncthreads();
	const int nT = b.nTiles;
	const int t = blockIdx.x * kScanThreads + threadIdx.x;
	uint32_t col = 0, v = 0;
	if (t < nT) {
#pragma unroll
		for (int g = 0; g < kCountGroups; ++g) col += tileCount[static_cast<long long>(g) * nT + t];
		v = clipHist[t]; }
	uint32_t tot, ctot;
	const uint32_t ex = block_exclusive_scan(col, sh, tot);
	const uint32_t cex = block_exclusive_scan(v, sh, ctot);
	if (t < nT) {
		const uint32_t cost = tileCost[t];
		tileCost[t] = 0u;
		const uint32_t nseg = vis_segments(b, col + v), nreg = vis_regions(b, cost, nseg);
		const int cls = cost_class(cost);
		atomicAdd(&classItems[cls], nseg * nreg);
		VisTile d;   // provisional: local prefixes, counts and the cost class
		d.item0 = nseg * nreg;
		d.frame = static_cast<uint32_t>(find_base(ftb, b.nFrames, t));
		d.ib = ex; d.cb = cex; d.ninl = col; d.ntot = v;
		d.nreg = nreg;
		d.multiSeg = (nseg > 1 ? 1u : 0u) | (static_cast<uint32_t>(cls) << 1);
		tileDesc[t] = d; }
	__syncthreads();
	ScanPart& P = part[blockIdx.x];
	for (int c = threadIdx.x; c < kCostClasses; c += kScanThreads) P.cls[c] = classItems[c];
	if (threadIdx.x == 0) { P.col = tot; P.clip = ctot; }}

// one block: each wave scans whole columns over the blocks (the list and
// clipped-list totals, and each cost class), heaviest class first; row nb
// of part receives the class bases
__global__ __launch_bounds__(kScanThreads) void k_scan_parts(ScanPart* __restrict__ part, int nb, int nT,
		uint32_t* __restrict__ tileOff, uint32_t* __restrict__ clipOff, uint32_t itemCap, Counters* __restrict__ ctr) {
	constexpr int nCols = 2 + kCostClasses;
	__shared__ uint32_t colTot[nCols];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	for (int k = wave; k < nCols; k += kScanThreads / 64) {
		uint32_t carry = 0;
		for (int i0 = 0; i0 < nb; i0 += 64) {
			const int i = i0 + lane;
			uint32_t* cell = i < nb ? (k == 0 ? &part[i].col : k == 1 ? &part[i].clip : &part[i].cls[k - 2]) : nullptr;
			const uint32_t v = cell ? *cell : 0u;
			uint32_t x = v;
			for (int o = 1; o < 64; o <<= 1) {
				const uint32_t y = __shfl_up(x, o, 64);
				if (lane >= o) x += y; }
			if (cell) *cell = carry + x - v;
			carry += __shfl(x, 63, 64); }
		if (lane == 0) colTot[k] = carry; }
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t run = 0;
		for (int c = kCostClasses - 1; c >= 0; --c) { part[nb].cls[c] = run; run += colTot[2 + c]; }
		tileOff[nT] = colTot[0]; clipOff[nT] = colTot[1];
		ctr->pairs = colTot[0]; ctr->clipPairs = colTot[1]; ctr->workItems = min(run, itemCap); ctr->workNext = 0; }}

__global__ __launch_bounds__(kScanThreads) void k_scan_final(BatchDesc b, const uint32_t* __restrict__ tileCount,
		uint32_t* __restrict__ tileCursor, uint32_t* __restrict__ tileOff, uint32_t* __restrict__ clipOff,
		uint32_t* __restrict__ clipCursor, uint32_t* __restrict__ itemTile, uint32_t itemCap, VisTile* __restrict__ tileDesc,
		uint32_t listCap, uint32_t clistCap, const ScanPart* __restrict__ part, int nb) {
	__shared__ uint32_t classCursor[kCostClasses];
	const ScanPart& P = part[blockIdx.x];
	for (int c = threadIdx.x; c < kCostClasses; c += kScanThreads) classCursor[c] = part[nb].cls[c] + P.cls[c];
	__syncthreads();
	const int nT = b.nTiles;
	const int t = blockIdx.x * kScanThreads + threadIdx.x;
	if (t >= nT) return;
	VisTile d = tileDesc[t];
	const uint32_t col = d.ninl, v = d.ntot, items = d.item0;
	const uint32_t off = P.col + d.ib, coff = P.clip + d.cb;
	tileOff[t] = off;
	uint32_t run = off;
#pragma unroll
	for (int g = 0; g < kCountGroups; ++g) {
		tileCursor[static_cast<long long>(g) * nT + t] = run;
		run += tileCount[static_cast<long long>(g) * nT + t]; }
	clipOff[t] = coff; clipCursor[t] = coff;
	const uint32_t item0 = atomicAdd(&classCursor[d.multiSeg >> 1], items);
	for (uint32_t j = item0, je = min(j + items, itemCap); j < je; ++j) itemTile[j] = static_cast<uint32_t>(t);
	d.item0 = item0;
	d.ib = min(off, listCap);
	d.cb = min(coff, clistCap);
	d.ninl = min(off + col, listCap) - d.ib;
	d.ntot = d.ninl + (min(coff + v, clistCap) - d.cb);
	d.multiSeg &= 1u;
	tileDesc[t] = d; }

// ---------------------------------------------------------------------------
// clipping: GPUBinImpl::ClipTriangles (rglv_gpu_impl.hxx:677-793),
// Sutherland-Hodgman against L, B, N, R, T, fan triangulation, per-tile
// counts (ForEachCoveredTile, :795-817).
// ---------------------------------------------------------------------------
// One wave per clipped triangle: lane c holds component c (x, y, z, w,
// then the varyings) of every vertex of the polygon, in an LDS column of
// its own, so each copy or interpolation of a vertex is one operation per
// lane; the wave's control flow, driven by the plane distances every lane
// reads alike, is uniform.  The arithmetic per component is the serial
// clipper's.
constexpr int kClipWaves = 1024;   // workgroups (one wave each) of the clipping launch

template <class P>
__global__ __launch_bounds__(64) void k_clip(const DrawDesc* __restrict__ draws, int nDraws,
		const uint32_t* __restrict__ clipq, uint32_t clipCap, float4* __restrict__ cvpos, float* __restrict__ cvvary,
		uint8_t* __restrict__ fcount, uint2* __restrict__ frect, uint32_t* __restrict__ clipHist, Counters* __restrict__ ctr,
		int varyStride) {
	constexpr int kComp = 4 + P::kVary;
	static_assert(kComp <= 64, "one lane per component");
	__shared__ float poly[2][kClipVerts][64];
	const int lane = threadIdx.x;
	const uint32_t nq = min(ctr->clipCount, clipCap);
	for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
		__syncthreads();
		const long long s = clipq[q];
		int lo = 0, hi = nDraws - 1;
		while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (draws[mid].triBase <= s) lo = mid; else hi = mid - 1; }
		const DrawDesc& d = draws[lo];
		if (d.program != P::kId) continue;   // another program's launch clips it (uniform)
		const long long n = s - d.triBase;
		const int iid = static_cast<int>(n / d.tris);
		const int t = static_cast<int>(n - static_cast<long long>(iid) * d.tris);
		// lanes 0..2 shade the triangle's vertices into rows 0..2
		if (lane < 3) {
			const int idx = d.elements ? d.indices[3 * t + lane] : 3 * t + lane;
			float pos[4], vo[P::kVary];
			P::vertex(d, iid, idx, pos, vo);
			for (int c = 0; c < 4; ++c) poly[0][lane][c] = pos[c];
			for (int j = 0; j < P::kVary; ++j) poly[0][lane][4 + j] = vo[j]; }
		__syncthreads();
		// lane k's distance of polygon vertex k to a clipping plane; every
		// lane reads vertex k's with readlane (a uniform value)
		auto lane_dist = [&](int buf, int plane) {
			if (lane >= kClipVerts) return 0.0F;
			const float w = poly[buf][lane][3];
			switch (plane) {
			case 0: return w + poly[buf][lane][0];
			case 1: return w + poly[buf][lane][1];
			case 2: return w + poly[buf][lane][2];
			case 3: return w - poly[buf][lane][0];
			default: return w - poly[buf][lane][1]; }};
		auto at = [](float v, int k) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k)); };
		// all three vertices outside one clipping plane: Sutherland-Hodgman
		// would emit an empty polygon for that plane
		bool rejected = false;
		for (int plane = 0; plane < 5; ++plane) {
			const float dk = lane_dist(0, plane);
			rejected |= at(dk, 0) < 0.0F && at(dk, 1) < 0.0F && at(dk, 2) < 0.0F; }
		if (rejected) { if (lane == 0) fcount[q] = 0; continue; }
		int cur = 0, na = 3;
		for (int plane = 0; plane < 5 && na > 0; ++plane) {
			const int nxt = cur ^ 1;
			const float dk = lane_dist(cur, plane);
			auto dist = [&](int k) { return at(dk, k); };
			bool hereIsInside = dist(0) >= 0.0F;
			int nb = 0;
			for (int h = 0; h < na; ++h) {
				const int nx = h + 1 == na ? 0 : h + 1;
				const bool nextIsInside = dist(nx) >= 0.0F;
				if (hereIsInside) { poly[nxt][nb][lane] = poly[cur][h][lane]; ++nb; }
				if (hereIsInside != nextIsInside) {
					const int from = hereIsInside ? h : nx;
					const int to = hereIsInside ? nx : h;
					const float fa = dist(from), fb = dist(to);
					const float dd = fa / (fa - fb);
					const float om = 1.0F - dd;
					poly[nxt][nb][lane] = om * poly[cur][from][lane] + dd * poly[cur][to][lane];
					++nb;
					hereIsInside = !hereIsInside; }}
			na = nb;
			cur = nxt; }
		if (na == 0) { if (lane == 0) fcount[q] = 0; continue; }
		// pdiv and viewport: lanes 0..3 (x, y, z, 1/w)
		for (int k = 0; k < na; ++k) {
			const float iw = oneover(poly[cur][k][3]);
			float v = poly[cur][k][lane < 4 ? lane : 0];
			if (lane == 0) v = (v * iw) * d.DS[0] + d.DO[0];
			else if (lane == 1) v = (v * iw) * d.DS[1] + d.DO[1];
			else if (lane == 2) v = v * iw;
			else if (lane == 3) v = iw;
			if (lane < 4) poly[cur][k][lane] = v; }
		const float d31x = poly[cur][2][0] - poly[cur][0][0], d31y = poly[cur][2][1] - poly[cur][0][1];
		const float d21x = poly[cur][1][0] - poly[cur][0][0], d21y = poly[cur][1][1] - poly[cur][0][1];
		const bool backfacing = (d31x*d21y - d31y*d21x) < 0.0F;
		bool willCull = true;
		if (backfacing) { if (!d.cullingEnabled || (d.cullFace & 2) == 0) willCull = false; }
		else { if (!d.cullingEnabled || (d.cullFace & 1) == 0) willCull = false; }
		if (willCull) { if (lane == 0) fcount[q] = 0; continue; }
		// the fan's vertices in output order (reversed when backfacing)
		auto src = [&](int k) { return backfacing ? na - 1 - k : k; };
		for (int k = 0; k < na; ++k) {
			const float v = poly[cur][src(k)][lane];
			if (lane < 4) reinterpret_cast<float*>(cvpos)[(static_cast<long long>(q) * kClipVerts + k) * 4 + lane] = v;
			else if (lane < kComp) cvvary[(static_cast<long long>(q) * kClipVerts + k) * varyStride + (lane - 4)] = v; }
		const int nf = na - 2;
		if (lane == 0) fcount[q] = static_cast<uint8_t>(nf);
		for (int j = 0; j < nf; ++j) {
			const int k0 = src(0), k1 = src(j + 1), k2 = src(j + 2);
			const int ix0 = cvtt(poly[cur][k0][0]), ix1 = cvtt(poly[cur][k1][0]), ix2 = cvtt(poly[cur][k2][0]);
			const int iy0 = cvtt(poly[cur][k0][1]), iy1 = cvtt(poly[cur][k1][1]), iy2 = cvtt(poly[cur][k2][1]);
			const int vminx = imax(imin(ix0, imin(ix1, ix2)), d.scissor[0]);
			const int vminy = imax(imin(iy0, imin(iy1, iy2)), d.scissor[1]);
			const int vmaxx = imin(wadd(imax(ix0, imax(ix1, ix2)), 1), d.scissor[2] - 1);
			const int vmaxy = imin(wadd(imax(iy0, imax(iy1, iy2)), 1), d.scissor[3] - 1);
			int tx0 = vminx / d.tileW, ty0 = vminy / d.tileH;
			int tx1 = vmaxx / d.tileW, ty1 = vmaxy / d.tileH;
			tx1 = imin(tx1, d.tilesX - 1); ty1 = imin(ty1, d.tilesY - 1);
			tx0 = imax(tx0, 0); ty0 = imax(ty0, 0);
			if (tx0 > tx1 || ty0 > ty1) { if (lane == 0) frect[static_cast<long long>(q) * kFanMax + j] = empty_rect(); continue; }
			if (lane == 0) frect[static_cast<long long>(q) * kFanMax + j] = pack_rect(tx0, ty0, tx1, ty1);
			// the fan triangle's tiles, shared out over the lanes
			const int w = tx1 - tx0 + 1, ntl = w * (ty1 - ty0 + 1);
			for (int i = lane; i < ntl; i += 64)
				atomicAdd(&clipHist[d.tileBase + (ty0 + i / w) * d.tilesX + tx0 + i % w], 1u); }}}

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
	long long vary[3];   // offset of each vertex's varyings, in floats (index times the batch's varying stride)
	bool clipped;
	int draw;
	uint32_t prim;
};

__device__ inline void inline_verts(const uint4 tv, const float4* __restrict__ vpos, TriVerts& tr, int varyStride = 0) {
	const uint32_t g[3] = { tv.x, tv.y, tv.z };
	for (int k = 0; k < 3; ++k) {
		const float4 p = vpos[g[k]];
		tr.X[k] = cvtt(16.0F * p.x);   // GPUTileImpl::DrawTriangles (:885-886)
		tr.Y[k] = cvtt(16.0F * p.y);
		tr.z[k] = p.z; tr.w[k] = p.w;
		tr.vary[k] = static_cast<long long>(g[k]) * varyStride; }
	tr.clipped = false; }

__device__ inline void clipped_verts(uint32_t fan, const float4* __restrict__ cvpos, TriVerts& tr, int varyStride = 0) {
	const uint32_t q = fan / kFanMax, j = fan - q * kFanMax;
	const long long v[3] = { static_cast<long long>(q) * kClipVerts, static_cast<long long>(q) * kClipVerts + j + 1,
	                         static_cast<long long>(q) * kClipVerts + j + 2 };
	for (int k = 0; k < 3; ++k) {
		const float4 p = cvpos[v[k]];
		tr.X[k] = cvtt(p.x * 16.0F);   // DrawClipped (:997-998)
		tr.Y[k] = cvtt(p.y * 16.0F);
		tr.z[k] = p.z; tr.w[k] = p.w;
		tr.vary[k] = v[k] * varyStride; }
	tr.clipped = true; }

struct RasterArgs {
	const DrawDesc* draws;
	const DrawLite* lite;       // the draws' compact table (the resolve reads it through the caches)
	const uint32_t* tileOff;    // inline records: tile t's pairs are [tileOff[t], tileOff[t + 1])
	uint32_t listCap;
	const uint4* tvert;
	const float4* vpos;
	const float* vvary;
	const uint32_t* clipq;
	const uint32_t* clipqOf;
	const float4* cvpos;
	const float* cvvary;
	int varyStride;             // floats per vertex in vvary and cvvary (the batch's largest varying count)
	const uint32_t* clipOff;    // clipped records: tile t's fans are [clipOff[t], clipOff[t + 1])
	uint32_t clistCap;
	uint32_t clipCap;
	Counters* ctr;
	unsigned long long* keys;   // per-pixel visibility key (smallest (depth, slot) of the frame)
	uint32_t* tileCost;         // per-tile quad work, for the next frame's region split
	Counters* hostCtr;          // pinned host copy of the counters, written by the last kernel (null: copied by the host)
};

constexpr unsigned long long kKeySentinel = ~0ull;

// visibility key of a covered fragment (see the header comment)
__device__ inline unsigned long long frag_key(float dz, uint32_t slot, bool leq) {
	return (static_cast<unsigned long long>(depth_order(dz)) << 32) |
	       (leq ? static_cast<unsigned long long>(~slot) : static_cast<unsigned long long>(slot)); }

// the key that a fragment must beat to pass the depth test against the
// clear depth: LESS needs depth < clear, LEQUAL depth <= clear (slots >= 1)
__device__ inline unsigned long long clear_key(const FrameDesc& f) {
	return (static_cast<unsigned long long>(depth_order(f.clearDepth)) << 32) | (f.depthFunc == 1 ? 0xffffffffull : 0ull); }

__device__ inline uint32_t key_slot(unsigned long long k, bool leq) {
	const uint32_t lo32 = static_cast<uint32_t>(k);
	return leq ? ~lo32 : lo32; }



// ---------------------------------------------------------------------------
// scatter + pair setup: each (triangle, tile) pair's rasteriser setup
// against the tile rect — VTriangleRasterizer (rglv_triangle.hxx:206-240)
// for inline triangles, TriangleRasterizer (:84-131) for clipped fans —
// stored as a 64-byte record at the pair's position in its tile's list
// (the per-tile append of BinTriangles, rglv_gpu_impl.hxx:651-661).  A
// tile's records are contiguous.  Their order within the tile is free: the
// visibility keys carry each triangle's submission slot.
// ---------------------------------------------------------------------------
struct alignas(16) VisRec {
	int32_t c1, c2, c3;
	int32_t dx12, dy12, dx23, dy23, dx31, dy31;
	uint32_t minxy;       // minx | miny << 16 (even, tile-clamped)
	uint32_t nq;          // quads across | quads down << 16
	float scale;
	float z0, z1, z2;
	uint32_t slot;
};
static_assert(sizeof(VisRec) == 64, "VisRec is four 16-byte loads");

__device__ inline VisRec make_rec(const TriSetup& t, const TriVerts& tr, uint32_t slot) {
	VisRec v;
	v.c1 = t.c1; v.c2 = t.c2; v.c3 = t.c3;
	v.dx12 = t.dx12; v.dy12 = t.dy12; v.dx23 = t.dx23; v.dy23 = t.dy23; v.dx31 = t.dx31; v.dy31 = t.dy31;
	v.minxy = static_cast<uint32_t>(t.minx & 0xffff) | (static_cast<uint32_t>(t.miny) << 16);
	v.nq = static_cast<uint32_t>(t.nqx) | (static_cast<uint32_t>(t.nqy) << 16);
	v.scale = t.scale; v.z0 = tr.z[0]; v.z1 = tr.z[1]; v.z2 = tr.z[2]; v.slot = slot;
	return v; }

// (triangle, tile) pairs of a tile, inline and clipped, as the scan counted them
__device__ inline uint32_t tile_pairs(const uint32_t* tileOff, const uint32_t* clipOff, int t) {
	return (tileOff[t + 1] - tileOff[t]) + (clipOff[t + 1] - clipOff[t]); }

// the rect a tile rasterises a draw's triangles in: the tile rect
// (GPU::TileRect, rglv_gpu.hxx:291-299), intersected with the scissor rect
// when the scissor test is on (DrawTriangles / DrawClipped,
// rglv_gpu_impl.hxx:848-849, 953-954)
__device__ inline Rect draw_tile_rect(const DrawLite& d, int tx, int ty) {
	Rect r;
	r.x0 = tx * d.tileW; r.y0 = ty * d.tileH;
	r.x1 = imin(r.x0 + d.tileW, d.fW); r.y1 = imin(r.y0 + d.tileH, d.fH);
	if (d.scissorOn) {
		r.x0 = imax(r.x0, d.scissor[0]); r.y0 = imax(r.y0, d.scissor[1]);
		r.x1 = imin(r.x1, d.scissor[2]); r.y1 = imin(r.y1, d.scissor[3]); }
	return r; }


// Batches of at most kFusedScanTiles tiles scan inside k_scatter: its
// block 0 lays out the tiles' lists and visibility work items (the scan
// kernel's work), and every other block computes the list offsets it
// needs from the pair counts itself, reserving positions from counters
// zeroed at the frame's start.  That takes a launch and its latency off
// the binning stage.
constexpr int kFusedScanTiles = 512;

struct ScatterArgs {
	const uint2* trect; const uint4* tvert; const float4* vpos;
	uint32_t* cursor;             // per (count group, tile) next free list position (fused: position within the group's block)
	VisRec* rec; uint32_t listCap;
	// clipped fans (blocks [units, units + clipBlocks))
	const uint32_t* clipq; const uint8_t* fcount; const uint2* frect; const float4* cvpos; uint32_t clipCap;
	uint32_t* clipCursor; VisRec* crec; uint32_t clistCap; int clipBlocks;
	// key reset of tiles cut into several segments (the remaining blocks)
	const uint32_t* tileOff; const uint32_t* clipOff; unsigned long long* keys;
	Counters* ctr;
	// fused scan
	int fused;
	const uint32_t* tileCount; const uint32_t* clipHist;
	ScanOut scan;
};

// the exclusive scan of v over the batch's tiles (at most
// kFusedScanTiles) into out, by one block
__device__ inline void block_tile_scan(int nT, const uint32_t* __restrict__ v, uint32_t* out, uint32_t* sh) {
	uint32_t carry = 0;
	for (int t0 = 0; t0 < nT; t0 += blockDim.x) {   // uniform
		const int t = t0 + threadIdx.x;
		const uint32_t x = t < nT ? v[t] : 0u;
		uint32_t tot;
		const uint32_t ex = block_exclusive_scan(x, sh, tot);
		if (t < nT) out[t] = carry + ex;
		carry += tot; }
	__syncthreads(); }

// the tiles of a packed tile rect (none when empty)
__device__ inline uint32_t rect_tiles(uint2 rc) {
	const int tx0 = rc.x & 0xffffu, tx1 = rc.x >> 16, ty0 = rc.y & 0xffffu, ty1 = rc.y >> 16;
	return tx0 <= tx1 && ty0 <= ty1 ? static_cast<uint32_t>((tx1 - tx0 + 1) * (ty1 - ty0 + 1)) : 0u; }

// clipped fans into their tiles' clipped lists (ForEachCoveredTile,
// rglv_gpu_impl.hxx:795-817).  A batch of the block's clipped triangles,
// one per thread, is expanded into its (fan triangle, tile) pairs, which
// the block's threads share out: a large clipped triangle's pairs take
// their list positions (an atomic each) in parallel.
__device__ inline void clip_scatter(const BatchDesc& b, const DrawLite* dl, int cb, const ScatterArgs& a, uint32_t* cbase,
                                    uint32_t* sh, uint32_t* pre) {
	const uint32_t nq = min(a.ctr->clipCount, a.clipCap);
	if (nq == 0) return;   // uniform
	// fused: the tiles' clipped-list offsets, reserved within from zeroed counters
	if (a.fused) block_tile_scan(b.nTiles, a.clipHist, cbase, sh);
	const uint32_t step = static_cast<uint32_t>(a.clipBlocks) * blockDim.x;
	for (uint32_t q0 = static_cast<uint32_t>(cb) * blockDim.x; q0 < nq; q0 += step) {   // uniform
		const uint32_t q = q0 + threadIdx.x;
		uint32_t n = 0;
		if (q < nq) {
			const int nf = a.fcount[q];
			for (int j = 0; j < nf; ++j) n += rect_tiles(a.frect[static_cast<long long>(q) * kFanMax + j]); }
		uint32_t tot;
		const uint32_t ex = block_exclusive_scan(n, sh, tot);
		pre[threadIdx.x] = ex;
		__syncthreads();
		for (uint32_t i = threadIdx.x; i < tot; i += blockDim.x) {
			// the pair's triangle: the last thread whose prefix is at or below i
			int lo = 0, hi = static_cast<int>(blockDim.x) - 1;
			while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (pre[mid] <= i) lo = mid; else hi = mid - 1; }
			const uint32_t qq = q0 + static_cast<uint32_t>(lo);
			uint32_t r = i - pre[lo];
			// its fan triangle and tile
			int j = 0;
			uint2 rc = a.frect[static_cast<long long>(qq) * kFanMax];
			for (uint32_t nt = rect_tiles(rc); r >= nt; nt = rect_tiles(rc)) { r -= nt; ++j; rc = a.frect[static_cast<long long>(qq) * kFanMax + j]; }
			const int tx0 = rc.x & 0xffffu, tx1 = rc.x >> 16, ty0 = rc.y & 0xffffu;
			const int w = tx1 - tx0 + 1;
			const int ty = ty0 + static_cast<int>(r) / w, tx = tx0 + static_cast<int>(r) % w;
			const uint32_t s = a.clipq[qq];
			const DrawLite& d = dl[find_draw_ll(dl, b.nDraws, s)];
			const uint32_t nTri = static_cast<uint32_t>(d.tris) * static_cast<uint32_t>(d.instances);
			const uint32_t slot = d.slotBase + nTri + static_cast<uint32_t>(s - d.triBase) * kFanMax + static_cast<uint32_t>(j);
			TriVerts tr;
			clipped_verts(qq * kFanMax + j, a.cvpos, tr);
			const int tile = d.tileBase + ty * d.tilesX + tx;
			const uint32_t pos = (a.fused ? cbase[tile] : 0u) + atomicAdd(&a.clipCursor[tile], 1u);
			if (pos < a.clistCap) {
				TriSetup t;
				setup_scalar(draw_tile_rect(d, tx, ty), tr.X, tr.Y, t);
				a.crec[pos] = make_rec(t, tr, slot); }
			else atomicOr(&a.ctr->overflow, 4u); }
		__syncthreads(); }}

// tiles cut into several segments merge their regions with a 64-bit
// atomic min, so their keys start the frame at the sentinel; tiles of one
// segment are overwritten whole by plain stores
__device__ inline void reset_keys(const BatchDesc& b, const int* ftb, int rb, int nrb, const ScatterArgs& a) {
	for (int tile = rb; tile < b.nTiles; tile += nrb) {
		uint32_t pairs;
		if (a.fused) {
			pairs = a.clipHist[tile];
			for (int g = 0; g < kCountGroups; ++g) pairs += a.tileCount[static_cast<long long>(g) * b.nTiles + tile]; }
		else pairs = tile_pairs(a.tileOff, a.clipOff, tile);
		if (vis_segments(b, pairs) == 1u) continue;
		const FrameDesc& f = b.frames[find_base(ftb, b.nFrames, tile)];
		const Rect R = tile_rect(f, tile - f.tileBase);
		const int rw = R.x1 - R.x0, n = rw * (R.y1 - R.y0);
		for (int i = threadIdx.x; i < n; i += blockDim.x) {
			const int ly = i / rw, lx = i - ly * rw;
			a.keys[f.keyBase + static_cast<long long>(R.y0 + ly) * f.W + R.x0 + lx] = kKeySentinel; }}}

// With the fused scan, block 0 is the scan's and the others shift by one.
// Blocks [0, units): one binning unit each.  The unit counts its pairs per
// tile in LDS, reserves one block of list positions per tile it touches
// from its count group's cursor, then each lane walks its triangle's
// tiles, takes a position from the unit's LDS cursor of the tile and
// writes the pair's setup record.  Blocks [units, units + clipBlocks): the
// clipped fans.  The rest: key resets.
__global__ __launch_bounds__(kUnitThreads) void k_scatter(BatchDesc b, const DrawLite* __restrict__ dl, ScatterArgs a) {
	// dynamic LDS: the unit's tile-window counts, then write cursors; the
	// draw table is read through the caches
	extern __shared__ uint4 dyn_lds[];
	__shared__ int ftb[kMaxFrames + 1];
	__shared__ uint32_t fbase[kFusedScanTiles];   // fused: per tile, this block's list (or clipped-list) base
	__shared__ uint32_t fsh[kUnitThreads / 64 + 1];
	__shared__ uint32_t cpre[kUnitThreads];      // clip blocks: the batch's pair prefixes
	uint32_t* base = reinterpret_cast<uint32_t*>(dyn_lds);
	int u = blockIdx.x;
	int nblk = static_cast<int>(gridDim.x);
	if (a.fused) {
		if (u == 0) { col_scan_block<kUnitThreads, kFusedScanTiles>(b, a.tileCount, a.clipHist, a.scan); return; }
		--u; --nblk; }
	if (u >= b.units + a.clipBlocks) {
		load_tile_bases(ftb, b);
		__syncthreads();
		reset_keys(b, ftb, u - b.units - a.clipBlocks, nblk - b.units - a.clipBlocks, a);
		return; }
	if (u >= b.units) {
		clip_scatter(b, dl, u - b.units, a, fbase, fsh, cpre);
		return; }
	const TileWindow win = unit_window(dl, b, u);
	if (win.n > b.unitWindow) return;   // flagged by the setup
	const int grp = u % kCountGroups;
	if (a.fused) {
		// per tile: its list offset (the scan of all groups' counts) plus
		// the counts of the groups below this unit's
		uint32_t carry = 0;
		for (int t0 = 0; t0 < b.nTiles; t0 += kUnitThreads) {   // uniform
			const int t = t0 + threadIdx.x;
			uint32_t col = 0, below = 0;
			if (t < b.nTiles) {
#pragma unroll
				for (int g = 0; g < kCountGroups; ++g) {
					const uint32_t c = a.tileCount[static_cast<long long>(g) * b.nTiles + t];
					col += c;
					below += g < grp ? c : 0u; }}
			uint32_t tot;
			const uint32_t ex = block_exclusive_scan(col, fsh, tot);
			if (t < b.nTiles) fbase[t] = carry + ex + below;
			carry += tot; }}
	for (int t = threadIdx.x; t < win.n; t += blockDim.x) base[t] = 0u;   // entry t: batch tile win.lo + t
	__syncthreads();
	const long long s0 = static_cast<long long>(u) * b.unitTris;
	for (int r = threadIdx.x; r < b.unitTris; r += blockDim.x) {
		const long long s = s0 + r;
		if (s >= b.totalTris) break;
		const uint2 rc = a.trect[s];
		const int tx0 = rc.x & 0xffffu, tx1 = rc.x >> 16, ty0 = rc.y & 0xffffu, ty1 = rc.y >> 16;
		if (tx0 > tx1 || ty0 > ty1) continue;
		const DrawLite& d = dl[find_draw_ll(dl, b.nDraws, s)];
		for (int ty = ty0; ty <= ty1; ++ty)
			for (int tx = tx0; tx <= tx1; ++tx) atomicAdd(&base[d.tileBase - win.lo + ty * d.tilesX + tx], 1u); }
	__syncthreads();
	uint32_t* cur = a.cursor + static_cast<long long>(grp) * b.nTiles + win.lo;
	for (int t = threadIdx.x; t < win.n; t += blockDim.x) {
		const uint32_t c = base[t];
		if (c) base[t] = (a.fused ? fbase[win.lo + t] : 0u) + atomicAdd(&cur[t], c); }
	__syncthreads();
	bool ovf = false;
	for (int r = threadIdx.x; r < b.unitTris; r += blockDim.x) {
		const long long s = s0 + r;
		if (s >= b.totalTris) break;
		const uint2 rc = a.trect[s];
		const int tx0 = rc.x & 0xffffu, tx1 = rc.x >> 16, ty0 = rc.y & 0xffffu, ty1 = rc.y >> 16;
		if (tx0 > tx1 || ty0 > ty1) continue;
		const DrawLite& d = dl[find_draw_ll(dl, b.nDraws, s)];
		TriVerts tr;
		inline_verts(a.tvert[s], a.vpos, tr);
		const uint32_t slot = d.slotBase + static_cast<uint32_t>(s - d.triBase);
		for (int ty = ty0; ty <= ty1; ++ty)
			for (int tx = tx0; tx <= tx1; ++tx) {
				const uint32_t pos = atomicAdd(&base[d.tileBase - win.lo + ty * d.tilesX + tx], 1u);
				if (pos < a.listCap) {
					TriSetup t;
					setup_vector(draw_tile_rect(d, tx, ty), tr.X, tr.Y, t);
					a.rec[pos] = make_rec(t, tr, slot); }
				else ovf = true; }}
	if (ovf) atomicOr(&a.ctr->overflow, 1u); }

// ---------------------------------------------------------------------------
// visibility: one workgroup per (tile, segment, region) work item, the
// region a 32x32 quadrant or its half or quarter.  Each item holds its
// region's keys in LDS, takes up to segPairs of its tile's setup records
// (inline records, then clipped fans, in any order; set up against the
// full tile rectangle, the reference's edge start values), rasterises the
// triangles' quads that fall in the region with ds_min_u64, and merges the
// region into the per-pixel key buffer: a plain store when the tile is one
// segment, a 64-bit atomic min otherwise.
// ---------------------------------------------------------------------------
struct alignas(16) VisSetup {
	int32_t c1, c2, c3;   // edge values at the triangle's first quad in the quadrant
	int32_t x0, y0;       // that quad's top-left pixel, quadrant-relative
	int32_t dx12, dy12, dx23, dy23, dx31, dy31;
	uint32_t nx;          // quads across inside the quadrant
	float invNx;          // 1 / nx: row = floor((k + 0.5) / nx), exact for k < 2^16
	float scale;
	float z0, z1, z2;
	uint32_t slot;
	uint32_t small;       // 1: every quad step fits 23 bits, 24-bit multiplies are exact
	uint32_t pad;
};
static_assert(sizeof(VisSetup) % 16 == 0, "VisSetup rows are read with aligned 16-byte loads");


constexpr int kQuadWindow = 2048;   // quad items mapped to their triangle per pass

__device__ inline VisRec vis_record(const VisRec* __restrict__ rec, const VisRec* __restrict__ crec, uint32_t ib, uint32_t cb,
                                    uint32_t ninl, uint32_t p) {
	return p < ninl ? rec[ib + p] : crec[cb + (p - ninl)]; }

// one quad of a triangle: the four lanes (c, c+dy, c+dx, c+dx+dy)
// (rglv_triangle.hxx:253-258) with their depth keys merged by a 64-bit LDS
// min; q1..q3 are the edge values at the quad's top-left pixel (lx, ly),
// quadrant-relative
__device__ inline void vis_quad(unsigned long long* keys, const VisSetup& v, int32_t q1, int32_t q2, int32_t q3,
                                int lx, int ly, bool leq, uint32_t& frags) {
#pragma unroll
	for (int l = 0; l < 4; ++l) {
		const int32_t e1 = (l & 2) ? wadd((l & 1) ? wadd(q1, v.dx12) : q1, (l & 1) ? v.dy12 : v.dx12) : ((l & 1) ? wadd(q1, v.dy12) : q1);
		const int32_t e2 = (l & 2) ? wadd((l & 1) ? wadd(q2, v.dx23) : q2, (l & 1) ? v.dy23 : v.dx23) : ((l & 1) ? wadd(q2, v.dy23) : q2);
		const int32_t e3 = (l & 2) ? wadd((l & 1) ? wadd(q3, v.dx31) : q3, (l & 1) ? v.dy31 : v.dx31) : ((l & 1) ? wadd(q3, v.dy31) : q3);
		const bool cov = (e1 | e2 | e3) >= 0;
		frags += cov ? 1u : 0u;
		const float bx = itof(e2) * v.scale;
		const float bz = itof(e1) * v.scale;
		const float by = (1.0F - bx) - bz;
		const float dz = (bx*v.z0 + by*v.z1) + bz*v.z2;
		if (cov && dz == dz) atomicMin(&keys[(ly + (l >> 1)) * kSubDim + (lx + (l & 1))], frag_key(dz, v.slot, leq)); }}

// one lane of a quad of a triangle setup: perspective-correct barycentrics
// BP and depth (TriangleProgram::Render, rglv_gpu_impl.hxx:169-191)
struct LaneBary { float bp0, bp1, bp2, depth; };

__device__ inline LaneBary lane_bary(const TriSetup& t, const TriVerts& tr, int32_t e1, int32_t e2, const uint32_t* rcpTab) {
	LaneBary r;
	const float b0 = itof(e2) * t.scale;
	const float b2 = itof(e1) * t.scale;
	const float b1 = (1.0F - b0) - b2;
	r.depth = (b0*tr.z[0] + b1*tr.z[1]) + b2*tr.z[2];
	const float fragW = oneover_tab((b0*tr.w[0] + b1*tr.w[1]) + b2*tr.w[2], rcpTab);
	r.bp0 = (tr.w[0] * b0) * fragW;
	r.bp2 = (tr.w[2] * b2) * fragW;
	r.bp1 = (1.0F - r.bp0) - r.bp2;
	return r; }

__device__ inline float lane_interp(const LaneBary& b, const float* v0, const float* v1, const float* v2, int k) {
	return (b.bp0*v0[k] + b.bp1*v1[k]) + b.bp2*v2[k]; }

// gl_FragCoord of lane l of the quad at (qx, qy) in a target h pixels
// high: the vector rasteriser's form (rglv_triangle.hxx:297) for inline
// triangles, the scalar one's (:155) for clipped ones (its y is one higher
// on the quad's top row)
__device__ inline FragCoord frag_coord(int qx, int qy, int l, int h, bool clipped, float depth) {
	FragCoord fc;
	fc.x = (static_cast<float>(qx) + 0.5F) + ((l & 1) ? 1.0F : 0.0F);
	if (clipped) fc.y = (static_cast<float>(h - qy) - 0.5F) + ((l >> 1) ? 0.0F : 1.0F);
	else fc.y = ((static_cast<float>(h) - 0.5F) - static_cast<float>(qy)) - ((l >> 1) ? 1.0F : 0.0F);
	fc.depth = depth;
	return fc; }

// shade one pixel (lane `lane` of quad (kx, ky)) of a winning triangle;
// h = the target's height (gl_FragCoord)
template <class P>
__device__ inline void shade_pixel(const DrawDesc& d, const TriSetup& t, const TriVerts& tr, const float* v0, const float* v1,
                                   const float* v2, int kx, int ky, int lane, const uint32_t* rcpTab, float color[3], float& depth,
                                   int h = 0) {
	const int32_t q1 = wadd(wadd(t.c1, wmul(kx, wmul(t.dy12, 2))), wmul(ky, wmul(t.dx12, 2)));
	const int32_t q2 = wadd(wadd(t.c2, wmul(kx, wmul(t.dy23, 2))), wmul(ky, wmul(t.dx23, 2)));
	// lane offsets of the rasteriser's quad: (c, c+dy, c+dx, c+dx+dy) (rglv_triangle.hxx:253-258)
	const int32_t e1 = (lane == 0) ? q1 : (lane == 1) ? wadd(q1, t.dy12) : (lane == 2) ? wadd(q1, t.dx12) : wadd(wadd(q1, t.dx12), t.dy12);
	const int32_t e2 = (lane == 0) ? q2 : (lane == 1) ? wadd(q2, t.dy23) : (lane == 2) ? wadd(q2, t.dx23) : wadd(wadd(q2, t.dx23), t.dy23);
	const LaneBary own = lane_bary(t, tr, e1, e2, rcpTab);
	depth = own.depth;
	float vv[P::kVary];
#pragma unroll
	for (int k = 0; k < P::kVary; ++k) vv[k] = lane_interp(own, v0, v1, v2, k);
	const int qx = t.minx + 2 * kx, qy = t.miny + 2 * ky;
	const FragCoord fc = frag_coord(qx, qy, lane, h, tr.clipped, own.depth);
	int lod = 0;
	if constexpr (P::kUV >= 0 || P::kUV == kUVFromFragCoord) {
		// the quad's LOD comes from the texture coordinates of lanes 0 and 1
		const LaneBary l0 = lane_bary(t, tr, q1, q2, rcpTab);
		const LaneBary l1 = lane_bary(t, tr, wadd(q1, t.dy12), wadd(q2, t.dy23), rcpTab);
		float a0[P::kVary], a1[P::kVary];
#pragma unroll
		for (int k = 0; k < P::kVary; ++k) { a0[k] = lane_interp(l0, v0, v1, v2, k); a1[k] = lane_interp(l1, v0, v1, v2, k); }
		float u0, w0, u1, w1;
		lod_texcoord<P>(d, a0, frag_coord(qx, qy, 0, h, tr.clipped, l0.depth), u0, w0);
		lod_texcoord<P>(d, a1, frag_coord(qx, qy, 1, h, tr.clipped, l1.depth), u1, w1);
		lod = texture_lod(d.tu0, u0, u1, w0, w1); }
	float c4[4];
	(void)P::fragment(d, vv, lod, fc, c4);
	color[0] = c4[0]; color[1] = c4[1]; color[2] = c4[2]; }

// resolve shaders: one program known at compile time, or a switch over
// the installed programs for frames that mix them.  load() takes the
// triangle's vertex varyings as soon as its vertex ids are known, so their
// gathers overlap the triangle setup; shade() then runs the program.
template <class P>
struct ShadeOne {
	struct Vary { float v[3][P::kVary]; };
	__device__ static void load(const float* __restrict__ vary, const TriVerts& tr, Vary& out) {
#pragma unroll
		for (int i = 0; i < 3; ++i) {
			const float* src = vary + tr.vary[i];
#pragma unroll
			for (int k = 0; k < P::kVary; ++k) out.v[i][k] = src[k]; }}
	__device__ static void shade(const DrawDesc& d, const TriSetup& t, const TriVerts& tr, const Vary& vy,
	                             int kx, int ky, int lane, const uint32_t* rcpTab, float color[3], float& depth, int h) {
		shade_pixel<P>(d, t, tr, vy.v[0], vy.v[1], vy.v[2], kx, ky, lane, rcpTab, color, depth, h); }};

// the program of the draw, chosen at run time over the program list
struct ShadeAny {
	struct Vary { const float* v[3]; };
	__device__ static void load(const float* __restrict__ vary, const TriVerts& tr, Vary& out) {
		for (int i = 0; i < 3; ++i) out.v[i] = vary + tr.vary[i]; }
	struct Call {
		const DrawDesc& d; const TriSetup& t; const TriVerts& tr; const Vary& vy;
		int kx, ky, lane; const uint32_t* rcpTab; float* color; float& depth; int h;
		template <class P> __device__ void run() {
			shade_pixel<P>(d, t, tr, vy.v[0], vy.v[1], vy.v[2], kx, ky, lane, rcpTab, color, depth, h); }};
	__device__ static void shade(const DrawDesc& d, const TriSetup& t, const TriVerts& tr, const Vary& vy,
	                             int kx, int ky, int lane, const uint32_t* rcpTab, float color[3], float& depth, int h) {
		Call c{ d, t, tr, vy, kx, ky, lane, rcpTab, color, depth, h };
		dispatch_program(d.program, c, AllPrograms{}); }};

// ---------------------------------------------------------------------------
// resolve: one thread per pixel.  The pixel's winning slot is shaded by
// recomputing its triangle's quad exactly as the reference's rasteriser
// produced it in the pixel's tile (tile-clamped edge start, barycentric
// scale, texture LOD from lanes 0 and 1), then post-shaded and
// sRGB-encoded (FilterTile, rglr_algorithm.hxx:31-104).  Colour, depth
// and primitive id are written once; the key is reset for the next frame.
// ---------------------------------------------------------------------------
// 256 threads = 16 x 16 pixels, a wave 16 x 4: compact blocks, whose
// pixels share more of their winning triangles' vertices and texels than
// 64 x 4 strips do (measured faster)
constexpr int kResolveW = 16, kResolveH = 16;

// the winner of pixel (x, y) drawn by draw d (table entry l), shaded:
// colour before the post program, depth and primitive id
template <class S>
__device__ inline void resolve_draw(const FrameDesc& f, const RasterArgs& a, const DrawLite& d, const DrawDesc& dd,
                                    int x, int y, uint32_t slot, const uint32_t* rcpTab, float col[3], float& dep, uint32_t& prim) {
	const uint32_t rel = slot - d.slotBase;
	const uint32_t nTri = static_cast<uint32_t>(d.tris) * static_cast<uint32_t>(d.instances);
	const int tx = x / f.tileW, ty = y / f.tileH;
	Rect R;
	R.x0 = tx * f.tileW; R.y0 = ty * f.tileH;
	R.x1 = imin(R.x0 + f.tileW, f.W); R.y1 = imin(R.y0 + f.tileH, f.H);
	TriVerts tr;
	TriSetup t;
	uint32_t srcTri;
	typename S::Vary vy;
	if (rel < nTri) {
		srcTri = rel;
		inline_verts(a.tvert[d.triBase + rel], a.vpos, tr, a.varyStride);
		S::load(a.vvary, tr, vy);
		setup_vector(R, tr.X, tr.Y, t); }
	else {
		const uint32_t fslot = rel - nTri;
		srcTri = fslot / kFanMax;
		const uint32_t j = fslot - srcTri * kFanMax;
		const uint32_t cq = a.clipqOf[d.triBase + srcTri];
		clipped_verts(cq * kFanMax + j, a.cvpos, tr, a.varyStride);
		S::load(a.cvvary, tr, vy);
		setup_scalar(R, tr.X, tr.Y, t); }
	const int kx = ((x & ~1) - t.minx) >> 1, ky = ((y & ~1) - t.miny) >> 1;
	const int lane = (x & 1) | ((y & 1) << 1);
	S::shade(dd, t, tr, vy, kx, ky, lane, rcpTab, col, dep, f.H);
	prim = d.primBase + srcTri; }

// the winner of pixel (x, y), shaded (clear values where nothing was
// drawn).  A batch of one draw reads that draw's table entries at a fixed
// index, and a wave whose pixels share one draw reads them by one lane's
// index: from restrict tables these are scalar loads, so the draw's
// uniforms and texture descriptors sit in scalar registers before the
// per-pixel gathers return.
template <class S>
__device__ inline bool resolve_key(const BatchDesc& b, const FrameDesc& f, const RasterArgs& a, const DrawLite* dl,
                                   const DrawDesc* draws, int x, int y, unsigned long long k, const uint32_t* rcpTab,
                                   float col[3], float& dep, uint32_t& prim) {
	col[0] = f.clearColor[0]; col[1] = f.clearColor[1]; col[2] = f.clearColor[2];
	dep = f.clearDepth;
	prim = 0xffffffffu;
	if (!(k < clear_key(f))) return false;
	const uint32_t slot = key_slot(k, f.depthFunc == 1);
	if (b.nDraws == 1) {
		const DrawLite d = dl[0];
		const DrawDesc dd = draws[0];
		resolve_draw<S>(f, a, d, dd, x, y, slot, rcpTab, col, dep, prim); }
	else {
		// a wave whose pixels share one draw reads it by one lane's index
		const int di = find_draw_slot(dl, b.nDraws, slot);
		const int d0 = __builtin_amdgcn_readfirstlane(di);
		if (__all(di == d0)) {
			const DrawLite d = dl[d0];
			const DrawDesc dd = draws[d0];
			resolve_draw<S>(f, a, d, dd, x, y, slot, rcpTab, col, dep, prim); }
		else resolve_draw<S>(f, a, dl[di], draws[di], x, y, slot, rcpTab, col, dep, prim); }
	return true; }

template <class S>
__device__ inline bool resolve_pixel(const BatchDesc& b, const FrameDesc& f, const RasterArgs& a, const DrawLite* dl, int x, int y,
                                     const uint32_t* rcpTab, float col[3], float& dep, uint32_t& prim) {
	return resolve_key<S>(b, f, a, dl, a.draws, x, y, a.keys[f.keyBase + static_cast<long long>(y) * f.W + x], rcpTab, col, dep, prim); }

// StoreTrueColor's ShadeCanvas + sRGB encode of one pixel (FilterTile,
// rglr_algorithm.hxx:31-104)
__device__ inline uint32_t post_pixel(const FrameDesc& f, const float col[3], int x, int y, const uint32_t* srgbTab) {
	float c[3] = { col[0], col[1], col[2] };
	if (f.postProgram != 1)
		post_shade(f.postProgram, f.exposure, c, f.W, f.H, (x / f.tileW) * f.tileW, (y / f.tileH) * f.tileH, x, y);
	return f.gamma ? ((to_srgb8_tab(c[0], srgbTab) << 16) | (to_srgb8_tab(c[1], srgbTab) << 8) | to_srgb8_tab(c[2], srgbTab))
	               : ((to_linear8(c[0]) << 16) | (to_linear8(c[1]) << 8) | to_linear8(c[2])); }

// resolve one pixel from its key and store colour, depth and primitive id
template <class S>
__device__ inline void resolve_store(const BatchDesc& b, const FrameDesc& f, const RasterArgs& a, const DrawLite* dl,
                                     const DrawDesc* draws, int x, int y, unsigned long long k, const uint32_t* rcpTab,
                                     const uint32_t* srgbTab) {
	float col[3], dep;
	uint32_t prim;
	resolve_key<S>(b, f, a, dl, draws, x, y, k, rcpTab, col, dep, prim);
	const long long o = static_cast<long long>(y) * f.W + x;
	// the outputs are not read again by this frame: streaming stores
	if (f.outColor) __builtin_nontemporal_store(post_pixel(f, col, x, y, srgbTab), &f.outColor[o]);
	if (f.outDepth) __builtin_nontemporal_store(dep, &f.outDepth[o]);
	if (f.outPrim) __builtin_nontemporal_store(prim, &f.outPrim[o]); }

// a resolve block's frame and pixel: blocks are numbered across the batch,
// each frame's blocks row-major over kResolveW x kResolveH pixel blocks
struct BlockPixel { int frame, x, y; };

__device__ inline BlockPixel block_pixel(const BatchDesc& b, const int* fbb, int blk = static_cast<int>(blockIdx.x)) {
	BlockPixel r;
	r.frame = find_base(fbb, b.nFrames, blk);
	const FrameDesc& f = b.frames[r.frame];
	const int lb = blk - f.blockBase;
	const int by = lb / f.blocksX, bx = lb - by * f.blocksX;
	r.x = bx * kResolveW + static_cast<int>(threadIdx.x % kResolveW);
	r.y = by * kResolveH + static_cast<int>(threadIdx.x / kResolveW);
	return r; }

// The resolve keeps no LDS: the draw table, the rcpps and sRGB tables and
// the frame table are read through the scalar and vector caches, so a
// workgroup starts on its pixels at once (it runs for a few microseconds).
template <class S>
__global__ __launch_bounds__(256) void k_resolve(BatchDesc b, RasterArgs a, const FrameDesc* __restrict__ frames,
                                                  const DrawLite* __restrict__ lite, const DrawDesc* __restrict__ draws) {
	// the frame and draw tables are read-only (restrict): uniform reads of
	// them are scalar loads
	const int blk = static_cast<int>(blockIdx.x);
	int lo = 0, hi = b.nFrames - 1;   // last frame whose first block is at or before blk
	while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (frames[mid].blockBase <= blk) lo = mid; else hi = mid - 1; }
	const FrameDesc f = frames[lo];
	const int lb = blk - f.blockBase;
	const int by = lb / f.blocksX, bx = lb - by * f.blocksX;
	const int x = bx * kResolveW + static_cast<int>(threadIdx.x % kResolveW);
	const int y = by * kResolveH + static_cast<int>(threadIdx.x / kResolveW);
	// frames of the in-order tile path store their own pixels
	if (x < f.W && y < f.H && !f.inOrder)
		resolve_store<S>(b, f, a, lite, draws, x, y, a.keys[f.keyBase + static_cast<long long>(y) * f.W + x], kRcpMantissa, kSrgb8Tab);
	// the frame's counters are final (no later kernel changes them): block 0
	// publishes them to the pinned host copy
	if (blockIdx.x == 0 && a.hostCtr)
		for (int i = threadIdx.x; i < static_cast<int>(sizeof(Counters) / 4); i += 256)
			reinterpret_cast<uint32_t*>(a.hostCtr)[i] = reinterpret_cast<const uint32_t*>(a.ctr)[i]; }

// visibility work items run one per workgroup; the grid is sized from the
// previous frame's item count and strides over any remainder, so no
// work-queue counter is shared between workgroups
__global__ __launch_bounds__(kRasterThreads) void k_vis(BatchDesc b, RasterArgs a, const VisRec* __restrict__ rec,
                                                       const VisRec* __restrict__ crec, const uint32_t* __restrict__ itemTile,
                                                       const VisTile* __restrict__ tileDesc, const FrameDesc* __restrict__ frames) {
	__shared__ unsigned long long keys[kSubDim * kSubDim];
	__shared__ VisSetup ts[kRasterThreads];
	__shared__ uint32_t qoff[kRasterThreads + 1];
	__shared__ uint32_t scan_sh[kRasterThreads / 64 + 1];
	__shared__ uint8_t owner[kQuadWindow];
	uint32_t frags = 0;
	const uint32_t nItems = a.ctr->workItems;
	// Workgroups are dealt round-robin over the eight XCDs.  A tile's items
	// are consecutive and come in multiples of four (its regions), so each
	// run of four items is taken by workgroups of one XCD: they read the
	// tile's records through that XCD's L2 instead of four.
	constexpr uint32_t kGroup = 4u * kXcds;
	const uint32_t full = nItems / kGroup * kGroup;
	for (uint32_t wi = blockIdx.x; wi < nItems; wi += gridDim.x) {
		__syncthreads();
		const uint32_t item = wi < full ? wi / kGroup * kGroup + (wi % kXcds) * 4u + (wi / kXcds) % 4u : wi;
		const int tile = static_cast<int>(itemTile[item]);
		const VisTile it = tileDesc[tile];
		const uint32_t local = item - it.item0;
		const uint32_t nreg = it.nreg, seg = local / nreg, reg = local - seg * nreg;
		const bool multiSeg = it.multiSeg != 0u;
		// the frame table is read-only (restrict): its fields come by scalar loads
		const FrameDesc& f = frames[it.frame];
		if (f.inOrder) continue;   // drawn by the in-order tile path (uniform)
		const bool leq = f.depthFunc == 1;
		const Rect R = tile_rect(f, tile - f.tileBase);
		// the item's region: a quadrant, or its half or quarter
		const uint32_t per = max(1u, nreg / kQuadrants), quad = reg / per, sub = reg - quad * per;
		Rect S;
		S.x0 = R.x0 + static_cast<int>(quad & 1) * kSubDim; S.y0 = R.y0 + static_cast<int>(quad >> 1) * kSubDim;
		int sw = kSubDim, sh = kSubDim;
		if (per == 2) { S.y0 += static_cast<int>(sub) * (kSubDim / 2); sh = kSubDim / 2; }
		else if (per == 4) { S.x0 += static_cast<int>(sub & 1) * (kSubDim / 2); S.y0 += static_cast<int>(sub >> 1) * (kSubDim / 2); sw = sh = kSubDim / 2; }
		S.x1 = imin(S.x0 + sw, R.x1); S.y1 = imin(S.y0 + sh, R.y1);
		if (S.x0 >= S.x1 || S.y0 >= S.y1) continue;   // region outside the screen (uniform)
		for (int i = threadIdx.x; i < kSubDim * kSubDim; i += kRasterThreads) keys[i] = kKeySentinel;
		const uint32_t ib = it.ib, cb = it.cb, ninl = it.ninl, ntot = it.ntot;
		const uint32_t pbeg = seg * static_cast<uint32_t>(b.segPairs), pend = min(pbeg + static_cast<uint32_t>(b.segPairs), ntot);
		uint32_t itemQuads = 0;   // quads dealt out over the workgroup: the item's cost for the next frame's split
		__syncthreads();
		// each thread holds one pair's record; the next batch's record is
		// loaded while this batch rasterises
		VisRec rn;
		if (f.drawEnabled && pbeg + threadIdx.x < pend) rn = vis_record(rec, crec, ib, cb, ninl, pbeg + threadIdx.x);
		for (uint32_t b0 = pbeg; f.drawEnabled && b0 < pend; b0 += kRasterThreads) {
			const uint32_t p = b0 + threadIdx.x;
			const VisRec r = rn;
			if (p + kRasterThreads < pend) rn = vis_record(rec, crec, ib, cb, ninl, p + kRasterThreads);
			uint32_t nq = 0;
			VisSetup v;
			if (p < pend) {
				const int minx = static_cast<int>(r.minxy & 0xffffu), miny = static_cast<int>(r.minxy >> 16);
				const int nqx = static_cast<int>(r.nq & 0xffffu), nqy = static_cast<int>(r.nq >> 16);
				// the triangle's quads inside the quadrant
				const int kx0 = imax(0, (S.x0 - minx) >> 1), ky0 = imax(0, (S.y0 - miny) >> 1);
				const int kx1 = imin(nqx, (S.x1 - minx + 1) >> 1), ky1 = imin(nqy, (S.y1 - miny + 1) >> 1);
				const int nx = imax(0, kx1 - kx0), ny = imax(0, ky1 - ky0);
				// edge values at the first quad, stepped exactly as the rasteriser loops do
				v.c1 = wadd(wadd(r.c1, wmul(kx0, wmul(r.dy12, 2))), wmul(ky0, wmul(r.dx12, 2)));
				v.c2 = wadd(wadd(r.c2, wmul(kx0, wmul(r.dy23, 2))), wmul(ky0, wmul(r.dx23, 2)));
				v.c3 = wadd(wadd(r.c3, wmul(kx0, wmul(r.dy31, 2))), wmul(ky0, wmul(r.dx31, 2)));
				v.dx12 = r.dx12; v.dy12 = r.dy12; v.dx23 = r.dx23; v.dy23 = r.dy23; v.dx31 = r.dx31; v.dy31 = r.dy31;
				const int32_t lim = 1 << 21;   // 2 * |d| < 2^22
				v.small = (r.dx12 > -lim && r.dx12 < lim && r.dy12 > -lim && r.dy12 < lim &&
				           r.dx23 > -lim && r.dx23 < lim && r.dy23 > -lim && r.dy23 < lim &&
				           r.dx31 > -lim && r.dx31 < lim && r.dy31 > -lim && r.dy31 < lim) ? 1u : 0u;
				v.pad = 0u;
				v.x0 = minx + 2 * kx0 - S.x0; v.y0 = miny + 2 * ky0 - S.y0;
				v.nx = static_cast<uint32_t>(nx);
				v.invNx = nx > 0 ? 1.0F / static_cast<float>(nx) : 0.0F;
				v.scale = r.scale; v.z0 = r.z0; v.z1 = r.z1; v.z2 = r.z2; v.slot = r.slot;
				nq = static_cast<uint32_t>(nx) * static_cast<uint32_t>(ny); }
			// the triangles' quads are dealt out over the block
			const bool coop = nq != 0u;
			if (coop) ts[threadIdx.x] = v;
			// 24-bit step products are exact when every dealt-out triangle's
			// steps fit 23 bits (nearly always); decided for the whole batch
			// so that the quad loop does not diverge on it
			const bool allSmall = __syncthreads_and(!coop || v.small != 0u) != 0;
			const uint32_t cq = coop ? nq : 0u;
			uint32_t total;
			const uint32_t ex = block_exclusive_scan(cq, scan_sh, total);
			itemQuads += total;
			qoff[threadIdx.x] = ex;
			for (uint32_t win0 = 0; win0 < total; win0 += kQuadWindow) {
				// owner table: which triangle each quad item of the window belongs to
				__syncthreads();
				{
					const uint32_t s0 = max(ex, win0), s1 = min(ex + cq, win0 + kQuadWindow);
					for (uint32_t q = s0; q < s1; ++q) owner[q - win0] = static_cast<uint8_t>(threadIdx.x);
				}
				__syncthreads();
				const uint32_t wend = min(total, win0 + kQuadWindow);
				for (uint32_t it = win0 + threadIdx.x; it < wend; it += kRasterThreads) {
					const int j = owner[it - win0];
					// the owner's setup row, read whole into registers
					VisSetup w;
					const uint4* src = reinterpret_cast<const uint4*>(&ts[j]);
					uint4* dstw = reinterpret_cast<uint4*>(&w);
#pragma unroll
					for (int c = 0; c < static_cast<int>(sizeof(VisSetup) / 16); ++c) dstw[c] = src[c];
					const uint32_t k = it - qoff[j];
					const uint32_t kyr = static_cast<uint32_t>((static_cast<float>(k) + 0.5F) * w.invNx);
					const uint32_t kxr = k - __umul24(kyr, w.nx);
					const int32_t kx = static_cast<int32_t>(kxr), ky = static_cast<int32_t>(kyr);
					int32_t q1, q2, q3;
					if (allSmall) {
						q1 = wadd(wadd(w.c1, __mul24(kx, shl(w.dy12, 1))), __mul24(ky, shl(w.dx12, 1)));
						q2 = wadd(wadd(w.c2, __mul24(kx, shl(w.dy23, 1))), __mul24(ky, shl(w.dx23, 1)));
						q3 = wadd(wadd(w.c3, __mul24(kx, shl(w.dy31, 1))), __mul24(ky, shl(w.dx31, 1))); }
					else {
						q1 = wadd(wadd(w.c1, wmul(kx, shl(w.dy12, 1))), wmul(ky, shl(w.dx12, 1)));
						q2 = wadd(wadd(w.c2, wmul(kx, shl(w.dy23, 1))), wmul(ky, shl(w.dx23, 1)));
						q3 = wadd(wadd(w.c3, wmul(kx, shl(w.dy31, 1))), wmul(ky, shl(w.dx31, 1))); }
					vis_quad(keys, w, q1, q2, q3, w.x0 + 2 * kx, w.y0 + 2 * ky, leq, frags); }}
			__syncthreads(); }
		// merge the region into the frame's key buffer
		const int rw = S.x1 - S.x0, rh = S.y1 - S.y0;
		for (int i = threadIdx.x; i < rw * rh; i += kRasterThreads) {
			const int ly = i / rw, lx = i - ly * rw;
			const unsigned long long k = keys[ly * kSubDim + lx];
			unsigned long long* dst = a.keys + f.keyBase + static_cast<long long>(S.y0 + ly) * f.W + S.x0 + lx;
			if (!multiSeg) *dst = k;
			else if (k != kKeySentinel) atomicMin(dst, k); }
		if (threadIdx.x == 0 && itemQuads) atomicAdd(&a.tileCost[tile], itemQuads); }
	for (int o = 32; o > 0; o >>= 1) frags += __shfl_xor(frags, o, 64);
	__syncthreads();
	if ((threadIdx.x & 63) == 0) scan_sh[threadIdx.x >> 6] = frags;
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long sum = 0;
		for (int w = 0; w < kRasterThreads / 64; ++w) sum += scan_sh[w];
		if (sum) atomicAdd(&a.ctr->slot[blockIdx.x % kCtrSlots].fragments, sum); }}

__global__ void k_fill_u32(uint32_t* p, long long n, uint32_t v) {
	for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<long long>(gridDim.x) * blockDim.x) p[i] = v; }

__global__ void k_fill_u64(unsigned long long* p, long long n, unsigned long long v) {
	for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<long long>(gridDim.x) * blockDim.x) p[i] = v; }

__global__ void k_zero_counters(ZeroArgs z) { zero_frame(z); }

}  // namespace rgl
