// rgl_inorder.h — the in-order tile path of the gfx950 rgl pipeline.
//
// Frames whose state the order-independent visibility keys cannot express
// are drawn here as the reference draws every tile (GPU::DrawImpl,
// rglv_gpu.cxx:263-432): each tile walks its commands in submission order.
// That covers alpha blending, the scissor test, depth test or depth writes
// off, GL_EQUAL, draws with different depth functions, and clears and stores
// between draws (a depth-only clear for a shadow pass, a store followed by
// more drawing).
//
//   k_tile_order  one workgroup per tile: the tile's (triangle, tile) setup
//                 records, inline and clipped, sorted by submission slot (an
//                 LDS bitonic sort; above kSortCap records, LDS-sorted runs
//                 merged in global memory).
//   k_inorder     one workgroup per 16 x 16-pixel region of a tile, one
//                 thread per 2x2 quad: the quad's colour, depth and
//                 primitive id stay in registers while the workgroup steps
//                 through the tile's triangles (from LDS, 64 at a time) and
//                 frame commands in slot order, applying
//                 TriangleProgram::Render (rglv_gpu_impl.hxx:166-222) to the
//                 quad: coverage, scissor, early or late depth test, depth
//                 write, shading, alpha blend, masked colour store.
#pragma once

#include "rgl_kernels.h"

namespace rgl {

// A frame command that every tile executes at one point of the submission
// order: CLEAR and the STORE_* commands that GPU::BinImpl broadcasts to all
// tiles (rglv_gpu.cxx:139-191).  pos is the first visibility slot it
// precedes: the triangles drawn before it have smaller slots.
enum FrameOpKind { kOpClear = 0, kOpStoreColor = 1, kOpStoreDepth = 2, kOpStorePrim = 3,
	kOpStoreColorFP = 4,      // CMD_STORE_COLOR_FULL_LINEAR_FP: RGBA float pixels (A = 0)
	kOpStoreColorHalf = 5,    // CMD_STORE_COLOR_HALF_LINEAR_FP: 2 x 2 box average (Downsample)
	kOpStoreColorQuads = 6 }; // CMD_STORE_COLOR_FULL_QUADS_FP: the quad-swizzled qfloat4 canvas
struct alignas(16) FrameOp {
	uint32_t pos;
	int kind;
	int bits;            // clear: GL_COLOR_BUFFER_BIT (1) | GL_DEPTH_BUFFER_BIT (2)
	int gamma;           // store colour: sRGB (1) or to_tc_basic (0)
	float color[4];      // clear colour (alpha: RB_RGBAF32's fourth channel)
	float depth;         // clear depth
	int post;            // store colour: post program (1 default, 2 exposure, 3 IQ)
	float exposure;
	int stride;          // float stores: the target's row stride, in pixels (quads for kOpStoreColorQuads)
	int attach;          // float stores: the colour attachment type (the quad store's alpha source)
	int pad[3];
	void* ptr;           // store target; truecolor, depth and id stores have row stride = the frame's width
};

// the in-order frames of a batch: their first k_tile_order block (one per
// tile) and first k_inorder block (one per region of a tile)
struct alignas(16) IoFrame {
	int frame;
	int tileStart;
	int blockStart;
	int regionsX, regionsY;
	int pad[3];
};

// colour attachment types (rglv_gl.hxx RB_COLOR_DEPTH, RB_RGBF32)
constexpr int kAttachColorDepth = 0, kAttachRGBF32 = 1;

constexpr int kSortCap = 4096;       // records a tile sorts in LDS in one piece
constexpr int kRegionDim = 16;       // k_inorder region: 16 x 16 pixels, 8 x 8 quads
constexpr int kIoThreads = (kRegionDim / 2) * (kRegionDim / 2);   // one thread per quad of the region

__device__ inline int find_io(const IoFrame* io, int n, int blk, bool tiles) {
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if ((tiles ? io[mid].tileStart : io[mid].blockStart) <= blk) lo = mid; else hi = mid - 1; }
	return lo; }

// ascending bitonic sort of m (a power of two) keys in LDS by the workgroup
__device__ inline void bitonic_sort(unsigned long long* s, uint32_t m) {
	for (uint32_t k = 2; k <= m; k <<= 1)
		for (uint32_t j = k >> 1; j > 0; j >>= 1) {
			for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
				const uint32_t ixj = i ^ j;
				if (ixj > i) {
					const unsigned long long x = s[i], y = s[ixj];
					if ((x > y) == ((i & k) == 0)) { s[i] = y; s[ixj] = x; }}}
			__syncthreads(); }}

// The pixel rect of the quads of every kIoThreads consecutive records of a
// tile's sorted list (one wave each): a k_inorder region skips the batches
// whose rect misses it.  boxes[ob + k] holds the rect of records k ..
// k + kIoThreads - 1 as (x0, y0, x1, y1), x1 and y1 exclusive.
static_assert(kIoThreads == 64, "a batch's rect is one wave's reduction");
__device__ inline void batch_boxes(uint32_t n, uint32_t ninl, const VisTile& d, uint32_t ob, const VisRec* __restrict__ rec,
                                   const VisRec* __restrict__ crec, const uint32_t* __restrict__ order, int4* __restrict__ boxes) {
	__syncthreads();   // the workgroup's order[] stores
	const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
	for (uint32_t k0 = wv * 64u; k0 < n; k0 += nw * 64u) {   // uniform over the wave
		int x0 = 0x7fffffff, y0 = 0x7fffffff, x1 = -0x7fffffff, y1 = -0x7fffffff;
		if (k0 + lane < n) {
			const uint32_t p = order[ob + k0 + lane];
			const VisRec& r = p < ninl ? rec[d.ib + p] : crec[d.cb + (p - ninl)];
			const int minx = static_cast<int>(r.minxy & 0xffffu), miny = static_cast<int>(r.minxy >> 16);
			x0 = minx; y0 = miny;
			x1 = minx + 2 * static_cast<int>(r.nq & 0xffffu); y1 = miny + 2 * static_cast<int>(r.nq >> 16); }
		for (int o = 32; o > 0; o >>= 1) {
			x0 = min(x0, __shfl_xor(x0, o, 64)); y0 = min(y0, __shfl_xor(y0, o, 64));
			x1 = max(x1, __shfl_xor(x1, o, 64)); y1 = max(y1, __shfl_xor(y1, o, 64)); }
		if (lane == 0) boxes[ob + k0] = make_int4(x0, y0, x1, y1); }}

// A tile's records in submission order.  Keys are (slot << 32 | p), p the
// record's position in the tile (inline records first, then clipped); every
// (triangle, tile) pair has its own slot, so keys are distinct.  order[ob +
// k] receives the k-th record's p, ob = the tile's first inline + first
// clipped list position (the tiles' ranges are disjoint).
__global__ __launch_bounds__(256) void k_tile_order(BatchDesc b, const IoFrame* __restrict__ io, int nIo,
		const FrameDesc* __restrict__ frames, const VisTile* __restrict__ tileDesc, const VisRec* __restrict__ rec,
		const VisRec* __restrict__ crec, uint32_t* __restrict__ order, unsigned long long* __restrict__ skeys0,
		unsigned long long* __restrict__ skeys1, int4* __restrict__ boxes) {
	__shared__ unsigned long long s[kSortCap];
	const int blk = static_cast<int>(blockIdx.x);
	const IoFrame e = io[find_io(io, nIo, blk, true)];
	const FrameDesc& f = frames[e.frame];
	const int tile = f.tileBase + (blk - e.tileStart);
	const VisTile d = tileDesc[tile];
	const uint32_t n = d.ntot, ninl = d.ninl, ob = d.ib + d.cb;
	const uint32_t t = threadIdx.x;
	auto key_of = [&](uint32_t p) -> unsigned long long {
		const uint32_t slot = p < ninl ? rec[d.ib + p].slot : crec[d.cb + (p - ninl)].slot;
		return (static_cast<unsigned long long>(slot) << 32) | p; };
	if (n <= static_cast<uint32_t>(kSortCap)) {
		uint32_t m = 1;
		while (m < n) m <<= 1;
		for (uint32_t i = t; i < m; i += blockDim.x) s[i] = i < n ? key_of(i) : ~0ull;
		__syncthreads();
		bitonic_sort(s, m);
		for (uint32_t i = t; i < n; i += blockDim.x) order[ob + i] = static_cast<uint32_t>(s[i]);
		batch_boxes(n, ninl, d, ob, rec, crec, order, boxes);
		return; }
	// runs of kSortCap sorted in LDS, then merged pairwise: each key's place
	// in the merged run is its index in its own run plus its rank (binary
	// search) in the other
	unsigned long long* src = skeys0 + ob;
	unsigned long long* dst = skeys1 + ob;
	for (uint32_t c0 = 0; c0 < n; c0 += kSortCap) {
		const uint32_t cn = min(static_cast<uint32_t>(kSortCap), n - c0);
		__syncthreads();
		for (uint32_t i = t; i < static_cast<uint32_t>(kSortCap); i += blockDim.x) s[i] = i < cn ? key_of(c0 + i) : ~0ull;
		__syncthreads();
		bitonic_sort(s, kSortCap);
		for (uint32_t i = t; i < cn; i += blockDim.x) src[c0 + i] = s[i]; }
	__syncthreads();
	for (uint32_t run = kSortCap; run < n; run <<= 1) {
		for (uint32_t i = t; i < n; i += blockDim.x) {
			const uint32_t a0 = i / (2 * run) * (2 * run);
			const uint32_t a1 = min(a0 + run, n), b1 = min(a0 + 2 * run, n);
			const unsigned long long key = src[i];
			const bool inA = i < a1;
			uint32_t lo = inA ? a1 : a0, hi = inA ? b1 : a1;
			const uint32_t own = inA ? i - a0 : i - a1;
			const uint32_t base = inA ? a1 : a0;
			while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (src[mid] < key) lo = mid + 1; else hi = mid; }
			dst[a0 + own + (lo - base)] = key; }
		__syncthreads();
		unsigned long long* x = src; src = dst; dst = x; }
	for (uint32_t i = t; i < n; i += blockDim.x) order[ob + i] = static_cast<uint32_t>(src[i]);
	batch_boxes(n, ninl, d, ob, rec, crec, order, boxes); }

// ShadeFragment of program P for a quad at (qx, qy): perspective-correct BP
// from BS and the vertices' 1/w (Render, rglv_gpu_impl.hxx:187-191),
// varyings, gl_FragCoord and gl_FragDepth, the quad's texture LOD from
// lanes 0 and 1; colour with alpha per lane, and the lanes the program
// keeps (discard clears them)
template <class P>
__device__ inline uint32_t shade_quad(const DrawDesc& d, const float w[3], const float* v0, const float* v1, const float* v2,
                                      const float bx[4], const float by[4], const float bz[4], const float fd[4], int qx, int qy,
                                      int h, bool clipped, float col[4][4]) {
	float vv[4][P::kVary];
	FragCoord fc[4];
#pragma unroll
	for (int l = 0; l < 4; ++l) {
		const float fragW = oneover((bx[l] * w[0] + by[l] * w[1]) + bz[l] * w[2]);
		const float p0 = (w[0] * bx[l]) * fragW;
		const float p2 = (w[2] * bz[l]) * fragW;
		const float p1 = (1.0F - p0) - p2;
#pragma unroll
		for (int k = 0; k < P::kVary; ++k) vv[l][k] = (p0 * v0[k] + p1 * v1[k]) + p2 * v2[k];
		fc[l] = frag_coord(qx, qy, l, h, clipped, fd[l]); }
	int lod = 0;
	float u0, w0, u1, w1;
	if (lod_texcoord<P>(d, vv[0], fc[0], u0, w0) && lod_texcoord<P>(d, vv[1], fc[1], u1, w1))
		lod = texture_lod(d.tu0, u0, u1, w0, w1);
	uint32_t keep = 0;
#pragma unroll
	for (int l = 0; l < 4; ++l) keep |= (P::fragment(d, vv[l], lod, fc[l], col[l]) ? 1u : 0u) << l;
	return keep; }

struct ShadeQuad {
	const DrawDesc& d; const float* w; const float* v0; const float* v1; const float* v2;
	const float* bx; const float* by; const float* bz; const float* fd; int qx, qy, h; bool clipped;
	float (*col)[4]; uint32_t keep; bool earlyZ;
	template <class P> __device__ void run() {
		keep = shade_quad<P>(d, w, v0, v1, v2, bx, by, bz, fd, qx, qy, h, clipped, col);
		earlyZ = P::kEarlyZ; }};

// depth functions of the fragment state (DepthLT / DepthLE / DepthEQ,
// rglv_gpu_impl.hxx:64-74)
__device__ inline bool depth_pass(int func, float frag, float dest) {
	return func == 0 ? frag < dest : func == 1 ? frag <= dest : frag == dest; }

// L: the programs the draws may use (a frame of one program compiles only
// that program's shading, which keeps the registers, and so the resident
// waves that hide each triangle's texture fetches, of that program alone)
template <class L>
__global__ __launch_bounds__(kIoThreads) void k_inorder(BatchDesc b, RasterArgs a, const IoFrame* __restrict__ io, int nIo,
		const FrameDesc* __restrict__ frames, const FrameOp* __restrict__ ops, const DrawDesc* __restrict__ draws,
		const DrawLite* __restrict__ lite, const VisTile* __restrict__ tileDesc, const VisRec* __restrict__ rec,
		const VisRec* __restrict__ crec, const uint32_t* __restrict__ order, const int4* __restrict__ boxes) {
	__shared__ VisRec srec[kIoThreads];
	__shared__ int sdraw[kIoThreads];
	__shared__ uint32_t sfrag[kIoThreads / 64];
	// each batch triangle's source triangle (bit 31: a clipped fan)
	__shared__ uint32_t ssrc[kIoThreads];
	__shared__ uint32_t scan_sh[kIoThreads / 64 + 1];
	// a window of kIoThreads (triangle, quad) pairs, shaded in parallel: the
	// lanes' depths and colours, the live and kept lanes; qoff[j]: the first
	// pair of batch triangle j
	__shared__ float wfd[kIoThreads][4];
	__shared__ float4 wcol[kIoThreads][4];
	__shared__ uint32_t wmask[kIoThreads];
	__shared__ uint32_t qoff[kIoThreads + 1];
	// per batch triangle, for the ordered pass: its draw's fragment state
	// (bit 0 depth test, 1 depth write, 2-3 depth function, 4 colour write,
	// 5 blend, 6 early depth test), its primitive id, and its pair rect in
	// the region (x0 | y0 << 16, cx | cy << 16)
	__shared__ uint32_t sstate[kIoThreads], sprim[kIoThreads], sxy[kIoThreads], scxy[kIoThreads];
	const int vs = a.varyStride;
	const int blk = static_cast<int>(blockIdx.x);
	const IoFrame e = io[find_io(io, nIo, blk, false)];
	const FrameDesc& f = frames[e.frame];
	const int rpt = e.regionsX * e.regionsY;
	const int lb = blk - e.blockStart;
	const int lt = lb / rpt, reg = lb - lt * rpt;
	const int tile = f.tileBase + lt;
	const Rect R = tile_rect(f, lt);
	const int rgy = reg / e.regionsX, rgx = reg - rgy * e.regionsX;
	const int t = static_cast<int>(threadIdx.x);
	// this thread's quad: top-left pixel (qx, qy); lane l = (l & 1, l >> 1)
	const int qx = R.x0 + rgx * kRegionDim + 2 * (t % (kRegionDim / 2));
	const int qy = R.y0 + rgy * kRegionDim + 2 * (t / (kRegionDim / 2));
	bool onScreen[4];
#pragma unroll
	for (int l = 0; l < 4; ++l) onScreen[l] = qx + (l & 1) < R.x1 && qy + (l >> 1) < R.y1;
	// the tile buffers start zeroed (the oracle's convention for a tile not
	// yet cleared; the reference leaves the previous tile's contents)
	float cr[4] = { 0.0F, 0.0F, 0.0F, 0.0F }, cg[4] = { 0.0F, 0.0F, 0.0F, 0.0F }, cb[4] = { 0.0F, 0.0F, 0.0F, 0.0F };
	float dz[4] = { 0.0F, 0.0F, 0.0F, 0.0F };
	float ca[4] = { 0.0F, 0.0F, 0.0F, 0.0F };   // RB_RGBAF32's alpha: set by clears, never drawn
	uint32_t pid[4] = { 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu };
	uint32_t frags = 0;
	const VisTile vt = tileDesc[tile];
	const uint32_t n = vt.ntot, ninl = vt.ninl, ob = vt.ib + vt.cb;
	uint32_t opi = static_cast<uint32_t>(f.opBase);
	const uint32_t ope = static_cast<uint32_t>(f.opBase + f.nOps);
	uint32_t nextOp = opi < ope ? ops[opi].pos : 0xffffffffu;   // the position of the next frame command

	auto apply_op = [&](const FrameOp& op) {
		if (op.kind == kOpClear) {
			if (op.bits & 1) {
#pragma unroll
				for (int l = 0; l < 4; ++l) {
					cr[l] = op.color[0]; cg[l] = op.color[1]; cb[l] = op.color[2]; ca[l] = op.color[3]; pid[l] = 0xffffffffu; }}
			if (op.bits & 2) {
#pragma unroll
				for (int l = 0; l < 4; ++l) dz[l] = op.depth; }}
		else if (op.kind == kOpStoreColorHalf || op.kind == kOpStoreColorQuads) {
			// one output per quad whose top-left pixel lies in the tile
			// rect (the reference's loops step 2 x 2 over the rect)
			if (!onScreen[0]) return;
			const long long q = static_cast<long long>(qy >> 1) * op.stride + (qx >> 1);
			if (op.kind == kOpStoreColorHalf) {
				// Downsample (rglr_algorithm.cxx:118-165): the transposed
				// quad's pixels summed in lane order, times 1/4; alpha 0
				const float z = ((0.0F + 0.0F) + 0.0F) + 0.0F;
				static_cast<float4*>(op.ptr)[q] = make_float4((((cr[0] + cr[1]) + cr[2]) + cr[3]) * 0.25F,
				                                              (((cg[0] + cg[1]) + cg[2]) + cg[3]) * 0.25F,
				                                              (((cb[0] + cb[1]) + cb[2]) + cb[3]) * 0.25F, z * 0.25F); }
			else {
				// Copy to a QFloat4Canvas (rglr_algorithm.cxx:320-368): alpha 1
				// from RB_RGBF32, the fourth channel (RB_COLOR_DEPTH: the depth)
				// otherwise
				float4* o = static_cast<float4*>(op.ptr) + 4 * q;
				float a[4];
#pragma unroll
				for (int l = 0; l < 4; ++l) a[l] = op.attach == kAttachRGBF32 ? 1.0F : op.attach == kAttachColorDepth ? dz[l] : ca[l];
				o[0] = make_float4(cr[0], cr[1], cr[2], cr[3]);
				o[1] = make_float4(cg[0], cg[1], cg[2], cg[3]);
				o[2] = make_float4(cb[0], cb[1], cb[2], cb[3]);
				o[3] = make_float4(a[0], a[1], a[2], a[3]); }}
		else {
#pragma unroll
			for (int l = 0; l < 4; ++l) {
				if (!onScreen[l]) continue;
				const long long o = static_cast<long long>(qy + (l >> 1)) * f.W + qx + (l & 1);
				if (op.kind == kOpStoreColor) {
					float c[3] = { cr[l], cg[l], cb[l] };
					if (op.post != 1)
						post_shade(op.post, op.exposure, c, f.W, f.H, R.x0, R.y0, qx + (l & 1), qy + (l >> 1));
					const uint32_t px = op.gamma
						? ((to_srgb8(c[0]) << 16) | (to_srgb8(c[1]) << 8) | to_srgb8(c[2]))
						: ((to_linear8(c[0]) << 16) | (to_linear8(c[1]) << 8) | to_linear8(c[2]));
					static_cast<uint32_t*>(op.ptr)[o] = px; }
				else if (op.kind == kOpStoreDepth) static_cast<float*>(op.ptr)[o] = dz[l];
				else if (op.kind == kOpStoreColorFP)   // Copy to a FloatingPointCanvas (rglr_algorithm.cxx:194-280): alpha 0
					static_cast<float4*>(op.ptr)[static_cast<long long>(qy + (l >> 1)) * op.stride + qx + (l & 1)] =
						make_float4(cr[l], cg[l], cb[l], 0.0F);
				else static_cast<uint32_t*>(op.ptr)[o] = pid[l]; }}};

	// the region's pixel rect: records whose quads miss it are dropped when
	// a batch is loaded (the rest keep their order), so the workgroup steps
	// only through the triangles that can touch its pixels
	const int rx0 = R.x0 + rgx * kRegionDim, ry0 = R.y0 + rgy * kRegionDim;
	const int rx1 = min(rx0 + kRegionDim, R.x1), ry1 = min(ry0 + kRegionDim, R.y1);
	for (uint32_t base = 0; base < n; base += kIoThreads) {
		// a batch whose records' quads all miss the region (uniform)
		const int4 bb = boxes[ob + base];
		if (bb.x >= rx1 || bb.z <= rx0 || bb.y >= ry1 || bb.w <= ry0) continue;
		__syncthreads();
		VisRec rr;
		bool hit = false;
		if (base + static_cast<uint32_t>(t) < n) {
			const uint32_t p = order[ob + base + t];
			rr = p < ninl ? rec[vt.ib + p] : crec[vt.cb + (p - ninl)];
			const int minx = static_cast<int>(rr.minxy & 0xffffu), miny = static_cast<int>(rr.minxy >> 16);
			const int nqx = static_cast<int>(rr.nq & 0xffffu), nqy = static_cast<int>(rr.nq >> 16);
			hit = minx < rx1 && minx + 2 * nqx > rx0 && miny < ry1 && miny + 2 * nqy > ry0; }
		uint32_t cnt;
		const uint32_t at = block_exclusive_scan(hit ? 1u : 0u, scan_sh, cnt);
		if (hit) {
			const uint32_t t2 = at;   // the record's place among the batch's hits
			srec[t2] = rr;
			const int di = find_draw_slot(lite, b.nDraws, rr.slot);
			sdraw[t2] = di;
			const DrawLite& dl = lite[di];
			const uint32_t rel = rr.slot - dl.slotBase;
			const uint32_t nTri = static_cast<uint32_t>(dl.tris) * static_cast<uint32_t>(dl.instances);
			const uint32_t srcTri = rel < nTri ? rel : (rel - nTri) / kFanMax;
			ssrc[t2] = srcTri | (rel < nTri ? 0u : 0x80000000u);
			const DrawDesc& dd = draws[di];
			sstate[t2] = (dd.depthTest ? 1u : 0u) | (dd.depthWrite ? 2u : 0u) | (static_cast<uint32_t>(dd.depthFunc) << 2) |
			             (dd.colorWrite ? 16u : 0u) | (dd.blend ? 32u : 0u) | (program_info(dd.program).earlyZ ? 64u : 0u);
			sprim[t2] = dl.primBase + srcTri; }
		__syncthreads();
		// the batch's (triangle, quad) pairs inside the region, triangle by
		// triangle in order: triangle j's quads of its quad rect that lie in
		// the region's quads (both grids step 2 from even origins)
		auto pair_rect = [&](int j, int& x0, int& y0, int& cx, int& cy) {
			const int minx = static_cast<int>(srec[j].minxy & 0xffffu), miny = static_cast<int>(srec[j].minxy >> 16);
			const int nqx = static_cast<int>(srec[j].nq & 0xffffu), nqy = static_cast<int>(srec[j].nq >> 16);
			x0 = max(minx, rx0); y0 = max(miny, ry0);
			const int x1 = min(minx + 2 * nqx, rx0 + kRegionDim), y1 = min(miny + 2 * nqy, ry0 + kRegionDim);
			cx = x1 > x0 ? (x1 - x0 + 1) >> 1 : 0;
			cy = y1 > y0 ? (y1 - y0 + 1) >> 1 : 0; };
		uint32_t np = 0;
		if (static_cast<uint32_t>(t) < cnt) {
			int x0, y0, cx, cy;
			pair_rect(t, x0, y0, cx, cy);
			np = static_cast<uint32_t>(cx * cy);
			sxy[t] = static_cast<uint32_t>(x0) | (static_cast<uint32_t>(y0) << 16);
			scxy[t] = static_cast<uint32_t>(cx) | (static_cast<uint32_t>(cy) << 16); }
		uint32_t total;
		const uint32_t poff = block_exclusive_scan(np, scan_sh, total);
		if (static_cast<uint32_t>(t) < cnt) qoff[t] = poff;
		if (t == 0) qoff[cnt] = total;
		__syncthreads();
		for (uint32_t w0 = 0; w0 < total; w0 += kIoThreads) {
			// shade: thread t takes pair w0 + t (coverage, scissor, the
			// program), in parallel over the window
			{
				const uint32_t p = w0 + static_cast<uint32_t>(t);
				uint32_t pm = 0;   // bits 0-3: live lanes the program keeps, bits 4-7: live lanes
				if (p < total) {
					int lo = 0, hi = static_cast<int>(cnt) - 1;   // the last triangle whose pairs start at or before p
					while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (qoff[mid] <= p) lo = mid; else hi = mid - 1; }
					const int j = lo;
					int x0, y0, cx, cy;
					pair_rect(j, x0, y0, cx, cy);
					const int k = static_cast<int>(p - qoff[j]), iy = k / cx, ix = k - iy * cx;
					const int px = x0 + 2 * ix, py = y0 + 2 * iy;
					const VisRec& r = srec[j];
					const int kx = (px - static_cast<int>(r.minxy & 0xffffu)) >> 1, ky = (py - static_cast<int>(r.minxy >> 16)) >> 1;
					// the quad's lanes (c, c+dy, c+dx, c+dx+dy), stepped as the
					// rasteriser loops step (rglv_triangle.hxx:253-304)
					const int32_t q1 = wadd(wadd(r.c1, wmul(kx, wmul(r.dy12, 2))), wmul(ky, wmul(r.dx12, 2)));
					const int32_t q2 = wadd(wadd(r.c2, wmul(kx, wmul(r.dy23, 2))), wmul(ky, wmul(r.dx23, 2)));
					const int32_t q3 = wadd(wadd(r.c3, wmul(kx, wmul(r.dy31, 2))), wmul(ky, wmul(r.dx31, 2)));
					const int32_t e1[4] = { q1, wadd(q1, r.dy12), wadd(q1, r.dx12), wadd(wadd(q1, r.dx12), r.dy12) };
					const int32_t e2[4] = { q2, wadd(q2, r.dy23), wadd(q2, r.dx23), wadd(wadd(q2, r.dx23), r.dy23) };
					const int32_t e3[4] = { q3, wadd(q3, r.dy31), wadd(q3, r.dx31), wadd(wadd(q3, r.dx31), r.dy31) };
					uint32_t live = 0;
#pragma unroll
					for (int l = 0; l < 4; ++l) live |= ((e1[l] | e2[l] | e3[l]) >= 0 ? 1u : 0u) << l;
					const int di = sdraw[j];
					const DrawDesc& dd = draws[di];
					if (live && dd.scissorOn) {
						// the scissor flags of the rasteriser loops mask exactly
						// the lanes outside the clipped raster rect (tile origins
						// are even)
						const Rect S = draw_tile_rect(lite[di], (R.x0 / f.tileW), (R.y0 / f.tileH));
#pragma unroll
						for (int l = 0; l < 4; ++l) {
							const int x = px + (l & 1), y = py + (l >> 1);
							if (x < S.x0 || x >= S.x1 || y < S.y0 || y >= S.y1) live &= ~(1u << l); }}
					if (live) {
						frags += static_cast<uint32_t>(__popc(live));
						float bx[4], by[4], bz[4], fd[4];
#pragma unroll
						for (int l = 0; l < 4; ++l) {
							bx[l] = itof(e2[l]) * r.scale;
							bz[l] = itof(e1[l]) * r.scale;
							by[l] = (1.0F - bx[l]) - bz[l];
							fd[l] = (bx[l] * r.z0 + by[l] * r.z1) + bz[l] * r.z2; }
						// the program on the triangle's vertices (1/w, varyings) from
						// LDS: it has no side effects, so it runs before the quad's
						// depth test, which the ordered pass below applies
						const DrawLite& dl = lite[di];
						const uint32_t src = ssrc[j], srcTri = src & 0x7fffffffu;
						TriVerts tr;
						const float* vary;
						if (!(src >> 31)) {
							inline_verts(a.tvert[dl.triBase + srcTri], a.vpos, tr, vs);
							vary = a.vvary; }
						else {
							const uint32_t fslot = r.slot - dl.slotBase - static_cast<uint32_t>(dl.tris) * static_cast<uint32_t>(dl.instances);
							clipped_verts(a.clipqOf[dl.triBase + srcTri] * kFanMax + (fslot - srcTri * kFanMax), a.cvpos, tr, vs);
							vary = a.cvvary; }
						float col[4][4];
						ShadeQuad sq{ dd, tr.w, vary + tr.vary[0], vary + tr.vary[1], vary + tr.vary[2],
						              bx, by, bz, fd, px, py, f.H, tr.clipped, col, 0xfu, true };
						dispatch_program(dd.program, sq, L{});
						pm = (live & sq.keep) | (live << 4);
#pragma unroll
						for (int l = 0; l < 4; ++l) {
							wfd[t][l] = fd[l];
							wcol[t][l] = make_float4(col[l][0], col[l][1], col[l][2], col[l][3]); }}}
				wmask[t] = pm;
			}
			__syncthreads();
			// apply: each thread walks its quad's pairs of the window in
			// triangle order with the frame commands between them:
			// TriangleProgram::Render's early or late depth test and depth
			// write around the program, then blend and the masked colour store
			int jb = 0, je = static_cast<int>(cnt) - 1;   // the first triangle whose pairs end after w0
			while (jb < je) { const int mid = (jb + je) >> 1; if (qoff[mid + 1] > w0) je = mid; else jb = mid + 1; }
			for (int j = jb; j < static_cast<int>(cnt) && qoff[j] < w0 + kIoThreads; ++j) {
				const uint32_t slot = srec[j].slot;
				while (nextOp <= slot) { apply_op(ops[opi]); ++opi; nextOp = opi < ope ? ops[opi].pos : 0xffffffffu; }
				const uint32_t xy = sxy[j], cxy = scxy[j];
				const int x0 = static_cast<int>(xy & 0xffffu), y0 = static_cast<int>(xy >> 16);
				const int cx = static_cast<int>(cxy & 0xffffu), cy = static_cast<int>(cxy >> 16);
				if (qx < x0 || qy < y0) continue;
				const int ix = (qx - x0) >> 1, iy = (qy - y0) >> 1;
				if (ix >= cx || iy >= cy) continue;
				const uint32_t p = qoff[j] + static_cast<uint32_t>(iy * cx + ix);
				if (p < w0 || p >= w0 + kIoThreads) continue;
				const int wi = static_cast<int>(p - w0);
				const uint32_t pm = wmask[wi];
				uint32_t mask = pm >> 4;
				if (!mask) continue;
				const uint32_t st = sstate[j];
				const bool depthTest = (st & 1u) != 0u, depthWrite = (st & 2u) != 0u, earlyZ = (st & 64u) != 0u;
				const int depthFunc = static_cast<int>((st >> 2) & 3u);
				float fd[4];
#pragma unroll
				for (int l = 0; l < 4; ++l) fd[l] = wfd[wi][l];
				if (depthTest && earlyZ) {
#pragma unroll
					for (int l = 0; l < 4; ++l) if (!depth_pass(depthFunc, fd[l], dz[l])) mask &= ~(1u << l);
					if (!mask) continue; }
				if (depthWrite && earlyZ) {
#pragma unroll
					for (int l = 0; l < 4; ++l) if (mask & (1u << l)) dz[l] = fd[l]; }
				mask &= pm & 0xfu;   // gl_triMask after ShadeFragment
				if (depthTest && !earlyZ) {
#pragma unroll
					for (int l = 0; l < 4; ++l) if (!depth_pass(depthFunc, fd[l], dz[l])) mask &= ~(1u << l);
					if (!mask) continue; }
				if (depthWrite && !earlyZ) {
#pragma unroll
					for (int l = 0; l < 4; ++l) if (mask & (1u << l)) dz[l] = fd[l]; }
				if (st & 16u) {
#pragma unroll
					for (int l = 0; l < 4; ++l) {
						if (!(mask & (1u << l))) continue;
						const float4 c = wcol[wi][l];
						if (st & 32u) {   // BlendAlpha (rglv_gpu_impl.hxx:80-84)
							const float al = c.w, om = 1.0F - al;
							cr[l] = c.x * al + cr[l] * om;
							cg[l] = c.y * al + cg[l] * om;
							cb[l] = c.z * al + cb[l] * om; }
						else { cr[l] = c.x; cg[l] = c.y; cb[l] = c.z; }}}
				const uint32_t prim = sprim[j];
#pragma unroll
				for (int l = 0; l < 4; ++l) if (mask & (1u << l)) pid[l] = prim; }
			__syncthreads(); }}
	while (opi < ope) { apply_op(ops[opi]); ++opi; }
	// fragments that reached the depth test, one atomic per workgroup
	for (int o = 32; o > 0; o >>= 1) frags += __shfl_xor(frags, o, 64);
	if ((t & 63) == 0) sfrag[t >> 6] = frags;
	__syncthreads();
	if (t == 0) {
		unsigned long long sum = 0;
		for (int w = 0; w < kIoThreads / 64; ++w) sum += sfrag[w];
		if (sum) atomicAdd(&a.ctr->slot[blockIdx.x % kCtrSlots].fragments, sum); }}

}  // namespace rgl
