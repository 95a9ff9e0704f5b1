// rgl_hip.hip — host side of the gfx950 rgl pipeline: the rglv::GL command
// recorder, frame planning (validation, draw descriptors, capacities),
// kernel launches and the C ABI declared in include/rgl_hip.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rgl_hip.h"
#include "rgl_mlp.h"
#include "rgl_inorder.h"
#include "rgl_postfx.h"
#include "rgl_assets.h"

using namespace rgl;

namespace {

// GL constants (rglv_gl.hxx:20-62)
constexpr int GL_CULL_FACE = 1, GL_SCISSOR_TEST = 2, GL_BLEND = 3, GL_DEPTH_TEST = 4;
constexpr int GL_BACK = 2, GL_FRONT = 1;
constexpr int GL_LESS = 0, GL_LEQUAL = 1, GL_EQUAL = 2;
constexpr int GL_COLOR_BUFFER_BIT = 1, GL_DEPTH_BUFFER_BIT = 2, GL_STENCIL_BUFFER_BIT = 4;
constexpr int GL_DEPTH_ATTACHMENT = 0, GL_COLOR_ATTACHMENT0 = 2;
constexpr int RB_COLOR_DEPTH = 0, RB_RGBF32 = 1, RB_RGBAF32 = 2, RB_F32 = 3;
constexpr int UNIFORM_BUFFER_SIZE = 32;
constexpr int kMaxOps = 1024;     // clears and stores per batch

struct TexState { const float* ptr; int width, height, stride, filter; };

// GLState (rglv_gl.hxx:80-179), with buffer lengths for bounds checks
struct GLState {
	float clearColor[4];
	float clearDepth;
	bool cullingEnabled; int cullFace;
	bool scissorEnabled; int scissorOrigin[2], scissorSize[2];
	int viewportOrigin[2]; bool hasViewportSize; int viewportSize[2];
	bool blendingEnabled, colorWriteMask, depthWriteMask, depthTestEnabled;
	int depthFunc;
	int programId, uniformsOfs;
	int color0AttachmentType, depthAttachmentType;
	const float* buffers[16];
	int bufferCounts[16];
	TexState tus[2];
	const float* tu3ptr; int tu3dim;   // GL::BindTexture3 (rglv_gl.hxx:275-278)
	float viewMatrix[16], normalMatrix[16], projectionMatrix[16];

	void reset() {
		std::memset(this, 0, sizeof(*this));
		clearColor[3] = 1.0F; clearDepth = 1.0F;
		cullFace = GL_BACK;
		colorWriteMask = true; depthWriteMask = true; depthTestEnabled = true;
		depthFunc = GL_LESS;
		uniformsOfs = -1;
		for (int i = 0; i < 2; ++i) { tus[i].ptr = nullptr; tus[i].width = tus[i].height = tus[i].stride = 8; tus[i].filter = 0; }
		tu3ptr = nullptr; tu3dim = 256; }

	uint32_t FragmentStateKey() const {  // rglv_gl.hxx:136-146
		uint32_t key = static_cast<uint32_t>(programId) << 24;
		key |= static_cast<uint32_t>(scissorEnabled);
		key |= static_cast<uint32_t>(depthTestEnabled) << 1;
		if (depthTestEnabled) key |= static_cast<uint32_t>(depthFunc) << 2;
		key |= static_cast<uint32_t>(blendingEnabled) << 4;
		key |= static_cast<uint32_t>(depthWriteMask) << 5;
		key |= static_cast<uint32_t>(colorWriteMask) << 6;
		key |= static_cast<uint32_t>(color0AttachmentType) << 7;
		key |= static_cast<uint32_t>(depthAttachmentType) << 9;
		return key; }};

struct Cmd {
	enum Kind { STATE, CLEAR, DRAW_ELEMENTS, DRAW_ARRAYS, DRAW_ELEMENTS_INSTANCED, DRAW_ARRAYS_INSTANCED,
	            STORE_TC, STORE_DEPTH, STORE_PRIM, STORE_SAMPLES, STORE_FP, STORE_QUADS } kind;
	int a, b;
	const void* ptr; };

// mat4 product in the reference's order (rmlm_mat4.hxx:196-207)
void mat4_mul(const float* l, const float* r, float* out) {
	for (int row = 0; row < 4; ++row)
		for (int col = 0; col < 4; ++col) {
			float ax = l[0*4 + row] * r[col*4 + 0];
			ax += l[1*4 + row] * r[col*4 + 1];
			ax += l[2*4 + row] * r[col*4 + 2];
			ax += l[3*4 + row] * r[col*4 + 3];
			out[col*4 + row] = ax; }}

int ilog2(unsigned x) { int p = 0; while (x != 0U) { x >>= 1; ++p; } return p - 1; }

// rmlm::inverse (rmlm_mat4.cxx:23-167): each element a signed sum of six
// triple products of the column-major elements, summed left to right, then
// scaled by 1 / det with det = m0 inv0 + m1 inv4 + m2 inv8 + m3 inv12
struct InvTerm { signed char s; unsigned char a, b, c; };
static const InvTerm kInvTerms[16][6] = {
	{ { 1, 5, 10, 15 }, { -1, 5, 11, 14 }, { -1, 9, 6, 15 }, { 1, 9, 7, 14 }, { 1, 13, 6, 11 }, { -1, 13, 7, 10 } },   // 0
	{ { -1, 1, 10, 15 }, { 1, 1, 11, 14 }, { 1, 9, 2, 15 }, { -1, 9, 3, 14 }, { -1, 13, 2, 11 }, { 1, 13, 3, 10 } },   // 1
	{ { 1, 1, 6, 15 }, { -1, 1, 7, 14 }, { -1, 5, 2, 15 }, { 1, 5, 3, 14 }, { 1, 13, 2, 7 }, { -1, 13, 3, 6 } },        // 2
	{ { -1, 1, 6, 11 }, { 1, 1, 7, 10 }, { 1, 5, 2, 11 }, { -1, 5, 3, 10 }, { -1, 9, 2, 7 }, { 1, 9, 3, 6 } },          // 3
	{ { -1, 4, 10, 15 }, { 1, 4, 11, 14 }, { 1, 8, 6, 15 }, { -1, 8, 7, 14 }, { -1, 12, 6, 11 }, { 1, 12, 7, 10 } },   // 4
	{ { 1, 0, 10, 15 }, { -1, 0, 11, 14 }, { -1, 8, 2, 15 }, { 1, 8, 3, 14 }, { 1, 12, 2, 11 }, { -1, 12, 3, 10 } },   // 5
	{ { -1, 0, 6, 15 }, { 1, 0, 7, 14 }, { 1, 4, 2, 15 }, { -1, 4, 3, 14 }, { -1, 12, 2, 7 }, { 1, 12, 3, 6 } },        // 6
	{ { 1, 0, 6, 11 }, { -1, 0, 7, 10 }, { -1, 4, 2, 11 }, { 1, 4, 3, 10 }, { 1, 8, 2, 7 }, { -1, 8, 3, 6 } },          // 7
	{ { 1, 4, 9, 15 }, { -1, 4, 11, 13 }, { -1, 8, 5, 15 }, { 1, 8, 7, 13 }, { 1, 12, 5, 11 }, { -1, 12, 7, 9 } },      // 8
	{ { -1, 0, 9, 15 }, { 1, 0, 11, 13 }, { 1, 8, 1, 15 }, { -1, 8, 3, 13 }, { -1, 12, 1, 11 }, { 1, 12, 3, 9 } },      // 9
	{ { 1, 0, 5, 15 }, { -1, 0, 7, 13 }, { -1, 4, 1, 15 }, { 1, 4, 3, 13 }, { 1, 12, 1, 7 }, { -1, 12, 3, 5 } },        // 10
	{ { -1, 0, 5, 11 }, { 1, 0, 7, 9 }, { 1, 4, 1, 11 }, { -1, 4, 3, 9 }, { -1, 8, 1, 7 }, { 1, 8, 3, 5 } },            // 11
	{ { -1, 4, 9, 14 }, { 1, 4, 10, 13 }, { 1, 8, 5, 14 }, { -1, 8, 6, 13 }, { -1, 12, 5, 10 }, { 1, 12, 6, 9 } },      // 12
	{ { 1, 0, 9, 14 }, { -1, 0, 10, 13 }, { -1, 8, 1, 14 }, { 1, 8, 2, 13 }, { 1, 12, 1, 10 }, { -1, 12, 2, 9 } },      // 13
	{ { -1, 0, 5, 14 }, { 1, 0, 6, 13 }, { 1, 4, 1, 14 }, { -1, 4, 2, 13 }, { -1, 12, 1, 6 }, { 1, 12, 2, 5 } },        // 14
	{ { 1, 0, 5, 10 }, { -1, 0, 6, 9 }, { -1, 4, 1, 10 }, { 1, 4, 2, 9 }, { 1, 8, 1, 6 }, { -1, 8, 2, 5 } } };          // 15

void mat4_inverse(const float* m, float* inv) {
	for (int i = 0; i < 16; ++i) {
		const InvTerm* t = kInvTerms[i];
		const float a0 = t[0].s < 0 ? -m[t[0].a] : m[t[0].a];
		float acc = (a0 * m[t[0].b]) * m[t[0].c];
		for (int k = 1; k < 6; ++k) {
			const float p = (m[t[k].a] * m[t[k].b]) * m[t[k].c];
			acc = t[k].s < 0 ? acc - p : acc + p; }
		inv[i] = acc; }
	const float det = ((m[0] * inv[0] + m[1] * inv[4]) + m[2] * inv[8]) + m[3] * inv[12];
	const float invdet = 1.0F / det;
	for (int i = 0; i < 16; ++i) inv[i] = inv[i] * invdet; }

// transpose (rmlm_mat4.hxx:211-216)
void mat4_transpose(const float* m, float* out) {
	for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) out[c * 4 + r] = m[r * 4 + c]; }


template <class T>
struct DevArray {
	T* p = nullptr;
	size_t cap = 0;
	hipError_t ensure(size_t n) {
		if (n <= cap) return hipSuccess;
		if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; cap = 0; }
		size_t want = std::max<size_t>(n + n / 4, 256);
		hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T));
		if (e == hipSuccess) cap = want;
		return e; }
	void release() { if (p) (void)hipFree(p); p = nullptr; cap = 0; }};

const char* kKernelNames[] = { "k_vertex", "k_setup_count", "k_clip", "k_scan", "k_scatter", "k_vis", "k_resolve",
                               "k_mlp_shade", "k_mlp_fallback" };
constexpr int kNumTimed = 9;

// a StoreTrainingSamples request (no reference counterpart)
struct SampleReq { float* out; uint32_t cap; uint32_t* count; float keep; uint32_t seed; };

// device buffers of one training pass over a batch of B samples
struct TrainBufs { float *X, *T, *H1, *H2, *H3, *Y, *dY, *dA, *dB, *Part, *Loss; };

// rgl_mlp_train_noalloc keeps its training buffers in one file-scope device
// array, laid out for batches of at most kNoAllocMaxBatch samples; every
// sub-buffer starts on a 256-byte boundary
constexpr int kNoAllocMaxBatch = 4096;
constexpr size_t kTrainChunk = 256;   // samples per split-K partial of the weight gradients
constexpr size_t align64(size_t n) { return (n + 63) / 64 * 64; }
constexpr size_t kScrX = 0;
constexpr size_t kScrT = kScrX + align64(static_cast<size_t>(kNoAllocMaxBatch) * kMlpK1);
constexpr size_t kScrH1 = kScrT + align64(static_cast<size_t>(kNoAllocMaxBatch) * 3);
constexpr size_t kScrH2 = kScrH1 + align64(static_cast<size_t>(kNoAllocMaxBatch) * kMlpHid);
constexpr size_t kScrH3 = kScrH2 + align64(static_cast<size_t>(kNoAllocMaxBatch) * kMlpHid);
constexpr size_t kScrY = kScrH3 + align64(static_cast<size_t>(kNoAllocMaxBatch) * kMlpHid);
constexpr size_t kScrdY = kScrY + align64(static_cast<size_t>(kNoAllocMaxBatch) * 3);
constexpr size_t kScrdA = kScrdY + align64(static_cast<size_t>(kNoAllocMaxBatch) * 3);
constexpr size_t kScrdB = kScrdA + align64(static_cast<size_t>(kNoAllocMaxBatch) * kMlpHid);
constexpr size_t kScrPart = kScrdB + align64(static_cast<size_t>(kNoAllocMaxBatch) * kMlpHid);
constexpr size_t kScrLoss = kScrPart + align64((kNoAllocMaxBatch + kTrainChunk - 1) / kTrainChunk * static_cast<size_t>(kMlpTrainable));
constexpr size_t kScratchFloats = kScrLoss + 64;   // 8.6 MB

}  // namespace

// The array as the host sees it is a stand-in symbol, not a device address:
// host code uses the per-device address that hipGetSymbolAddress returns
// (rgl_gpu::scratchDev), never this name.
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {
// the device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array)
hipError_t scratch_address(int device, float** out) {
	static std::mutex m;
	static float* addr[64] = {};
	if (device < 0 || device >= 64) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lock(m);
	if (!addr[device]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess) return e;
		addr[device] = static_cast<float*>(p); }
	*out = addr[device];
	return hipSuccess; }

// dst[0] = src[0] * scale (the loss read-back of the no-allocation training)
__global__ void k_scale_copy(const float* __restrict__ src, float scale, float* __restrict__ dst) {
	if (threadIdx.x == 0) dst[0] = src[0] * scale; }
}  // namespace

struct rgl_gpu {
	int device = 0;
	int lastWorkItems = 0;          // visibility work items of the last completed frame
	std::string guid;
	std::string err;

	// recorder (rglv::GL)
	GLState cs{};
	bool dirty = true;
	std::vector<float> ubuf;
	std::deque<GLState> states;
	std::vector<Cmd> commands;
	std::vector<uint32_t> installed;
	std::vector<SampleReq> sampleReqs;
	long long frameTriStart = 0;    // batch triangle index of this recorder's first triangle (planning)
	bool fixedGuardBand = false;    // rgl_gpu_set_guard_band

	// learned shading (rgl_gpu_set_learned_shading)
	const float* mlpParams = nullptr;
	float mlpMaxRms = 2.0F / 255.0F;
	DevArray<uint32_t> mlpImg;      // the network image of learned frames
	DevArray<uint32_t> fwdImg;      // the network image of rgl_mlp_forward
	DevArray<float> mlpQueue;
	DevArray<uint32_t> probeWant;   // the probes' plain colours, kMlpProbes per tile
	DevArray<uint32_t> quadList;    // tile quadrants the fallback re-shades
	DevArray<uint32_t> qCount;      // network queue counters, kQueueSegs x kQueueStride
	int mlpShadeBlocks = 256;       // persistent network workgroups: one per CU
	bool mlpShadeBlocksSet = false;
	uint32_t mlpSegCap = 0;         // queue entries per segment
	// training scratch
	DevArray<float> tX, tT, tH1, tH2, tH3, tY, tdY, tdA, tdB, tPart, tLoss;

	// render target (GPU::Reset / Retile, rglv_gpu.cxx:50-87)
	int W = 0, H = 0, tileW = 64, tileH = 64, tilesX = 0, tilesY = 0;

	// device buffers
	DevArray<float4> vpos; DevArray<uint8_t> vflags; DevArray<float> vvary;
	DevArray<uint2> trect; DevArray<uint4> tvert;
	DevArray<uint32_t> tileCount, tileCursor, tileOff;
	DevArray<uint32_t> clipq, clipqOf; DevArray<float4> cvpos; DevArray<float> cvvary;
	DevArray<uint8_t> fcount; DevArray<uint2> frect;
	DevArray<uint32_t> clipHist, clipOff, clipCursor, itemTile, tileCost;
	DevArray<VisTile> tileDesc;
	DevArray<ScanPart> scanPart;
	size_t costReady = 0;          // tileCost entries cleared since allocation
	uint32_t itemCap = 0;
	DevArray<unsigned long long> keys;
	DevArray<VisRec> rec, crec;
	size_t keysReady = 0;           // pixels of keys holding the sentinel
	DevArray<DrawDesc> dDraws;
	DevArray<DrawLite> dLite;
	Counters* dCtr = nullptr;
	Counters* hCtr = nullptr;       // pinned
	Counters* hCtrDev = nullptr;    // hCtr as the device sees it (null: the host copies the counters)
	DrawDesc* hDraws = nullptr;     // pinned, kMaxDraws
	DrawLite* hLite = nullptr;      // pinned, kMaxDraws
	std::vector<uint8_t> lastLite;
	FrameDesc* hFrames = nullptr;   // pinned, kMaxFrames
	DevArray<FrameDesc> dFrames;
	std::vector<uint8_t> lastFrames;
	// the in-order tile path: frame commands, in-order frames, sorted tile lists
	FrameOp* hOps = nullptr;        // pinned, kMaxOps
	DevArray<FrameOp> dOps;
	std::vector<uint8_t> lastOps;
	IoFrame* hIo = nullptr;         // pinned, kMaxFrames
	DevArray<IoFrame> dIo;
	std::vector<uint8_t> lastIo;
	int nIo = 0, nIoTiles = 0, nIoBlocks = 0;
	DevArray<uint32_t> order;
	DevArray<int4> ioBoxes;         // the pixel rect of every 64 sorted records of an in-order tile
	DevArray<unsigned long long> skeys0, skeys1;
	int segPairs = kSegPairs;       // (triangle, tile) pairs per visibility work item
	int splitQuads = kSplitQuads;   // quad work per quadrant from which visibility regions are halved
	// graph replay of the launch sequence (rgl_gpu_set_graph)
	bool graphMode = false, graphValid = false;
	uint64_t graphKey = 0;
	hipGraphExec_t graphExec = nullptr;
	hipStream_t capStream = nullptr;
	int captures = 0;
	uint32_t listCap = 0, clipCap = 0, clistCap = 0;
	std::vector<uint8_t> lastTable;

	// last frame
	bool pending = false;
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	// the last network forward / gradient / training call: its own stream,
	// ordered against frames by an event (frames never run on it implicitly)
	bool auxPending = false;
	hipStream_t auxStream = nullptr;
	hipEvent_t auxDone = nullptr;
	float* scratchDev = nullptr;    // this device's address of `scratch` (set up by rgl_gpu_create)
	// double-buffered runs (rgl_gpu_set_double_buffer): further executors
	// with their own per-frame buffers; frames go round the executors, each
	// on an engine stream, so that one frame's binning overlaps the previous
	// frame's raster while the frames' stores stay in submission order
	bool dbMode = false;
	static constexpr int kMaxExec = 4;
	int nExec = 1;
	rgl_gpu* exec[kMaxExec] = {};   // exec[0] is this object
	hipStream_t dbStream[kMaxExec] = {};
	hipEvent_t dbIn = nullptr;      // recorded on the caller's stream at Run
	int dbNext = 0, dbLast = -1;    // executor of the next and of the last frame
	int dbPrev = -1;                // executor of the frame before the last
	rgl_stats dbStats{};            // double-buffered: the stats rgl_gpu_stats reports (the waited frame's)
	hipEvent_t startAfter = nullptr, rasterAfter = nullptr;   // waits of the frame being launched
	rgl_stats stats{};
	long long submitted = 0;
	int lastCommands = 0;

	// timing: a ring of event sets so timed frames stay queued back to back
	static constexpr int kRing = 64;
	bool timing = false;
	hipEvent_t ev[kRing][kNumTimed + 1] = {};
	bool evUsed[kRing] = {};
	int evNext = 0;
	hipEvent_t* cur = nullptr;
	// the visibility and resolve kernels of a timed frame carry start and
	// stop events recorded by their own dispatch, so their times are the
	// kernels' execution alone (the other stages are timed between events)
	static constexpr int kDispatchTimed = 2;   // k_vis, k_resolve
	static constexpr int kFirstDispatchTimed = 5;
	hipEvent_t kev[kRing][2 * kDispatchTimed] = {};
	bool kevUsed[kRing][kDispatchTimed] = {};
	int curSlot = -1;
	double acc[kNumTimed] = {};
	int accRuns = 0;

	void Accumulate(int slot) {
		if (!evUsed[slot]) return;
		if (hipEventSynchronize(ev[slot][kNumTimed]) == hipSuccess) {
			for (int i = 0; i < kNumTimed; ++i) {
				float ms = 0.0F;
				const int k = i - kFirstDispatchTimed;
				if (k >= 0 && k < kDispatchTimed && kevUsed[slot][k]) {
					if (hipEventElapsedTime(&ms, kev[slot][2 * k], kev[slot][2 * k + 1]) == hipSuccess) acc[i] += ms; }
				else if (hipEventElapsedTime(&ms, ev[slot][i], ev[slot][i + 1]) == hipSuccess) acc[i] += ms; }
			++accRuns; }
		evUsed[slot] = false;
		for (bool& u : kevUsed[slot]) u = false; }

	// a kernel launch; in a timed frame the dispatch records start and stop
	// events of dispatch-timed kernel k
	template <typename... KArgs, typename... Args>
	void Dispatch(int k, void (*kernel)(KArgs...), dim3 grid, dim3 block, uint32_t shm, Args... args) {
		if (cur && curSlot >= 0 && k >= 0) {
			hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, kev[curSlot][2 * k], kev[curSlot][2 * k + 1], 0, args...);
			kevUsed[curSlot][k] = true; }
		else hipLaunchKernelGGL(kernel, grid, block, shm, stream, args...); }

	~rgl_gpu() {
		for (auto* a : { &qCount, &tileCount, &tileCursor, &tileOff, &clipq, &clipqOf, &clipHist, &clipOff, &clipCursor, &itemTile, &tileCost }) a->release();
		tileDesc.release(); scanPart.release();
		keys.release(); rec.release(); crec.release();
		probeWant.release(); quadList.release();
		for (auto* a : { &mlpQueue, &tX, &tT, &tH1, &tH2, &tH3, &tY, &tdY, &tdA, &tdB, &tPart, &tLoss }) a->release();
		mlpImg.release(); fwdImg.release();
		vpos.release(); vflags.release(); vvary.release(); trect.release(); tvert.release();
		cvpos.release(); cvvary.release(); fcount.release(); frect.release(); dDraws.release(); dLite.release();
		if (dCtr) (void)hipFree(dCtr);
		if (hCtr) (void)hipHostFree(hCtr);
		if (hDraws) (void)hipHostFree(hDraws);
		if (hLite) (void)hipHostFree(hLite);
		if (hFrames) (void)hipHostFree(hFrames);
		dFrames.release();
		if (hOps) (void)hipHostFree(hOps);
		if (hIo) (void)hipHostFree(hIo);
		dOps.release(); dIo.release(); order.release(); ioBoxes.release(); skeys0.release(); skeys1.release();
		if (auxDone) (void)hipEventDestroy(auxDone);
		if (graphExec) (void)hipGraphExecDestroy(graphExec);
		if (capStream) (void)hipStreamDestroy(capStream);
		if (done) (void)hipEventDestroy(done);
		for (auto& set : ev) for (auto& e : set) if (e) (void)hipEventDestroy(e);
		for (auto& set : kev) for (auto& e : set) if (e) (void)hipEventDestroy(e); }

	int fail(int code, const std::string& msg) { err = msg; return code; }
	int hipfail(hipError_t e, const char* what) {
		err = std::string(what) + ": " + hipGetErrorString(e);
		return RGL_E_HIP; }

	void MaybeUpdateState() {  // GL::MaybeUpdateState (rglv_gl.cxx:100-106)
		if (!dirty) return;
		states.push_back(cs);
		commands.push_back({ Cmd::STATE, static_cast<int>(states.size()) - 1, 0, nullptr });
		dirty = false; }

	bool Installed(uint32_t key) const { return std::find(installed.begin(), installed.end(), key) != installed.end(); }

	void ResetGL() { dirty = true; states.clear(); cs.reset(); ubuf.clear(); commands.clear(); sampleReqs.clear(); }

	// ---- frame planning: this recorder's commands as frame fi of a batch;
	// draws are appended with batch-wide triangle and vertex numbering, the
	// frame's clears and stores to ops (its tile commands, in order) ----
	//
	// A frame takes the order-independent key path when it is one full
	// clear, then draws that all test depth LESS (or all LEQUAL) and write
	// depth and colour without blending or scissor, then at most one store
	// of each kind.  Any other command sequence is drawn in order per tile.
	int PlanFrame(FrameDesc& f, std::vector<DrawDesc>& draws, std::vector<FrameOp>& ops, long long& triBase, long long& vBase,
	              int fi, int tileBase) {
		std::memset(&f, 0, sizeof(f));
		f.W = W; f.H = H; f.tileW = tileW; f.tileH = tileH; f.tilesX = tilesX; f.tilesY = tilesY;
		f.nTiles = tilesX * tilesY;
		f.tileBase = tileBase;
		f.depthFunc = -1;
		f.drawEnabled = 1;
		f.opBase = static_cast<int>(ops.size());
		const GLState* st = nullptr;
		// the frame needs the in-order tile path (the key path's
		// visibility tiles are at most kMaxTileDim pixels square)
		bool ordered = tileW > kMaxTileDim || tileH > kMaxTileDim;
		int clears = 0, stores[3] = { 0, 0, 0 };
		bool stored = false, drawn = false;
		frameTriStart = triBase;
		submitted = 0;
		if (W <= 0 || H <= 0) return fail(RGL_E_STATE, "Reset() must set the target size before Run");
		for (const Cmd& c : commands) {
			const uint32_t pos = static_cast<uint32_t>(1 + kSlotsPerTri * triBase);   // the next triangle's first slot
			FrameOp op;
			std::memset(&op, 0, sizeof(op));
			op.pos = pos;
			switch (c.kind) {
			case Cmd::STATE: st = &states[c.a]; break;
			case Cmd::CLEAR: {   // GPU::DrawImpl CMD_CLEAR (rglv_gpu.cxx:292-324)
				if (!st) return fail(RGL_E_STATE, "clear without state");
				if (c.a & GL_STENCIL_BUFFER_BIT) return fail(RGL_E_UNSUPPORTED, "CLEAR on STENCIL not implemented");
				if (st->color0AttachmentType == RB_COLOR_DEPTH && c.a != (GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT))
					return fail(RGL_E_STATE, "must clear color and depth when using RB_COLOR_DEPTH");
				if ((c.a & GL_COLOR_BUFFER_BIT) && st->color0AttachmentType != RB_COLOR_DEPTH && st->color0AttachmentType != RB_RGBF32 &&
				    st->color0AttachmentType != RB_RGBAF32) return fail(RGL_E_UNSUPPORTED, "can't clear color buffer of this type");
				if ((c.a & GL_DEPTH_BUFFER_BIT) && st->color0AttachmentType != RB_COLOR_DEPTH && st->depthAttachmentType != RB_F32)
					return fail(RGL_E_UNSUPPORTED, "can't clear depth buffer of this type");
				op.kind = kOpClear; op.bits = c.a & 3;
				for (int k = 0; k < 4; ++k) op.color[k] = st->clearColor[k];
				op.depth = st->clearDepth;
				if (drawn || stored || clears > 0 || op.bits != 3) ordered = true;
				++clears;
				f.doClearColor = (c.a & GL_COLOR_BUFFER_BIT) != 0;
				f.doClearDepth = (c.a & GL_DEPTH_BUFFER_BIT) != 0;
				for (int k = 0; k < 3; ++k) f.clearColor[k] = st->clearColor[k];
				f.clearDepth = st->clearDepth;
				ops.push_back(op);
				break; }
			case Cmd::STORE_TC:
				if (!st) return fail(RGL_E_STATE, "store without state");
				if (!Installed((static_cast<uint32_t>(st->programId) << 24) | 0x40000000u))
					return fail(RGL_E_DISPATCH, "no dispatch entry found for CMD_STORE_COLOR_FULL_LINEAR_TC");
				if (st->programId < 1 || st->programId > 3) return fail(RGL_E_UNSUPPORTED, "post program not implemented");
				op.kind = kOpStoreColor; op.ptr = const_cast<void*>(c.ptr); op.gamma = c.a; op.post = st->programId;
				op.exposure = (st->programId == 2 && st->uniformsOfs >= 0) ? ubuf[st->uniformsOfs] : 1.0F;
				f.outColor = static_cast<uint32_t*>(const_cast<void*>(c.ptr));
				f.gamma = op.gamma; f.postProgram = op.post; f.exposure = op.exposure;
				if (stores[0]++) ordered = true;
				stored = true;
				ops.push_back(op);
				break;
			case Cmd::STORE_DEPTH:
				if (!st) return fail(RGL_E_STATE, "store without state");
				if (st->depthAttachmentType != RB_F32 && st->depthAttachmentType != RB_COLOR_DEPTH)
					return fail(RGL_E_UNSUPPORTED, "STORE_DEPTH_FULL_LINEAR_FP not implemented for attachment type");
				op.kind = kOpStoreDepth; op.ptr = const_cast<void*>(c.ptr);
				f.outDepth = static_cast<float*>(const_cast<void*>(c.ptr));
				if (stores[1]++) ordered = true;
				stored = true;
				ops.push_back(op);
				break;
			case Cmd::STORE_PRIM:
				op.kind = kOpStorePrim; op.ptr = const_cast<void*>(c.ptr);
				f.outPrim = static_cast<uint32_t*>(const_cast<void*>(c.ptr));
				if (stores[2]++) ordered = true;
				stored = true;
				ops.push_back(op);
				break;
			case Cmd::STORE_FP:      // CMD_STORE_COLOR_{FULL,HALF}_LINEAR_FP (rglv_gpu.cxx:345-370)
			case Cmd::STORE_QUADS: { // CMD_STORE_COLOR_FULL_QUADS_FP (:371-384)
				if (!st) return fail(RGL_E_STATE, "store without state");
				if (st->color0AttachmentType != RB_COLOR_DEPTH && st->color0AttachmentType != RB_RGBF32 &&
				    st->color0AttachmentType != RB_RGBAF32)
					return fail(RGL_E_UNSUPPORTED, "float colour store not implemented for attachment type");
				const bool half = c.kind == Cmd::STORE_FP && c.a != 0;
				op.kind = c.kind == Cmd::STORE_QUADS ? kOpStoreColorQuads : half ? kOpStoreColorHalf : kOpStoreColorFP;
				op.ptr = const_cast<void*>(c.ptr);
				op.stride = c.b > 0 ? c.b : (c.kind == Cmd::STORE_QUADS || half) ? (W + 1) / 2 : W;
				op.attach = st->color0AttachmentType;
				ordered = true;   // drawn by the in-order tile path
				stored = true;
				ops.push_back(op);
				break; }
			case Cmd::STORE_SAMPLES:
				stored = true; break;
			default: {
				if (!st) return fail(RGL_E_STATE, "draw without state");
				if (static_cast<int>(draws.size()) >= kMaxDraws) return fail(RGL_E_CAPACITY, "too many draws in one batch");
				const size_t nd = draws.size();
				const int r = PlanDraw(*st, c, triBase, vBase, f, draws, fi, ordered);
				if (r != 0) return r;
				if (draws.size() != nd) {
					if (stored || clears == 0) ordered = true;
					drawn = true; }
				break; }}}
		if (ops.size() > static_cast<size_t>(kMaxOps)) return fail(RGL_E_CAPACITY, "too many clears and stores in one batch");
		f.nOps = static_cast<int>(ops.size()) - f.opBase;
		f.inOrder = ordered ? 1 : 0;
		if (!ordered) {
			if (clears == 0) { f.doClearColor = 0; f.doClearDepth = 0; }
			if (f.depthFunc < 0) f.depthFunc = GL_LESS;
			if (std::isnan(f.clearDepth)) f.drawEnabled = 0; }
		if (ordered && !sampleReqs.empty()) return fail(RGL_E_UNSUPPORTED, "training samples need a frame of the key path");
		if (triBase * kSlotsPerTri + 1 >= 0xffffffffLL) return fail(RGL_E_CAPACITY, "too many triangles in one batch");
		return 0; }

	int PlanDraw(const GLState& s, const Cmd& c, long long& triBase, long long& vBase, FrameDesc& f, std::vector<DrawDesc>& draws, int fi,
	             bool& ordered) {
		const uint32_t vkey = (static_cast<uint32_t>(s.programId) << 24) | 0x80000000u;
		if (!Installed(vkey)) return fail(RGL_E_DISPATCH, "BIN: no dispatch entry for program");
		if (!Installed(s.FragmentStateKey())) return fail(RGL_E_DISPATCH, "no dispatch entry found for CMD_DRAW_INLINE");
		const ProgramInfo info = program_info(s.programId);
		const int nv = info.nvary;
		if (nv < 0) return fail(RGL_E_UNSUPPORTED, "program not implemented on the device");
		if (s.tu3ptr && (s.tu3dim <= 0 || (s.tu3dim & (s.tu3dim - 1)) != 0 || s.tu3dim > 8192))
			return fail(RGL_E_ARG, "depth texture dimension must be a power of two");
		if (s.depthFunc != GL_LESS && s.depthFunc != GL_LEQUAL && s.depthFunc != GL_EQUAL)
			return fail(RGL_E_UNSUPPORTED, "depth func not implemented");
		if (s.color0AttachmentType != RB_COLOR_DEPTH && s.color0AttachmentType != RB_RGBF32 && s.color0AttachmentType != RB_RGBAF32)
			return fail(RGL_E_UNSUPPORTED, "RDR:STATE unknown color0 Renderbuffer type");
		if (s.depthAttachmentType != RB_COLOR_DEPTH && s.depthAttachmentType != RB_F32)
			return fail(RGL_E_UNSUPPORTED, "RDR:STATE unknown depth Renderbuffer type");
		const bool elements = c.kind == Cmd::DRAW_ELEMENTS || c.kind == Cmd::DRAW_ELEMENTS_INSTANCED;
		const bool inst = c.kind == Cmd::DRAW_ELEMENTS_INSTANCED || c.kind == Cmd::DRAW_ARRAYS_INSTANCED;
		const int count = c.a;
		const int instances = inst ? c.b : 1;
		if (count < 0 || instances < 0) return fail(RGL_E_ARG, "negative count");
		if (instances > 65536) return fail(RGL_E_CAPACITY, "instance id must be < 65536");
		const int tris = count / 3;
		if (tris == 0 || instances == 0) return 0;
		// the key path takes depth-tested, depth- and colour-writing draws of
		// one depth function (LESS or LEQUAL) without blending or scissor
		if (!s.depthTestEnabled || !s.depthWriteMask || !s.colorWriteMask || s.blendingEnabled || s.scissorEnabled ||
		    s.depthFunc == GL_EQUAL || (f.depthFunc >= 0 && f.depthFunc != s.depthFunc)) ordered = true;
		// programs that test depth after shading or discard fragments
		if (!info.earlyZ || info.discard) ordered = true;
		if (f.depthFunc < 0 && (s.depthFunc == GL_LESS || s.depthFunc == GL_LEQUAL)) f.depthFunc = s.depthFunc;
		DrawDesc d;
		std::memset(&d, 0, sizeof(d));
		d.program = s.programId;
		d.instances = instances;
		d.tris = tris;
		d.elements = elements ? 1 : 0;
		d.indices = static_cast<const uint16_t*>(c.ptr);
		if (elements && d.indices == nullptr) return fail(RGL_E_ARG, "null index buffer");
		int verts;
		if (elements) {
			if (!s.buffers[0] || s.bufferCounts[0] <= 0) return fail(RGL_E_ARG, "indexed draw needs a position buffer with a count");
			verts = s.bufferCounts[0]; }
		else {
			verts = count; }
		if (verts > 32768) return fail(RGL_E_CAPACITY, "vertex index must be <= 32767");
		const bool ggx = s.programId == ProgGGXEnv::kId;
		for (int k = 0; k < (ggx ? 11 : 15); ++k)
			if (s.buffers[k] && s.bufferCounts[k] < (elements ? verts : count))
				return fail(RGL_E_ARG, "vertex buffer shorter than the position buffer");
		if (ggx) {
			// heavy-shader tables: environment cube, taps, lights, per-instance roughness
			const int need[4] = { 6 * kEnvDim * kEnvDim * 4, 4 * kEnvTaps, 8 * kGGXLights, instances };
			for (int k = 0; k < 4; ++k)
				if (s.buffers[11 + k] && s.bufferCounts[11 + k] < need[k])
					return fail(RGL_E_ARG, "heavy-shader table buffer too short"); }
		// a program that reads no instance data (its LoadInstance does
		// nothing, rglv_gpu_impl.hxx:363-366) draws every instance in place,
		// as BaseProgram does in the $layer shadow pass of a $many node
		if (info.instanced && (!s.buffers[15] || s.bufferCounts[15] < 16 * instances))
			return fail(RGL_E_ARG, "instance matrix buffer too short");
		d.verts = verts;
		d.nvary = nv;
		d.triBase = triBase;
		d.vBase = vBase;
		d.primBase = static_cast<uint32_t>(triBase - frameTriStart);   // ids count from the frame's first triangle
		d.slotBase = static_cast<uint32_t>(1 + kSlotsPerTri * triBase);
		for (int k = 0; k < 16; ++k) d.buffers[k] = s.buffers[k];
		// MakeMatrices (rglv_gpu_impl.hxx:44-51): vm, nm = transpose(inverse(vm)), vpm = pm * vm
		mat4_mul(s.projectionMatrix, s.viewMatrix, d.vpm);
		std::memcpy(d.vm, s.viewMatrix, sizeof(d.vm));
		float inv[16];
		mat4_inverse(s.viewMatrix, inv);
		mat4_transpose(inv, d.nm);
		d.tu3 = s.tu3ptr; d.tu3dim = s.tu3dim;
		// GPU::DSDO (rglv_gpu.hxx:265-270)
		const int vx = s.hasViewportSize ? s.viewportSize[0] : W;
		const int vy = s.hasViewportSize ? s.viewportSize[1] : H;
		d.DS[0] = static_cast<float>(vx / 2); d.DS[1] = static_cast<float>(-vy / 2);
		d.DO[0] = static_cast<float>(vx / 2 + s.viewportOrigin[0]);
		d.DO[1] = static_cast<float>(H - (vy / 2 + s.viewportOrigin[1]));
		// CalcGuardBandFactor (rglv_view_frustum.hxx:36-39)
		const int half = std::max(W, H) / 2;
		d.gbFactor = (2048.0F - static_cast<float>(half)) / static_cast<float>(half);
		if (fixedGuardBand && d.gbFactor < 1.0F) d.gbFactor = 1.0F;
		// GPU::ScissorRect (rglv_gpu.hxx:258-276): left, top, right, bottom
		// with bottom = H - origin.y - 1; clamped to the target, where the
		// reference would index tiles outside its grid
		d.scissor[0] = 0; d.scissor[1] = 0; d.scissor[2] = W; d.scissor[3] = H;
		d.scissorOn = s.scissorEnabled ? 1 : 0;
		if (s.scissorEnabled) {
			const int left = s.scissorOrigin[0], right = left + s.scissorSize[0];
			const int bottom = H - s.scissorOrigin[1] - 1, top = bottom - s.scissorSize[1];
			d.scissor[0] = std::min(std::max(left, 0), W); d.scissor[1] = std::min(std::max(top, 0), H);
			d.scissor[2] = std::min(std::max(right, 0), W); d.scissor[3] = std::min(std::max(bottom, 0), H); }
		d.depthTest = s.depthTestEnabled ? 1 : 0;
		d.depthFunc = s.depthFunc;
		d.depthWrite = s.depthWriteMask ? 1 : 0;
		d.colorWrite = s.colorWriteMask ? 1 : 0;
		d.blend = s.blendingEnabled ? 1 : 0;
		d.frame = fi;
		d.tileW = tileW; d.tileH = tileH; d.tilesX = tilesX; d.tilesY = tilesY; d.tileBase = f.tileBase;
		d.fW = W; d.fH = H;
		d.keepBacks = !(s.cullingEnabled && s.cullFace == GL_BACK);
		d.keepFronts = !(s.cullingEnabled && s.cullFace == GL_FRONT);
		d.cullingEnabled = s.cullingEnabled; d.cullFace = s.cullFace;
		if (s.uniformsOfs >= 0) std::memcpy(d.uniforms, ubuf.data() + s.uniformsOfs, sizeof(d.uniforms));
		for (int u = 0; u < 2; ++u) {
			SamplerDesc& sd = u == 0 ? d.tu0 : d.tu1;
			const TexState& t = s.tus[u];
			sd.buf = t.ptr; sd.kind = -1;
			if (!t.ptr) continue;
			const int power = ilog2(static_cast<unsigned>(t.width));
			const bool isPow2 = (1 << power) == t.width && t.width == t.height && t.stride == t.width;
			sd.width = t.width; sd.height = t.height; sd.stride = t.stride;
			if (!isPow2) { sd.kind = 2; sd.texels = t.stride * t.height; continue; }
			if (t.width < 4 || t.width > 1024) return fail(RGL_E_UNSUPPORTED, "can't make TextureUnit for pow2 size");
			sd.power = power; sd.kind = t.filter ? 1 : 0; sd.texels = 2 * t.width * t.width; }
		draws.push_back(d);
		triBase += static_cast<long long>(tris) * instances;
		vBase += static_cast<long long>(verts) * instances;
		submitted += static_cast<long long>(tris) * instances;
		return 0; }

	// a buffer that must grow is reallocated only once the frame in flight
	// (which may still use it) has finished; otherwise frames queue back to
	// back on the stream without a host round trip between them
	template <class T>
	hipError_t Grow(DevArray<T>& a, size_t n) {
		if (n <= a.cap) return hipSuccess;
		if (pending) { const hipError_t e = hipEventSynchronize(done); if (e != hipSuccess) return e; }
		return a.ensure(n); }

	int Ensure(const BatchDesc& b, long long vtot, long long totalPx) {
		hipError_t e;
		const long long S = std::max<long long>(b.totalTris, 1);
		if ((e = Grow(vpos, vtot + 1)) != hipSuccess) return hipfail(e, "hipMalloc vpos");
		if ((e = Grow(vflags, vtot + 1)) != hipSuccess) return hipfail(e, "hipMalloc vflags");
		if ((e = Grow(vvary, (vtot + 1) * kMaxVary)) != hipSuccess) return hipfail(e, "hipMalloc vvary");
		if ((e = Grow(trect, S)) != hipSuccess) return hipfail(e, "hipMalloc trect");
		if ((e = Grow(tvert, S)) != hipSuccess) return hipfail(e, "hipMalloc tvert");
		if ((e = Grow(tileCount, static_cast<long long>(kCountGroups) * b.nTiles + 1)) != hipSuccess) return hipfail(e, "hipMalloc tile counts");
		if ((e = Grow(tileCursor, static_cast<long long>(kCountGroups) * b.nTiles + 1)) != hipSuccess) return hipfail(e, "hipMalloc tile cursors");
		if ((e = Grow(tileOff, b.nTiles + 1)) != hipSuccess) return hipfail(e, "hipMalloc tile offsets");
		if (listCap < 2 * S + 4096) listCap = static_cast<uint32_t>(std::min<long long>(2 * S + 4096, 0x7fffffffLL));
		if ((e = Grow(rec, listCap)) != hipSuccess) return hipfail(e, "hipMalloc setup records");
		if (clipCap < std::min<long long>(S, 4096)) clipCap = static_cast<uint32_t>(std::min<long long>(S, 4096));
		if ((e = Grow(clipq, clipCap)) != hipSuccess) return hipfail(e, "hipMalloc clipq");
		if ((e = Grow(clipqOf, S)) != hipSuccess) return hipfail(e, "hipMalloc clipqOf");
		if ((e = Grow(cvpos, static_cast<size_t>(clipCap) * kClipVerts)) != hipSuccess) return hipfail(e, "hipMalloc cvpos");
		if ((e = Grow(cvvary, static_cast<size_t>(clipCap) * kClipVerts * kMaxVary)) != hipSuccess) return hipfail(e, "hipMalloc cvvary");
		if ((e = Grow(fcount, clipCap)) != hipSuccess) return hipfail(e, "hipMalloc fcount");
		if ((e = Grow(frect, static_cast<size_t>(clipCap) * kFanMax)) != hipSuccess) return hipfail(e, "hipMalloc frect");
		if ((e = Grow(clipHist, b.nTiles + 1)) != hipSuccess) return hipfail(e, "hipMalloc clipHist");
		if ((e = Grow(clipOff, b.nTiles + 1)) != hipSuccess) return hipfail(e, "hipMalloc clipOff");
		if ((e = Grow(clipCursor, b.nTiles + 1)) != hipSuccess) return hipfail(e, "hipMalloc clipCursor");
		if (clistCap < 4096) clistCap = 65536;
		if ((e = Grow(crec, clistCap)) != hipSuccess) return hipfail(e, "hipMalloc clipped setup records");
		// work items: at most kMaxRegions per segment, and a tile's segments
		// number at most its pairs / segPairs + 1
		itemCap = static_cast<uint32_t>(std::min<long long>(
			static_cast<long long>(kMaxRegions) * (b.nTiles + (static_cast<long long>(listCap) + clistCap) / b.segPairs + 1), 1LL << 30));
		if ((e = Grow(itemTile, itemCap)) != hipSuccess) return hipfail(e, "hipMalloc item tiles");
		if ((e = Grow(tileDesc, b.nTiles)) != hipSuccess) return hipfail(e, "hipMalloc tile descriptors");
		if ((e = Grow(scanPart, (b.nTiles + kScanThreads - 1) / kScanThreads + 1)) != hipSuccess) return hipfail(e, "hipMalloc scan parts");
		if ((e = Grow(tileCost, b.nTiles)) != hipSuccess) return hipfail(e, "hipMalloc tile costs");
		const size_t px = static_cast<size_t>(std::max<long long>(totalPx, 1));
		if (px > keys.cap) keysReady = 0;
		if ((e = Grow(keys, px)) != hipSuccess) return hipfail(e, "hipMalloc keys");
		if (mlpParams) {
			if ((e = Grow(mlpImg, kImgFloats)) != hipSuccess) return hipfail(e, "hipMalloc network image");
			if ((e = Grow(probeWant, static_cast<size_t>(b.nTiles) * kMlpProbes + 1)) != hipSuccess) return hipfail(e, "hipMalloc probes");
			if ((e = Grow(quadList, static_cast<size_t>(b.nTiles) * 4 + 1)) != hipSuccess) return hipfail(e, "hipMalloc fallback list");
			// each queue segment takes the resolve blocks with index % kQueueSegs == seg
			mlpSegCap = static_cast<uint32_t>((static_cast<long long>(b.nBlocks) + kQueueSegs - 1) / kQueueSegs * 256);
			if ((e = Grow(mlpQueue, static_cast<size_t>(mlpSegCap) * kQueueSegs * kQueueWords)) != hipSuccess) return hipfail(e, "hipMalloc network queue");
			if (!mlpShadeBlocksSet) {
				int cus = 0;
				if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) mlpShadeBlocks = cus;
				mlpShadeBlocksSet = true; }}
		if ((e = Grow(qCount, kQueueSegs * kQueueStride)) != hipSuccess) return hipfail(e, "hipMalloc queue counters");
		if (!dDraws.p) {
			if ((e = Grow(dDraws, kMaxDraws)) != hipSuccess) return hipfail(e, "hipMalloc draws");
			if ((e = Grow(dLite, kMaxDraws)) != hipSuccess) return hipfail(e, "hipMalloc draw table");
			lastTable.clear(); }
		if (!dFrames.p) {
			if ((e = Grow(dFrames, kMaxFrames)) != hipSuccess) return hipfail(e, "hipMalloc frames");
			lastFrames.clear(); }
		return 0; }

	void Mark(int i) { if (cur) (void)hipEventRecord(cur[i], stream); }

	// uploads a host table to its device copy through pinned staging when it
	// changed; staging is rewritten only after the last frame reading it ended
	template <class T>
	int UploadTable(const std::vector<T>& v, std::vector<uint8_t>& last, T* pinned, T* dev) {
		const size_t bytes = sizeof(T) * v.size();
		if (bytes == last.size() && (bytes == 0 || std::memcmp(last.data(), v.data(), bytes) == 0)) return 0;
		hipError_t e;
		if (pending && (e = hipEventSynchronize(done)) != hipSuccess) return hipfail(e, "hipEventSynchronize");
		std::memcpy(pinned, v.data(), bytes);
		if (bytes && (e = hipMemcpyAsync(dev, pinned, bytes, hipMemcpyHostToDevice, stream)) != hipSuccess) return hipfail(e, "hipMemcpyAsync table");
		last.assign(reinterpret_cast<const uint8_t*>(v.data()), reinterpret_cast<const uint8_t*>(v.data()) + bytes);
		graphValid = false;
		return 0; }

	// Run of a batch: recs[i]'s recorded commands draw frame i; this object's
	// buffers and stream execute all of them with one launch sequence
	int RunBatch(rgl_gpu* const* recs, int n, hipStream_t s) {
		err.clear();
		if (n < 1 || n > kMaxFrames) return fail(RGL_E_ARG, "a batch holds 1 to 64 frames");
		std::vector<FrameDesc> frames(n);
		std::vector<DrawDesc> draws;
		std::vector<FrameOp> ops;
		long long triBase = 0, vBase = 0, keyBase = 0;
		int tileBase = 0, blockBase = 0;
		submitted = 0;
		for (int i = 0; i < n; ++i) {
			rgl_gpu* r = recs[i];
			if (!r) return fail(RGL_E_ARG, "null frame in batch");
			for (int j = 0; j < i; ++j) if (recs[j] == r) return fail(RGL_E_ARG, "a recorder appears twice in one batch");
			const long long sub0 = submitted;
			const int rc = r->PlanFrame(frames[i], draws, ops, triBase, vBase, i, tileBase);
			if (rc != 0) { err = r->err; return rc; }
			submitted = sub0 + r->submitted;
			FrameDesc& f = frames[i];
			f.keyBase = keyBase; f.blockBase = blockBase;
			f.blocksX = (f.W + kResolveW - 1) / kResolveW;
			tileBase += f.nTiles;
			keyBase += static_cast<long long>(f.W) * f.H;
			blockBase += f.blocksX * ((f.H + kResolveH - 1) / kResolveH); }
		if (n > 1) for (int i = 0; i < n; ++i)
			if (!recs[i]->sampleReqs.empty()) return fail(RGL_E_UNSUPPORTED, "training samples are stored from single-frame runs");
		// the in-order frames: one k_tile_order block per tile, one k_inorder
		// block per kRegionDim x kRegionDim region of a tile
		std::vector<IoFrame> io;
		int ioTiles = 0, ioBlocks = 0;
		for (int i = 0; i < n; ++i) {
			if (!frames[i].inOrder) continue;
			IoFrame e;
			std::memset(&e, 0, sizeof(e));
			e.frame = i; e.tileStart = ioTiles; e.blockStart = ioBlocks;
			e.regionsX = (frames[i].tileW + kRegionDim - 1) / kRegionDim;
			e.regionsY = (frames[i].tileH + kRegionDim - 1) / kRegionDim;
			io.push_back(e);
			ioTiles += frames[i].nTiles;
			ioBlocks += frames[i].nTiles * e.regionsX * e.regionsY; }
		BatchDesc b;
		b.nFrames = n; b.nTiles = tileBase; b.nDraws = static_cast<int>(draws.size());
		// binning units: kUnitTris triangles, grown in steps of kUnitTris
		// when the tiles a unit's frames span (its LDS table, cleared and
		// scanned once per unit) outnumber its triangles by more than 2 to 1
		b.totalTris = triBase;
		b.unitTris = kUnitTris;
		b.unitWindow = UnitWindow(draws, triBase, kUnitTris);
		if (b.unitWindow > 2 * kUnitTris) {
			b.unitTris = kUnitTris * std::min(32, b.unitWindow / (2 * kUnitTris));
			b.unitWindow = UnitWindow(draws, triBase, b.unitTris); }
		b.units = static_cast<int>((triBase + b.unitTris - 1) / b.unitTris);
		// the setup's LDS holds the draw table and the unit's tile window
		if (sizeof(DrawLite) * draws.size() + sizeof(uint32_t) * static_cast<size_t>(b.unitWindow) > kUnitLdsBytes)
			return fail(RGL_E_UNSUPPORTED, "a binning unit's frames span too many tiles for its LDS table; use larger tiles");
		b.nBlocks = blockBase; b.segPairs = segPairs; b.splitQuads = splitQuads; b.frames = nullptr;
		long long vtot = 0;
		for (const auto& d : draws) vtot += static_cast<long long>(d.instances) * d.verts;
		hipError_t e;
		// frames of one rgl_gpu queue back to back on one stream; a frame on
		// another stream first waits for the one in flight
		if (pending && s != stream && (e = hipEventSynchronize(done)) != hipSuccess) return hipfail(e, "hipEventSynchronize");
		// a network call on another stream (which may update the parameters a
		// learned frame reads) completes before this frame's kernels start
		if (auxPending && auxStream != s && (e = hipStreamWaitEvent(s, auxDone, 0)) != hipSuccess) return hipfail(e, "hipStreamWaitEvent");
		int r = Ensure(b, vtot, keyBase);
		if (r != 0) return r;
		if (!io.empty()) {
			const size_t nl = static_cast<size_t>(listCap) + clistCap;
			if ((e = Grow(order, nl)) != hipSuccess) return hipfail(e, "hipMalloc tile order");
			if ((e = Grow(ioBoxes, nl)) != hipSuccess) return hipfail(e, "hipMalloc tile order rects");
			if ((e = Grow(skeys0, nl)) != hipSuccess) return hipfail(e, "hipMalloc sort keys");
			if ((e = Grow(skeys1, nl)) != hipSuccess) return hipfail(e, "hipMalloc sort keys"); }
		if (!dOps.p) {
			if ((e = Grow(dOps, kMaxOps)) != hipSuccess) return hipfail(e, "hipMalloc frame commands");
			if ((e = Grow(dIo, kMaxFrames)) != hipSuccess) return hipfail(e, "hipMalloc in-order frames");
			lastOps.clear(); lastIo.clear(); }
		b.frames = dFrames.p;
		stream = s;
		if (!dCtr) {
			if ((e = hipMalloc(reinterpret_cast<void**>(&dCtr), sizeof(Counters))) != hipSuccess) return hipfail(e, "hipMalloc counters");
			if ((e = hipHostMalloc(reinterpret_cast<void**>(&hCtr), sizeof(Counters), hipHostMallocMapped)) != hipSuccess) return hipfail(e, "hipHostMalloc");
			if (hipHostGetDevicePointer(reinterpret_cast<void**>(&hCtrDev), hCtr, 0) != hipSuccess) hCtrDev = nullptr;
			if ((e = hipHostMalloc(reinterpret_cast<void**>(&hDraws), sizeof(DrawDesc) * kMaxDraws, hipHostMallocDefault)) != hipSuccess) return hipfail(e, "hipHostMalloc");
			if ((e = hipHostMalloc(reinterpret_cast<void**>(&hLite), sizeof(DrawLite) * kMaxDraws, hipHostMallocDefault)) != hipSuccess) return hipfail(e, "hipHostMalloc");
			if ((e = hipHostMalloc(reinterpret_cast<void**>(&hFrames), sizeof(FrameDesc) * kMaxFrames, hipHostMallocDefault)) != hipSuccess) return hipfail(e, "hipHostMalloc");
			if ((e = hipHostMalloc(reinterpret_cast<void**>(&hOps), sizeof(FrameOp) * kMaxOps, hipHostMallocDefault)) != hipSuccess) return hipfail(e, "hipHostMalloc");
			if ((e = hipHostMalloc(reinterpret_cast<void**>(&hIo), sizeof(IoFrame) * kMaxFrames, hipHostMallocDefault)) != hipSuccess) return hipfail(e, "hipHostMalloc");
			if ((e = hipEventCreateWithFlags(&done, hipEventDisableTiming)) != hipSuccess) return hipfail(e, "hipEventCreate"); }
		cur = nullptr;
		curSlot = -1;
		if (timing) {
			if (!ev[0][0]) {
				for (auto& set : ev) for (auto& x : set) (void)hipEventCreate(&x);
				for (auto& set : kev) for (auto& x : set) (void)hipEventCreate(&x); }
			Accumulate(evNext);
			cur = ev[evNext];
			curSlot = evNext;
			evUsed[evNext] = true;
			evNext = (evNext + 1) % kRing; }
		if ((r = UploadTable(draws, lastTable, hDraws, dDraws.p)) != 0) return r;
		std::vector<DrawLite> lite(draws.size());
		for (size_t i = 0; i < draws.size(); ++i) lite[i] = make_lite(draws[i]);
		if ((r = UploadTable(lite, lastLite, hLite, dLite.p)) != 0) return r;
		if ((r = UploadTable(frames, lastFrames, hFrames, dFrames.p)) != 0) return r;
		if ((r = UploadTable(ops, lastOps, hOps, dOps.p)) != 0) return r;
		if ((r = UploadTable(io, lastIo, hIo, dIo.p)) != 0) return r;
		nIo = static_cast<int>(io.size()); nIoTiles = ioTiles; nIoBlocks = ioBlocks;
		if (costReady < tileCost.cap) {
			k_fill_u32<<<8, 256, 0, stream>>>(tileCost.p, static_cast<long long>(tileCost.cap), 0u);
			costReady = tileCost.cap; }
		if (keysReady < keys.cap) {
			// defined contents after allocation; each frame then resets the
			// keys of multi-segment tiles in k_pair_setup and overwrites the rest
			k_fill_u64<<<1024, 256, 0, stream>>>(keys.p, static_cast<long long>(keys.cap), kKeySentinel);
			keysReady = keys.cap; }
		int pid = draws.empty() ? ProgManyBlinnPhong::kId : draws[0].program;
		for (const auto& d : draws) if (d.program != pid) pid = -1;
		// the network's output is the sRGB colour of the default post program:
		// frames stored linear or through another post program shade plainly
		bool learned = mlpParams != nullptr && pid == ProgGGXEnv::kId;
		for (const FrameDesc& f : frames)
			if (f.inOrder || (f.outColor && (f.gamma != 1 || f.postProgram != 1))) learned = false;
		const std::vector<SampleReq>& reqs = recs[0]->sampleReqs;
		if (!reqs.empty() && pid != ProgGGXEnv::kId) return fail(RGL_E_UNSUPPORTED, "training samples need the heavy-shader program");
		// the launch sequence, replayed from a captured graph when enabled and
		// nothing it captured has changed
		const bool useGraph = graphMode && !timing && reqs.empty();
		const int items = std::max<int>(WorkItemsBucket(), b.nTiles * kQuadrants);
		if (useGraph) {
			const uint64_t key = GraphKey(b, learned, items);
			if (!graphValid || key != graphKey) {
				if ((r = Capture(b, draws, learned, pid, items, key)) != 0) return r; }
			if ((e = hipGraphLaunch(graphExec, stream)) != hipSuccess) return hipfail(e, "hipGraphLaunch"); }
		else {
			if ((r = Launch(b, draws, learned, pid, items, reqs)) != 0) return r; }
		if ((e = hipEventRecord(done, stream)) != hipSuccess) return hipfail(e, "hipEventRecord");
		stats = rgl_stats{};
		stats.tiles = b.nTiles;
		stats.commands = 0;
		for (int i = 0; i < n; ++i) stats.commands += static_cast<int64_t>(recs[i]->commands.size());
		stats.primitives_submitted = submitted;
		pending = true;
		return 0; }

	int Run(hipStream_t s) {
		if (dbMode) return RunDouble(s);
		rgl_gpu* self = this;
		return RunBatch(&self, 1, s); }

	// GPU::RunImpl with doubleBuffer (rglv_gpu.cxx:96-116): frame N's
	// binning runs beside frame N-1's drawing.  Here frame N goes to the
	// executor frame N-1 did not use: its kernels start after the work the
	// caller queued on s before Run, and its first kernel that writes or
	// samples caller memory (in-order tiles, resolve) waits for frame N-1's
	// end.  The caller's stream does not wait for the frame: rgl_gpu_wait does.
	int RunDouble(hipStream_t s) {
		if (mlpParams || !sampleReqs.empty()) return fail(RGL_E_UNSUPPORTED, "learned shading and training samples run single-buffered");
		if (timing || graphMode) return fail(RGL_E_UNSUPPORTED, "double-buffered frames launch eagerly and untimed");
		// a stream wait costs the host about as much as a kernel launch: none
		// for a caller stream that has finished its work, or for a previous
		// frame that has ended
		hipError_t e = hipStreamQuery(s);
		const bool callerBusy = e != hipSuccess;
		if (callerBusy && e != hipErrorNotReady) return hipfail(e, "hipStreamQuery");
		if (callerBusy && (e = hipEventRecord(dbIn, s)) != hipSuccess) return hipfail(e, "hipEventRecord");
		rgl_gpu* x = exec[dbNext];
		rgl_gpu* o = dbLast >= 0 ? exec[dbLast] : nullptr;
		x->segPairs = segPairs; x->splitQuads = splitQuads;
		x->startAfter = callerBusy ? dbIn : nullptr;
		x->rasterAfter = o && o != x && o->pending && hipEventQuery(o->done) == hipErrorNotReady ? o->done : nullptr;
		rgl_gpu* self = this;
		const int r = x->RunBatch(&self, 1, dbStream[dbNext]);
		x->startAfter = x->rasterAfter = nullptr;
		if (r != 0) { if (x != this) err = x->err; return r; }
		dbPrev = dbLast;
		dbLast = dbNext;
		dbNext = (dbNext + 1) % nExec;
		return 0; }

	// the frame before the last one queued (double-buffered), and so every
	// frame before it: each frame's stores waited for the previous frame's
	// end.  The stats are that frame's.
	int WaitPrevious() {
		if (!dbMode) return WaitAll();
		if (dbPrev < 0) return 0;
		rgl_gpu* p = exec[dbPrev];
		const int r = p->Wait();
		dbStats = p->stats;
		if (p != this && r != 0) err = p->err;
		return r; }

	// every frame in flight; the stats are the last frame's.  A capacity
	// grown by one executor is grown in the other too.
	int WaitAll() {
		if (!dbMode) return Wait();
		int rc = 0;
		std::string e0;
		for (int k = 0; k < nExec; ++k) {
			const int r = exec[k]->Wait();
			if (r != 0 && rc == 0) { rc = r; e0 = exec[k]->err; }}
		if (dbLast >= 0) dbStats = exec[dbLast]->stats;
		for (int k = 1; k < nExec; ++k) {
			listCap = std::max(listCap, exec[k]->listCap);
			clipCap = std::max(clipCap, exec[k]->clipCap);
			clistCap = std::max(clistCap, exec[k]->clistCap); }
		for (int k = 1; k < nExec; ++k) { exec[k]->listCap = listCap; exec[k]->clipCap = clipCap; exec[k]->clistCap = clistCap; }
		if (rc != 0) err = e0;
		return rc; }

	int BeginDouble(int n) {
		hipError_t e;
		nExec = n;
		exec[0] = this;
		for (int k = 1; k < n; ++k) {
			auto* t = new rgl_gpu();
			t->device = device; t->guid = guid + "/" + std::to_string(k); t->scratchDev = scratchDev;
			exec[k] = t; }
		for (int k = 0; k < n; ++k)
			if ((e = hipStreamCreateWithFlags(&dbStream[k], hipStreamNonBlocking)) != hipSuccess) { EndDouble(); return hipfail(e, "hipStreamCreate"); }
		if ((e = hipEventCreateWithFlags(&dbIn, hipEventDisableTiming)) != hipSuccess) { EndDouble(); return hipfail(e, "hipEventCreate"); }
		dbNext = 0; dbLast = -1; dbPrev = -1;
		dbStats = stats;
		dbMode = true;
		return 0; }

	// after WaitAll
	void EndDouble() {
		for (int k = 1; k < kMaxExec; ++k) { delete exec[k]; exec[k] = nullptr; }
		exec[0] = nullptr;
		nExec = 1;
		for (hipStream_t& st : dbStream) { if (st) (void)hipStreamDestroy(st); st = nullptr; }
		if (dbIn) (void)hipEventDestroy(dbIn);
		dbIn = nullptr;
		stream = nullptr;
		if (dbMode && dbLast >= 0) stats = dbStats;
		dbMode = false; }

	// The network entry points run on their caller's stream ms and share the
	// network image and training scratch with nothing a frame writes except
	// the image, so: a frame still in flight is finished first, and a call on
	// another stream than the previous network call's waits for that call
	// (its scratch may be regrown).
	// AuxEnd records the event that a later frame on another stream waits for.
	int AuxBegin(hipStream_t ms) {
		hipError_t e;
		if (pending && (e = hipEventSynchronize(done)) != hipSuccess) return hipfail(e, "hipEventSynchronize");
		if (auxPending && auxStream != ms && (e = hipEventSynchronize(auxDone)) != hipSuccess)
			return hipfail(e, "hipEventSynchronize");
		if (!auxDone) return fail(RGL_E_STATE, "network calls need the event rgl_gpu_create makes");
		return 0; }
	int AuxEnd(hipStream_t ms) {
		hipError_t e;
		if ((e = hipEventRecord(auxDone, ms)) != hipSuccess) return hipfail(e, "hipEventRecord");
		auxPending = true;
		auxStream = ms;
		return 0; }

	// the most tiles any binning unit of ut triangles can touch: from the
	// first tile of its first triangle's frame to the last tile of its last
	// triangle's frame, with the draw of a triangle chosen as the kernels'
	// find_draw_ll does (the last draw whose first triangle is at or before it)
	static int UnitWindow(const std::vector<DrawDesc>& draws, long long total, long long ut) {
		if (draws.empty() || total <= 0) return 1;
		const int nd = static_cast<int>(draws.size());
		auto drawOf = [&](long long s, int from) {
			int d = from;
			while (d + 1 < nd && draws[d + 1].triBase <= s) ++d;
			return d; };
		int best = 1, d0 = 0, d1 = 0;
		for (long long s0 = 0; s0 < total; s0 += ut) {
			d0 = drawOf(s0, d0);
			d1 = drawOf(std::min(s0 + ut, total) - 1, std::max(d1, d0));
			const DrawDesc& a = draws[d0];
			const DrawDesc& z = draws[d1];
			best = std::max(best, z.tileBase + z.tilesX * z.tilesY - a.tileBase); }
		return best; }

	// visibility grid: the last completed frame's item count, rounded up to a
	// power of two so that a captured graph survives small changes
	int WorkItemsBucket() const {
		int v = 1;
		while (v < lastWorkItems && v < (1 << 20)) v <<= 1;
		return lastWorkItems > 0 ? v : 0; }

	uint64_t GraphKey(const BatchDesc& b, bool learned, int items) const {
		uint64_t h = 1469598103934665603ull;
		auto mix = [&h](const void* p, size_t n) {
			const uint8_t* c = static_cast<const uint8_t*>(p);
			for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }};
		mix(&b, sizeof(b));
		mix(lastTable.data(), lastTable.size());
		mix(lastFrames.data(), lastFrames.size());
		mix(lastOps.data(), lastOps.size());
		mix(lastIo.data(), lastIo.size());
		const void* ptrs[] = { vpos.p, vflags.p, vvary.p, trect.p, tvert.p, tileCount.p, tileCursor.p, tileOff.p, clipq.p, clipqOf.p,
		                       cvpos.p, cvvary.p, fcount.p, frect.p, clipHist.p, clipOff.p, clipCursor.p, itemTile.p, tileDesc.p, scanPart.p, tileCost.p,
		                       keys.p, rec.p, crec.p, mlpImg.p, probeWant.p, quadList.p, mlpQueue.p, qCount.p, mlpParams,
		                       dOps.p, dIo.p, order.p, ioBoxes.p, skeys0.p, skeys1.p };
		mix(ptrs, sizeof(ptrs));
		const uint32_t caps[] = { listCap, clipCap, clistCap, mlpSegCap, static_cast<uint32_t>(items), learned ? 1u : 0u,
		                          static_cast<uint32_t>(mlpShadeBlocks) };
		mix(caps, sizeof(caps));
		const float lim = mlpMaxRms;
		mix(&lim, sizeof(lim));
		return h; }

	int Capture(const BatchDesc& b, const std::vector<DrawDesc>& draws, bool learned, int pid, int items, uint64_t key) {
		hipError_t e;
		if (pending && (e = hipEventSynchronize(done)) != hipSuccess) return hipfail(e, "hipEventSynchronize");
		if (graphExec) { (void)hipGraphExecDestroy(graphExec); graphExec = nullptr; }
		if (!capStream && (e = hipStreamCreateWithFlags(&capStream, hipStreamNonBlocking)) != hipSuccess) return hipfail(e, "hipStreamCreate");
		hipStream_t user = stream;
		stream = capStream;
		if ((e = hipStreamBeginCapture(capStream, hipStreamCaptureModeThreadLocal)) != hipSuccess) { stream = user; return hipfail(e, "hipStreamBeginCapture"); }
		static const std::vector<SampleReq> none;
		const int r = Launch(b, draws, learned, pid, items, none);
		hipGraph_t graph = nullptr;
		e = hipStreamEndCapture(capStream, &graph);
		stream = user;
		if (r != 0) { if (graph) (void)hipGraphDestroy(graph); return r; }
		if (e != hipSuccess) return hipfail(e, "hipStreamEndCapture");
		e = hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0);
		(void)hipGraphDestroy(graph);
		if (e != hipSuccess) { graphExec = nullptr; return hipfail(e, "hipGraphInstantiate"); }
		graphKey = key;
		graphValid = true;
		++captures;
		return 0; }

	// per-program launches, generated from the program list
	struct VertexLaunch {
		rgl_gpu& g; const std::vector<DrawDesc>& draws; const BatchDesc& b; const ZeroArgs& za; unsigned vblocks; long long vtot;
		int zeroHere; int varyStride;
		template <class P> void run() {
			bool present = false;
			for (const auto& d : draws) present |= d.program == P::kId;
			if (!present || vtot <= 0) return;
			k_vertex<P><<<vblocks, 256, 0, g.stream>>>(g.dDraws.p, b.nDraws, vtot, g.vpos.p, g.vflags.p, g.vvary.p, za, zeroHere,
			                                           varyStride);
			zeroHere = 0; }};
	struct ClipLaunch {
		rgl_gpu& g; const std::vector<DrawDesc>& draws; const BatchDesc& b; int varyStride;
		template <class P> void run() {
			bool present = false;
			for (const auto& d : draws) present |= d.program == P::kId;
			if (!present) return;
			const unsigned waves = std::max(1u, std::min<unsigned>(kClipWaves, g.clipCap));
			k_clip<P><<<waves, 64, 0, g.stream>>>(g.dDraws.p, b.nDraws, g.clipq.p, g.clipCap, g.cvpos.p, g.cvvary.p, g.fcount.p,
				g.frect.p, g.clipHist.p, g.dCtr, varyStride); }};
	// k_inorder compiled for the batch's one program (or for all of them)
	struct InorderLaunch {
		rgl_gpu& g; const BatchDesc& b; const RasterArgs& ra; size_t lds;
		template <class P> void run() { launch<ProgramList<P>>(); }
		template <class L> void launch() {
			k_inorder<L><<<g.nIoBlocks, kIoThreads, lds, g.stream>>>(b, ra, g.dIo.p, g.nIo, b.frames, g.dOps.p, g.dDraws.p, g.dLite.p, g.tileDesc.p,
				g.rec.p, g.crec.p, g.order.p, g.ioBoxes.p); }};
	struct ResolveLaunch {
		rgl_gpu& g; const BatchDesc& b; const RasterArgs& ra; unsigned rblocks; const DrawLite* lite; const DrawDesc* dd;
		template <class P> void run() {
			g.Dispatch(1, k_resolve<ShadeOne<P>>, dim3(rblocks), dim3(256), 0, b, ra, b.frames, lite, dd); }};

	// every kernel of one batch, on `stream`; ends with the counters' copy
	// to pinned host memory
	int Launch(const BatchDesc& b, const std::vector<DrawDesc>& draws, bool learned, int pid, int items,
	           const std::vector<SampleReq>& reqs) {
		hipError_t e;
		const size_t dlBytes = sizeof(DrawLite) * static_cast<size_t>(std::max(b.nDraws, 1));
		// small batches scan inside k_scatter (kFusedScanTiles)
		const bool fusedScan = b.nTiles <= kFusedScanTiles && b.units > 0;
		const ZeroArgs za{ dCtr, clipHist.p, b.nTiles, tileCount.p, qCount.p, fusedScan ? tileCursor.p : nullptr,
		                   fusedScan ? clipCursor.p : nullptr };
		// a double-buffered frame reads the caller's buffers only after the
		// work the caller queued before Run
		if (startAfter && (e = hipStreamWaitEvent(stream, startAfter, 0)) != hipSuccess) return hipfail(e, "hipStreamWaitEvent");
		Mark(0);
		// vertex and clip stages: one launch per program present in the batch
		long long vtot = 0;
		for (const auto& d : draws) vtot += static_cast<long long>(d.instances) * d.verts;
		const unsigned vblocks = static_cast<unsigned>((vtot + 255) / 256);
		// varyings are stored at the batch's largest varying count per vertex
		int varyStride = 1;
		for (const auto& d : draws) varyStride = std::max(varyStride, d.nvary);
		// the first vertex launch also clears the frame's counters
		VertexLaunch vl{ *this, draws, b, za, vblocks, vtot, 1, varyStride };
		for_each_program(vl, AllPrograms{});
		int zeroHere = vl.zeroHere;
		if (zeroHere) k_zero_counters<<<std::max(8, std::min((kCountGroups * b.nTiles + 255) / 256, 256)), 256, 0, stream>>>(za);
		Mark(1);
		if (b.units > 0) {
			k_setup_count<<<b.units, kUnitThreads, dlBytes + b.unitWindow * sizeof(uint32_t), stream>>>(dDraws.p, b, vpos.p, vflags.p, trect.p, tvert.p,
				tileCount.p, clipq.p, clipqOf.p, clipCap, dCtr); }
		Mark(2);
		// clipping: the clip queue is complete after setup; the clipped
		// per-tile counts feed the scan below
		ClipLaunch cl{ *this, draws, b, varyStride };
		for_each_program(cl, AllPrograms{});
		Mark(3);
		const ScanOut so{ tileCursor.p, tileOff.p, clipOff.p, clipCursor.p, itemTile.p, itemCap, tileDesc.p, listCap, clistCap,
		                  tileCost.p, dCtr };
		if (fusedScan) {}
		else if (b.nTiles > kScanSplitTiles) {
			const int nb = (b.nTiles + kScanThreads - 1) / kScanThreads;
			k_scan_local<<<nb, kScanThreads, 0, stream>>>(b, tileCount.p, clipHist.p, tileCost.p, tileDesc.p, scanPart.p);
			k_scan_parts<<<1, kScanThreads, 0, stream>>>(scanPart.p, nb, b.nTiles, tileOff.p, clipOff.p, itemCap, dCtr);
			k_scan_final<<<nb, kScanThreads, 0, stream>>>(b, tileCount.p, tileCursor.p, tileOff.p, clipOff.p, clipCursor.p,
				itemTile.p, itemCap, tileDesc.p, listCap, clistCap, scanPart.p, nb); }
		else k_col_scan<<<1, kScanThreads, 0, stream>>>(b, tileCount.p, clipHist.p, so);
		Mark(4);
		RasterArgs ra;
		ra.draws = dDraws.p; ra.lite = dLite.p; ra.tileOff = tileOff.p; ra.listCap = listCap;
		ra.tvert = tvert.p; ra.vpos = vpos.p; ra.vvary = vvary.p;
		ra.clipq = clipq.p; ra.clipqOf = clipqOf.p; ra.cvpos = cvpos.p; ra.cvvary = cvvary.p; ra.varyStride = varyStride;
		ra.clipOff = clipOff.p; ra.clistCap = clistCap; ra.clipCap = clipCap; ra.ctr = dCtr;
		ra.keys = keys.p; ra.tileCost = tileCost.p;
		// plain frames end with k_resolve, which publishes the counters itself
		ra.hostCtr = learned ? nullptr : hCtrDev;
		if (b.units > 0) {
			ScatterArgs sa;
			sa.trect = trect.p; sa.tvert = tvert.p; sa.vpos = vpos.p; sa.cursor = tileCursor.p; sa.rec = rec.p; sa.listCap = listCap;
			sa.clipq = clipq.p; sa.fcount = fcount.p; sa.frect = frect.p; sa.cvpos = cvpos.p; sa.clipCap = clipCap;
			sa.clipCursor = clipCursor.p; sa.crec = crec.p; sa.clistCap = clistCap; sa.clipBlocks = kClipScatterBlocks;
			sa.tileOff = tileOff.p; sa.clipOff = clipOff.p; sa.keys = keys.p; sa.ctr = dCtr;
			sa.fused = fusedScan ? 1 : 0; sa.tileCount = tileCount.p; sa.clipHist = clipHist.p;
			sa.scan = so;
			sa.scan.tileCursor = nullptr; sa.scan.clipCursor = nullptr;
			k_scatter<<<(fusedScan ? 1 : 0) + b.units + kClipScatterBlocks + kResetBlocks, kUnitThreads, b.unitWindow * sizeof(uint32_t),
				stream>>>(b, dLite.p, sa); }
		Mark(5);
		if (b.nTiles > 0) {
			// one workgroup per work item, sized from the last completed frame;
			// the kernel strides over any remainder
			Dispatch(0, k_vis, dim3(std::min(items, 1 << 20)), dim3(kRasterThreads), 0, b, ra,
			         static_cast<const VisRec*>(rec.p), static_cast<const VisRec*>(crec.p), static_cast<const uint32_t*>(itemTile.p),
			         static_cast<const VisTile*>(tileDesc.p), b.frames); }
		// frames drawn in order: the tiles' records sorted by slot, then
		// every tile region walks them (before k_resolve, whose first block
		// publishes the frame's counters)
		if (nIo > 0)
			k_tile_order<<<nIoTiles, 256, 0, stream>>>(b, dIo.p, nIo, b.frames, tileDesc.p, rec.p, crec.p, order.p, skeys0.p, skeys1.p,
				ioBoxes.p);
		// everything above writes this executor's buffers alone; what
		// follows writes or samples the caller's, after the previous
		// double-buffered frame has finished
		if (rasterAfter && (e = hipStreamWaitEvent(stream, rasterAfter, 0)) != hipSuccess) return hipfail(e, "hipStreamWaitEvent");
		if (nIo > 0) {
			InorderLaunch il{ *this, b, ra, 0 };
			if (pid < 0 || !dispatch_program(pid, il, AllPrograms{})) il.launch<AllPrograms>(); }
		Mark(6);
		const unsigned rblocks = static_cast<unsigned>(std::max(b.nBlocks, 1));
		const DrawLite* cLite = dLite.p;
		const DrawDesc* cDraws = dDraws.p;
		// a limit of 1 or more (255 per channel) re-shades no tile: no probes
		const bool fallback = learned && b.nTiles > 0 && mlpMaxRms < 1.0F;
		if (learned) {
			// the gather's leading workgroups compute the tiles' probes (their
			// plain colours need the keys, not the network)
			const int probeBlocks = fallback ? (b.nTiles + kProbeTilesPerBlock - 1) / kProbeTilesPerBlock : 0;
			k_mlp_pack<<<16, 256, 0, stream>>>(mlpParams, mlpImg.p);
			k_mlp_gather<ProgGGXEnv><<<rblocks + probeBlocks, 256, dlBytes, stream>>>(b, ra, mlpQueue.p, qCount.p, mlpSegCap,
				probeWant.p, probeBlocks, mlpImg.p); }
		else {
			// a batch of one program resolves with it compiled in, others
			// dispatch over the program list per pixel
			ResolveLaunch rl{ *this, b, ra, rblocks, cLite, cDraws };
			if (pid < 0 || !dispatch_program(pid, rl, AllPrograms{}))
				Dispatch(1, k_resolve<ShadeAny>, dim3(rblocks), dim3(256), 0, b, ra, b.frames, cLite, cDraws); }
		Mark(7);
		if (learned) k_mlp_shade<<<mlpShadeBlocks, 64 * kMlpWaves, 0, stream>>>(b, mlpQueue.p, qCount.p, mlpSegCap, mlpImg.p);
		Mark(8);
		if (fallback) {
			// plain re-shading of the tiles whose probes show an error over
			// the limit
			const float lim = mlpMaxRms * 255.0F;
			k_mlp_decide<<<(b.nTiles + 3) / 4, 256, 0, stream>>>(b, probeWant.p, lim * lim, quadList.p, &dCtr->fallbackQuads);
			// 16 x 16 pixel workgroups per listed quadrant: a quadrant is at
			// most 4 probe cells of ceil(tile / 8) pixels across
			int tw = 8, th = 8;
			for (int k = 0; k < b.nFrames; ++k) { tw = std::max(tw, hFrames[k].tileW); th = std::max(th, hFrames[k].tileH); }
			const int subX = (4 * ((tw + 7) / 8) + 15) / 16, subY = (4 * ((th + 7) / 8) + 15) / 16;
			k_mlp_fallback<ProgGGXEnv><<<4 * b.nTiles * subX * subY, 256, dlBytes, stream>>>(b, ra, quadList.p, &dCtr->fallbackQuads,
				subX, subY); }
		for (const SampleReq& q : reqs)
			k_mlp_samples<ProgGGXEnv><<<rblocks, 256, dlBytes, stream>>>(b, ra, q.out, q.cap, q.count, q.keep, q.seed);
		Mark(9);
		if ((e = hipGetLastError()) != hipSuccess) return hipfail(e, "kernel launch");
		if (!ra.hostCtr && (e = hipMemcpyAsync(hCtr, dCtr, sizeof(Counters), hipMemcpyDeviceToHost, stream)) != hipSuccess)
			return hipfail(e, "hipMemcpyAsync counters");
		return 0; }

	static constexpr int kClipScatterBlocks = 32;   // k_scatter blocks that place the clipped fans
	static constexpr size_t kUnitLdsBytes = 160 * 1024;   // LDS per workgroup on gfx950
	static constexpr int kResetBlocks = 128;        // k_scatter blocks that reset the keys of multi-segment tiles

	// ---- learned shading: one forward + backward pass over a batch ----
	static void Gemm(hipStream_t ts, const float* A, const float* B, float* C, int M, int N, int K, int lda, int ldb, int ldc,
	                 int tA, int tB, const float* bias, int relu, const float* mask, int ldm, int kChunk, int nz) {
		GemmArgs g{ A, B, C, M, N, K, lda, ldb, ldc, tA, tB, bias, relu, mask, ldm, kChunk, kMlpTrainable };
		k_gemm_bf16<<<dim3((N + 63) / 64, (M + 63) / 64, nz), 256, 0, ts>>>(g); }

	// seed 0xffffffff: the first B samples in order.  lr > 0: Adam step on
	// prm/opt; lr == 0: the summed gradient goes to gradOut.
	// the training buffers of a batch of B samples, allocated on demand
	int TrainBuffers(int B, TrainBufs& tb) {
		hipError_t e;
		const int nz = (B + static_cast<int>(kTrainChunk) - 1) / static_cast<int>(kTrainChunk);
		for (auto* a : { &tX }) if ((e = a->ensure(static_cast<size_t>(B) * kMlpK1)) != hipSuccess) return hipfail(e, "hipMalloc training");
		for (auto* a : { &tH1, &tH2, &tH3, &tdA, &tdB }) if ((e = a->ensure(static_cast<size_t>(B) * kMlpHid)) != hipSuccess) return hipfail(e, "hipMalloc training");
		for (auto* a : { &tT, &tY, &tdY }) if ((e = a->ensure(static_cast<size_t>(B) * 3)) != hipSuccess) return hipfail(e, "hipMalloc training");
		if ((e = tPart.ensure(static_cast<size_t>(nz) * kMlpTrainable)) != hipSuccess) return hipfail(e, "hipMalloc training");
		if ((e = tLoss.ensure(1)) != hipSuccess) return hipfail(e, "hipMalloc training");
		tb = TrainBufs{ tX.p, tT.p, tH1.p, tH2.p, tH3.p, tY.p, tdY.p, tdA.p, tdB.p, tPart.p, tLoss.p };
		return 0; }

	// the same buffers inside this device's `scratch` array (no allocation)
	TrainBufs ScratchBuffers() const {
		float* b = scratchDev;
		return TrainBufs{ b + kScrX, b + kScrT, b + kScrH1, b + kScrH2, b + kScrH3, b + kScrY, b + kScrdY, b + kScrdA, b + kScrdB,
		                  b + kScrPart, b + kScrLoss }; }

	// seed 0xffffffff: the first B samples in order.  lr > 0: Adam step on
	// prm/opt; lr == 0: the summed gradient goes to gradOut.
	int TrainPass(hipStream_t ts, const TrainBufs& tb, const float* samples, long long n, int B, float* prm, float* opt, float lr,
	              float b1, float b2, int t, uint32_t seed, float* gradOut) {
		hipError_t e;
		const int kChunk = static_cast<int>(kTrainChunk);
		const int nz = (B + kChunk - 1) / kChunk;
		k_mlp_batch<<<(B + 255) / 256, 256, 0, ts>>>(samples, n, B, seed, static_cast<uint32_t>(t), prm, tb.X, tb.T);
		// forward: H = relu(X W^T + b), Y = H3 W4^T + b4
		Gemm(ts, tb.X, prm + kW1, tb.H1, B, kMlpHid, kMlpK1, kMlpK1, kMlpK1, kMlpHid, 0, 1, prm + kB1, 1, nullptr, 0, kMlpK1, 1);
		Gemm(ts, tb.H1, prm + kW2, tb.H2, B, kMlpHid, kMlpHid, kMlpHid, kMlpHid, kMlpHid, 0, 1, prm + kB2, 1, nullptr, 0, kMlpHid, 1);
		Gemm(ts, tb.H2, prm + kW3, tb.H3, B, kMlpHid, kMlpHid, kMlpHid, kMlpHid, kMlpHid, 0, 1, prm + kB3, 1, nullptr, 0, kMlpHid, 1);
		Gemm(ts, tb.H3, prm + kW4, tb.Y, B, kMlpOut, kMlpHid, kMlpHid, kMlpHid, kMlpOut, 0, 1, prm + kB4, 0, nullptr, 0, kMlpHid, 1);
		k_fill_u32<<<1, 64, 0, ts>>>(reinterpret_cast<uint32_t*>(tb.Loss), 1, 0u);
		k_mlp_loss<<<(3 * B + 255) / 256, 256, 0, ts>>>(tb.Y, tb.T, B, tb.dY, tb.Loss);
		// backward: weight gradients split over K = samples into nz partials
		float* P = tb.Part;
		Gemm(ts, tb.dY, tb.H3, P + kW4, kMlpOut, kMlpHid, B, kMlpOut, kMlpHid, kMlpHid, 1, 0, nullptr, 0, nullptr, 0, kChunk, nz);
		k_mlp_colsum<<<nz, 64, 0, ts>>>(tb.dY, B, kMlpOut, kChunk, P + kB4, kMlpTrainable);
		Gemm(ts, tb.dY, prm + kW4, tb.dA, B, kMlpHid, kMlpOut, kMlpOut, kMlpHid, kMlpHid, 0, 0, nullptr, 0, tb.H3, kMlpHid, kMlpOut, 1);
		Gemm(ts, tb.dA, tb.H2, P + kW3, kMlpHid, kMlpHid, B, kMlpHid, kMlpHid, kMlpHid, 1, 0, nullptr, 0, nullptr, 0, kChunk, nz);
		k_mlp_colsum<<<nz, 64, 0, ts>>>(tb.dA, B, kMlpHid, kChunk, P + kB3, kMlpTrainable);
		Gemm(ts, tb.dA, prm + kW3, tb.dB, B, kMlpHid, kMlpHid, kMlpHid, kMlpHid, kMlpHid, 0, 0, nullptr, 0, tb.H2, kMlpHid, kMlpHid, 1);
		Gemm(ts, tb.dB, tb.H1, P + kW2, kMlpHid, kMlpHid, B, kMlpHid, kMlpHid, kMlpHid, 1, 0, nullptr, 0, nullptr, 0, kChunk, nz);
		k_mlp_colsum<<<nz, 64, 0, ts>>>(tb.dB, B, kMlpHid, kChunk, P + kB2, kMlpTrainable);
		Gemm(ts, tb.dB, prm + kW2, tb.dA, B, kMlpHid, kMlpHid, kMlpHid, kMlpHid, kMlpHid, 0, 0, nullptr, 0, tb.H1, kMlpHid, kMlpHid, 1);
		Gemm(ts, tb.dA, tb.X, P + kW1, kMlpHid, kMlpK1, B, kMlpHid, kMlpK1, kMlpK1, 1, 0, nullptr, 0, nullptr, 0, kChunk, nz);
		k_mlp_colsum<<<nz, 64, 0, ts>>>(tb.dA, B, kMlpHid, kChunk, P + kB1, kMlpTrainable);
		const float bc1 = 1.0F - std::pow(b1, static_cast<float>(t)), bc2 = 1.0F - std::pow(b2, static_cast<float>(t));
		k_mlp_adam<<<(kMlpTrainable + 255) / 256, 256, 0, ts>>>(P, nz, prm, opt, opt ? opt + kMlpTrainable : nullptr,
			lr, b1, b2, 1e-8F, bc1 > 0.0F ? bc1 : 1.0F, bc2 > 0.0F ? bc2 : 1.0F, gradOut);
		if ((e = hipGetLastError()) != hipSuccess) return hipfail(e, "training launch");
		return 0; }

	int Wait() {
		if (!pending) return 0;
		pending = false;
		const hipError_t e = hipEventSynchronize(done);
		if (e != hipSuccess) return hipfail(e, "hipEventSynchronize");
		stats.primitives_clipped = hCtr->clipCount;
		uint32_t binned = 0;
		unsigned long long frags = 0;
		for (const CtrSlot& c : hCtr->slot) { binned += c.binned; frags += c.fragments; }
		stats.primitives_binned = binned;
		stats.tile_pairs = hCtr->pairs;
		stats.clipped_tile_pairs = hCtr->clipPairs;
		stats.fragments = static_cast<int64_t>(frags);
		stats.overflow = hCtr->overflow;
		stats.fallback_quads = hCtr->fallbackQuads;
		lastWorkItems = static_cast<int>(std::min<uint32_t>(hCtr->workItems, 1u << 20));
		for (int k = 0; k < kRing; ++k) Accumulate(k);
		if (hCtr->overflow & 8u) return fail(RGL_E_ARG, "vertex index outside the position buffer");
		if (hCtr->overflow & 16u) return fail(RGL_E_STATE, "a binning unit spans more tiles than its table holds");
		if (hCtr->overflow) {
			// grow and report; the caller runs the frame again
			if (hCtr->overflow & 1u) listCap = static_cast<uint32_t>(std::min<long long>(2LL * hCtr->pairs + 4096, 0x7fffffffLL));
			if (hCtr->overflow & 2u) clipCap = static_cast<uint32_t>(std::min<long long>(2LL * hCtr->clipCount + 1024, 0x7fffffffLL));
			if (hCtr->overflow & 4u) clistCap = static_cast<uint32_t>(std::min<long long>(2LL * hCtr->clipPairs + 4096, 0x7fffffffLL));
			return fail(RGL_E_CAPACITY, "a per-frame buffer overflowed; capacities were grown, run the frame again"); }
		return 0; }
};

namespace {
int check(rgl_gpu* g) { return g ? 0 : RGL_E_ARG; }
}

extern "C" {

int rgl_gpu_create(int device_ordinal, int concurrency, const char* guid, rgl_gpu** out) {
	if (!out) return RGL_E_ARG;
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || device_ordinal < 0 || device_ordinal >= n) return RGL_E_HIP;
	if (hipSetDevice(device_ordinal) != hipSuccess) return RGL_E_HIP;
	// set-up that later calls rely on without allocating: the network calls'
	// event, and this device's address of the no-allocation scratch array
	float* scr = nullptr;
	if (scratch_address(device_ordinal, &scr) != hipSuccess) return RGL_E_HIP;
	hipEvent_t aux = nullptr;
	if (hipEventCreateWithFlags(&aux, hipEventDisableTiming) != hipSuccess) return RGL_E_HIP;
	auto* g = new rgl_gpu();
	g->device = device_ordinal;
	g->guid = guid ? guid : "unnamed";
	g->scratchDev = scr;
	g->auxDone = aux;
	(void)concurrency;
	g->cs.reset();
	*out = g;
	return 0; }

int rgl_gpu_destroy(rgl_gpu* g) {
	if (!g) return RGL_E_ARG;
	(void)g->WaitAll();
	if (g->dbMode) g->EndDouble();
	if (g->auxPending) (void)hipEventSynchronize(g->auxDone);
	delete g;
	return 0; }

int rgl_gpu_install(rgl_gpu* g, int program_id, uint32_t state_key, int kind) {
	if (check(g)) return RGL_E_ARG;
	if (kind < 0 || kind > 2) return g->fail(RGL_E_ARG, "bad install kind");
	uint32_t key = (static_cast<uint32_t>(program_id) << 24) | state_key;
	if (kind == RGL_BIN_PROGRAM) key |= 0x80000000u;
	if (kind == RGL_BLT_PROGRAM) key |= 0x40000000u;
	if (std::find(g->installed.begin(), g->installed.end(), key) == g->installed.end()) g->installed.push_back(key);
	return 0; }

int rgl_gpu_reset(rgl_gpu* g, int width, int height, int tbx, int tby) {
	if (check(g)) return RGL_E_ARG;
	if (width <= 0 || height <= 0 || tbx <= 0 || tby <= 0) return g->fail(RGL_E_ARG, "bad size");
	// the reference's fragment cursors step 64 quads a row: tiles of at
	// most 128 x 128 pixels (rglr_fragmentcursor.hxx:14)
	if (tbx * 8 > kMaxAnyTileDim || tby * 8 > kMaxAnyTileDim)
		return g->fail(RGL_E_UNSUPPORTED, "tiles larger than 128x128 pixels are not supported");
	if (width > 65535 * tbx * 8 || height > 65535 * tby * 8) return g->fail(RGL_E_ARG, "target too large");
	// a double-buffered frame in flight holds copies of what it draws: the
	// next frame records while it runs
	if (g->pending && !g->dbMode) { const int r = g->Wait(); if (r != 0 && r != RGL_E_CAPACITY) return r; }
	g->W = width; g->H = height; g->tileW = tbx * 8; g->tileH = tby * 8;
	g->tilesX = (width + g->tileW - 1) / g->tileW; g->tilesY = (height + g->tileH - 1) / g->tileH;
	g->ResetGL();
	return 0; }

int rgl_gpu_run(rgl_gpu* g, void* hip_stream) {
	if (check(g)) return RGL_E_ARG;
	if (g->W <= 0) return g->fail(RGL_E_STATE, "Run before Reset");
	return g->Run(static_cast<hipStream_t>(hip_stream)); }

int rgl_gpu_wait(rgl_gpu* g) { if (check(g)) return RGL_E_ARG; return g->WaitAll(); }

int rgl_gpu_wait_previous(rgl_gpu* g) { if (check(g)) return RGL_E_ARG; return g->WaitPrevious(); }

int rgl_gpu_set_double_buffer(rgl_gpu* g, int enable) {
	if (check(g)) return RGL_E_ARG;
	if (enable < 0 || enable > rgl_gpu::kMaxExec) return g->fail(RGL_E_ARG, "enable is 0, 1, or 2 to 4 frames in flight");
	const int n = enable == 0 ? 1 : std::max(enable, 2);
	if (n == g->nExec && g->dbMode == (n > 1)) return 0;
	const int r = g->WaitAll();
	if (r != 0 && r != RGL_E_CAPACITY) return r;
	if (g->dbMode) g->EndDouble();
	return n > 1 ? g->BeginDouble(n) : 0; }

int rgl_gpu_stats(rgl_gpu* g, rgl_stats* out) {
	if (check(g) || !out) return RGL_E_ARG;
	*out = g->dbMode ? g->dbStats : g->stats;
	return 0; }

int rgl_gpu_set_timing(rgl_gpu* g, int enable) {
	if (check(g)) return RGL_E_ARG;
	for (int k = 0; k < rgl_gpu::kRing; ++k) { g->evUsed[k] = false; for (bool& u : g->kevUsed[k]) u = false; }
	g->timing = enable != 0;
	for (double& a : g->acc) a = 0.0;
	g->accRuns = 0;
	return 0; }

int rgl_gpu_kernel_times(rgl_gpu* g, double* ms_out, int max_kernels, char* names_out, int names_len) {
	if (check(g)) return RGL_E_ARG;
	for (int k = 0; k < rgl_gpu::kRing; ++k) g->Accumulate(k);
	std::string names;
	for (int i = 0; i < kNumTimed; ++i) {
		if (i < max_kernels && ms_out) ms_out[i] = g->accRuns ? g->acc[i] / g->accRuns : 0.0;
		names += (i ? "," : ""); names += kKernelNames[i]; }
	if (names_out && names_len > 0) { std::strncpy(names_out, names.c_str(), names_len - 1); names_out[names_len - 1] = 0; }
	for (double& a : g->acc) a = 0.0;
	g->accRuns = 0;
	return kNumTimed; }

const char* rgl_gpu_last_error(rgl_gpu* g) { return g ? g->err.c_str() : "null gpu"; }

#define RGL_STATE_SETTER(body) do { if (check(g)) return RGL_E_ARG; g->dirty = true; body; return 0; } while (0)

int rgl_gl_enable(rgl_gpu* g, int v) {
	if (check(g)) return RGL_E_ARG;
	g->dirty = true;
	if (v == GL_CULL_FACE) g->cs.cullingEnabled = true;
	else if (v == GL_SCISSOR_TEST) g->cs.scissorEnabled = true;
	else if (v == GL_BLEND) g->cs.blendingEnabled = true;
	else if (v == GL_DEPTH_TEST) g->cs.depthTestEnabled = true;
	else return g->fail(RGL_E_ARG, "unknown glEnable value");
	return 0; }
int rgl_gl_disable(rgl_gpu* g, int v) {
	if (check(g)) return RGL_E_ARG;
	g->dirty = true;
	if (v == GL_CULL_FACE) g->cs.cullingEnabled = false;
	else if (v == GL_SCISSOR_TEST) g->cs.scissorEnabled = false;
	else if (v == GL_BLEND) g->cs.blendingEnabled = false;
	else if (v == GL_DEPTH_TEST) g->cs.depthTestEnabled = false;
	else return g->fail(RGL_E_ARG, "unknown glEnable value");
	return 0; }
int rgl_gl_depth_func(rgl_gpu* g, int v) { RGL_STATE_SETTER(g->cs.depthFunc = v); }
int rgl_gl_depth_write_mask(rgl_gpu* g, int v) { RGL_STATE_SETTER(g->cs.depthWriteMask = v != 0); }
int rgl_gl_color_write_mask(rgl_gpu* g, int v) { RGL_STATE_SETTER(g->cs.colorWriteMask = v != 0); }
int rgl_gl_cull_face(rgl_gpu* g, int v) { RGL_STATE_SETTER(g->cs.cullFace = v); }
int rgl_gl_scissor(rgl_gpu* g, int x, int y, int w, int h) {
	RGL_STATE_SETTER(g->cs.scissorOrigin[0] = x; g->cs.scissorOrigin[1] = y; g->cs.scissorSize[0] = w; g->cs.scissorSize[1] = h); }
int rgl_gl_viewport(rgl_gpu* g, int x, int y, int w, int h) {
	RGL_STATE_SETTER(g->cs.viewportOrigin[0] = x; g->cs.viewportOrigin[1] = y; g->cs.hasViewportSize = true;
	                 g->cs.viewportSize[0] = w; g->cs.viewportSize[1] = h); }
int rgl_gl_use_program(rgl_gpu* g, int v) { RGL_STATE_SETTER(g->cs.programId = v); }
int rgl_gl_use_uniforms(rgl_gpu* g, int ofs) { RGL_STATE_SETTER(g->cs.uniformsOfs = ofs); }
int rgl_gl_alloc_uniform_buffer(rgl_gpu* g, const float* values, int n) {
	if (check(g)) return RGL_E_ARG;
	const int idx = static_cast<int>(g->ubuf.size());
	g->ubuf.resize(g->ubuf.size() + UNIFORM_BUFFER_SIZE, 0.0F);
	for (int i = 0; values && i < n && i < UNIFORM_BUFFER_SIZE; ++i) g->ubuf[idx + i] = values[i];
	return idx; }
float* rgl_gl_uniform_buffer_addr(rgl_gpu* g, int ofs) {
	if (!g || ofs < 0 || ofs >= static_cast<int>(g->ubuf.size())) return nullptr;
	return g->ubuf.data() + ofs; }
int rgl_gl_renderbuffer_type(rgl_gpu* g, int attachment, int type) {
	if (check(g)) return RGL_E_ARG;
	g->dirty = true;
	if (attachment == GL_DEPTH_ATTACHMENT) g->cs.depthAttachmentType = type;
	else if (attachment == GL_COLOR_ATTACHMENT0) g->cs.color0AttachmentType = type;
	return 0; }
int rgl_gl_bind_texture(rgl_gpu* g, int unit, const float* ptr, int w, int h, int stride, int mode) {
	if (check(g)) return RGL_E_ARG;
	if (unit < 0 || unit > 1) return g->fail(RGL_E_ARG, "texture unit must be 0 or 1");
	if (mode != 0 && mode != 1) return g->fail(RGL_E_ARG, "texture filter must be GL_NEAREST/LINEAR_MIPMAP_NEAREST");
	g->dirty = true;
	g->cs.tus[unit] = { ptr, w, h, stride, mode };
	return 0; }
int rgl_gl_bind_texture3(rgl_gpu* g, const float* ptr, int dim) {
	if (check(g)) return RGL_E_ARG;
	// the unit addresses (row * dim + column) & (dim^2 - 1), which wraps
	// around the map for a power-of-two dim
	if (ptr && (dim <= 0 || (dim & (dim - 1)) != 0 || dim > 8192))
		return g->fail(RGL_E_ARG, "depth texture dimension must be a power of two <= 8192");
	g->dirty = true;
	g->cs.tu3ptr = ptr; g->cs.tu3dim = dim;
	return 0; }
int rgl_gl_use_buffer(rgl_gpu* g, int slot, const float* ptr, int count) {
	if (check(g)) return RGL_E_ARG;
	if (slot < 0 || slot > 15) return g->fail(RGL_E_ARG, "buffer slot must be 0..15");
	g->dirty = true;
	g->cs.buffers[slot] = ptr;
	g->cs.bufferCounts[slot] = ptr ? count : 0;
	return 0; }
int rgl_gl_clear_color(rgl_gpu* g, float r, float gg, float b) {
	RGL_STATE_SETTER(g->cs.clearColor[0] = r; g->cs.clearColor[1] = gg; g->cs.clearColor[2] = b; g->cs.clearColor[3] = 1.0F); }
int rgl_gl_clear_depth(rgl_gpu* g, float v) { RGL_STATE_SETTER(g->cs.clearDepth = v); }
int rgl_gl_view_matrix(rgl_gpu* g, const float* m) { if (!m) return RGL_E_ARG; RGL_STATE_SETTER(std::memcpy(g->cs.viewMatrix, m, 64)); }
int rgl_gl_projection_matrix(rgl_gpu* g, const float* m) { if (!m) return RGL_E_ARG; RGL_STATE_SETTER(std::memcpy(g->cs.projectionMatrix, m, 64)); }
int rgl_gl_normal_matrix(rgl_gpu* g, const float* m) { if (!m) return RGL_E_ARG; RGL_STATE_SETTER(std::memcpy(g->cs.normalMatrix, m, 64)); }

int rgl_gl_draw_elements(rgl_gpu* g, int mode, int count, int type, const uint16_t* idx, int hint) {
	if (check(g)) return RGL_E_ARG;
	if (mode != 0 || type != 1) return g->fail(RGL_E_UNSUPPORTED, "only GL_TRIANGLES with GL_UNSIGNED_SHORT indices");
	(void)hint;
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ELEMENTS, count, 1, idx }); return 0; }
int rgl_gl_draw_arrays(rgl_gpu* g, int mode, int start, int count) {
	if (check(g)) return RGL_E_ARG;
	if (mode != 0 || start != 0) return g->fail(RGL_E_UNSUPPORTED, "only GL_TRIANGLES from 0");
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ARRAYS, count, 1, nullptr }); return 0; }
int rgl_gl_draw_elements_instanced(rgl_gpu* g, int mode, int count, int type, const uint16_t* idx, int n) {
	if (check(g)) return RGL_E_ARG;
	if (mode != 0 || type != 1) return g->fail(RGL_E_UNSUPPORTED, "only GL_TRIANGLES with GL_UNSIGNED_SHORT indices");
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ELEMENTS_INSTANCED, count, n, idx }); return 0; }
int rgl_gl_draw_arrays_instanced(rgl_gpu* g, int mode, int start, int count, int n) {
	if (check(g)) return RGL_E_ARG;
	if (mode != 0 || start != 0) return g->fail(RGL_E_UNSUPPORTED, "only GL_TRIANGLES from 0");
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ARRAYS_INSTANCED, count, n, nullptr }); return 0; }
int rgl_gl_clear(rgl_gpu* g, int bits) {
	if (check(g)) return RGL_E_ARG;
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::CLEAR, bits, 0, nullptr }); return 0; }
int rgl_gl_store_color_tc(rgl_gpu* g, uint32_t* dst, int gamma) {
	if (check(g)) return RGL_E_ARG;
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_TC, gamma ? 1 : 0, 0, dst }); return 0; }
int rgl_gl_store_depth(rgl_gpu* g, float* dst) {
	if (check(g)) return RGL_E_ARG;
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_DEPTH, 0, 0, dst }); return 0; }
int rgl_gl_store_color_fp(rgl_gpu* g, float* dst, int stride_px, int downsample) {
	if (check(g)) return RGL_E_ARG;
	if (dst == nullptr || stride_px < 0) return g->fail(RGL_E_ARG, "bad float colour target");
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_FP, downsample ? 1 : 0, stride_px, dst }); return 0; }
int rgl_gl_store_color_quads(rgl_gpu* g, float* dst, int stride_quads) {
	if (check(g)) return RGL_E_ARG;
	if (dst == nullptr || stride_quads < 0) return g->fail(RGL_E_ARG, "bad quad canvas target");
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_QUADS, 0, stride_quads, dst }); return 0; }
int rgl_kawase_blur(const float* src, float* dst, int width, int height, int src_stride, int dst_stride, int d,
                    void* hip_stream) {
	if (!src || !dst || src == dst || width <= 0 || height <= 0 || src_stride < width || dst_stride < width || d < 0)
		return RGL_E_ARG;
	const dim3 grid(static_cast<unsigned>((width + 63) / 64), static_cast<unsigned>((height + 3) / 4));
	hipLaunchKernelGGL(k_kawase, grid, dim3(256), 0, static_cast<hipStream_t>(hip_stream),
	                   reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), width, height, src_stride,
	                   dst_stride, d);
	return hipGetLastError() == hipSuccess ? 0 : RGL_E_HIP; }
int rgl_gl_store_primitive_id(rgl_gpu* g, uint32_t* dst) {
	if (check(g)) return RGL_E_ARG;
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_PRIM, 0, 0, dst }); return 0; }
int rgl_gl_reset(rgl_gpu* g) { if (check(g)) return RGL_E_ARG; g->ResetGL(); return 0; }

// Texture::maybe_make_mipmap (rglr_texture.cxx:33-81)
int rgl_texture_make_mipmap(float* buf, int dim) {
	if (!buf || dim < 1 || (dim & (dim - 1)) != 0 || dim > 4096) return RGL_E_ARG;
	const int stride = dim;
	int srcSize = dim, src = 0, dst = dim * dim;
	while (srcSize > 1) {
		for (int sy = 0; sy < srcSize; sy += 2) {
			int dstRow = dst;
			for (int sx = 0; sx < srcSize; sx += 2) {
				const int r1 = src + sy * stride + sx, r2 = src + (sy + 1) * stride + sx;
				for (int c = 0; c < 4; ++c) {
					const float sum = ((buf[r1 * 4 + c] + buf[(r1 + 1) * 4 + c]) + buf[r2 * 4 + c]) + buf[(r2 + 1) * 4 + c];
					buf[dstRow * 4 + c] = sum / 4.0F; }
				++dstRow; }
			dst += stride; }
		src += srcSize * stride;
		srcSize >>= 1; }
	return 0; }

/* ---- the asset path's texture preparation on the device ---- */
static const Srgb8Table& srgb8_table() {
	// ryg::srgb8_to_float (ryg-srgb.h:33-56): c = i / 255 in float, c / 12.92
	// up to 0.04045, else pow of the float base and exponent, to float
	static Srgb8Table tab = [] {
		Srgb8Table t;
		for (int i = 0; i < 256; ++i) {
			const float c = static_cast<float>(i) * (1.0F / 255.0F);
			t.v[i] = c <= 0.04045F ? c / 12.92F
			                       : static_cast<float>(std::pow(static_cast<double>((c + 0.055F) / 1.055F), static_cast<double>(2.4F))); }
		return t; }();
	return tab; }

int rgl_texture_from_rgba8(const uint8_t* src, int width, int height, int premultiply, float* dst, void* hip_stream) {
	if (!src || !dst || width <= 0 || height <= 0) return RGL_E_ARG;
	const long long n = static_cast<long long>(width) * height;
	hipLaunchKernelGGL(k_texture_from_rgba8, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
	                   static_cast<hipStream_t>(hip_stream), reinterpret_cast<const uchar4*>(src), reinterpret_cast<float4*>(dst),
	                   n, premultiply ? 1 : 0, srgb8_table());
	return hipGetLastError() == hipSuccess ? 0 : RGL_E_HIP; }

int rgl_texture_pow2(const float* src, int width, int height, float* dst, int dim, void* hip_stream) {
	if (!src || !dst || width <= 0 || height <= 0 || dim < 2 || (dim & (dim - 1)) != 0 || dim > 16384 ||
	    dim < width || dim < height) return RGL_E_ARG;
	const long long n = static_cast<long long>(dim) * dim;
	hipLaunchKernelGGL(k_texture_pow2, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
	                   static_cast<hipStream_t>(hip_stream), reinterpret_cast<const float4*>(src), width, height,
	                   reinterpret_cast<float4*>(dst), dim);
	return hipGetLastError() == hipSuccess ? 0 : RGL_E_HIP; }

int rgl_texture_mipmap(float* buf, int dim, void* hip_stream) {
	if (!buf || dim < 1 || (dim & (dim - 1)) != 0 || dim > 4096) return RGL_E_ARG;
	int srcRow = 0, dstRow = dim;
	for (int size = dim / 2; size >= 1; size >>= 1) {
		hipLaunchKernelGGL(k_mip_level, dim3(static_cast<unsigned>((size * size + 255) / 256)), dim3(256), 0,
		                   static_cast<hipStream_t>(hip_stream), reinterpret_cast<float4*>(buf), dim, srcRow, dstRow, size);
		srcRow = dstRow; dstRow += size; }
	return hipGetLastError() == hipSuccess ? 0 : RGL_E_HIP; }

/* ---- batches, graph replay and raster parameters (no reference counterpart) ---- */
int rgl_gpu_run_batch(rgl_gpu* g, rgl_gpu* const* frames, int n, void* hip_stream) {
	if (check(g)) return RGL_E_ARG;
	if (!frames) return g->fail(RGL_E_ARG, "null frame list");
	if (g->dbMode) return g->fail(RGL_E_UNSUPPORTED, "batches run single-buffered");
	return g->RunBatch(frames, n, static_cast<hipStream_t>(hip_stream)); }

int rgl_gpu_set_graph(rgl_gpu* g, int enable) {
	if (check(g)) return RGL_E_ARG;
	if (g->pending || g->dbMode) { const int r = g->WaitAll(); if (r != 0 && r != RGL_E_CAPACITY) return r; }
	g->graphMode = enable != 0;
	g->graphValid = false;
	// leaving graph replay frees the capture stream: a process has few
	// hardware queues (four by default), and streams beyond them share one
	if (!g->graphMode) {
		if (g->graphExec) (void)hipGraphExecDestroy(g->graphExec);
		g->graphExec = nullptr;
		if (g->capStream) (void)hipStreamDestroy(g->capStream);
		g->capStream = nullptr; }
	return 0; }

int rgl_gpu_graph_captures(rgl_gpu* g) { return g ? g->captures : RGL_E_ARG; }

int rgl_gpu_set_guard_band(rgl_gpu* g, int mode) {
	if (check(g)) return RGL_E_ARG;
	if (mode != 0 && mode != 1) return g->fail(RGL_E_ARG, "guard band mode is 0 (reference) or 1 (fixed)");
	g->fixedGuardBand = mode == 1;
	return 0; }

int rgl_gpu_set_raster_params(rgl_gpu* g, int seg_pairs, int split_quads) {
	if (check(g)) return RGL_E_ARG;
	if (seg_pairs < 64 || seg_pairs > 4096) return g->fail(RGL_E_ARG, "seg_pairs must be in [64, 4096]");
	if (split_quads < 0 || split_quads > (1 << 20)) return g->fail(RGL_E_ARG, "split_quads must be in [0, 2^20]");
	g->segPairs = seg_pairs;
	g->splitQuads = split_quads;
	g->graphValid = false;
	return 0; }

/* ---- learned shading (no reference counterpart) ---- */
int rgl_mlp_param_count(void) { return kMlpParams; }
int rgl_mlp_trainable_count(void) { return kMlpTrainable; }

int rgl_gpu_set_learned_shading(rgl_gpu* g, const float* dev_params, float max_rms_err) {
	if (check(g)) return RGL_E_ARG;
	if (!(max_rms_err >= 0.0F)) return g->fail(RGL_E_ARG, "max_rms_err must be >= 0");
	g->mlpParams = dev_params;
	g->mlpMaxRms = max_rms_err;
	return 0; }

int rgl_gl_store_training_samples(rgl_gpu* g, float* dev_out, int capacity, uint32_t* dev_count, float keep, uint32_t seed) {
	if (check(g)) return RGL_E_ARG;
	if (!dev_out || !dev_count || capacity < 0 || !(keep >= 0.0F && keep <= 1.0F)) return g->fail(RGL_E_ARG, "bad sample request");
	g->MaybeUpdateState();
	g->sampleReqs.push_back({ dev_out, static_cast<uint32_t>(capacity), dev_count, keep, seed });
	g->commands.push_back({ Cmd::STORE_SAMPLES, static_cast<int>(g->sampleReqs.size()) - 1, 0, nullptr });
	return 0; }

int rgl_mlp_forward(rgl_gpu* g, const float* dev_samples, int n, const float* dev_params, float* dev_out, void* hip_stream) {
	if (check(g)) return RGL_E_ARG;
	if (!dev_samples || !dev_params || !dev_out || n < 0) return g->fail(RGL_E_ARG, "bad forward arguments");
	if (n == 0) return 0;
	const hipStream_t ms = static_cast<hipStream_t>(hip_stream);
	int r = g->AuxBegin(ms);
	if (r != 0) return r;
	hipError_t e;
	if ((e = g->fwdImg.ensure(kImgFloats)) != hipSuccess) return g->hipfail(e, "hipMalloc network image");
	k_mlp_pack<<<16, 256, 0, ms>>>(dev_params, g->fwdImg.p);
	const int per = 64 * kMlpWaves;
	k_mlp_forward<<<(n + per - 1) / per, per, 0, ms>>>(dev_samples, n, g->fwdImg.p, dev_out);
	if ((e = hipGetLastError()) != hipSuccess) return g->hipfail(e, "forward launch");
	return g->AuxEnd(ms); }

int rgl_mlp_gradient(rgl_gpu* g, const float* dev_samples, int n, float* dev_params, float* dev_grad, float* loss_out,
                     void* hip_stream) {
	if (check(g)) return RGL_E_ARG;
	if (!dev_samples || !dev_params || !dev_grad || n <= 0) return g->fail(RGL_E_ARG, "bad gradient arguments");
	const hipStream_t ms = static_cast<hipStream_t>(hip_stream);
	int r = g->AuxBegin(ms);
	if (r != 0) return r;
	TrainBufs tb;
	if ((r = g->TrainBuffers(n, tb)) != 0) return r;
	r = g->TrainPass(ms, tb, dev_samples, n, n, dev_params, nullptr, 0.0F, 0.9F, 0.999F, 1, 0xffffffffu, dev_grad);
	if (r != 0) return r;
	if ((r = g->AuxEnd(ms)) != 0) return r;
	if (loss_out) {
		hipError_t e;
		float s = 0.0F;
		if ((e = hipMemcpyAsync(&s, g->tLoss.p, sizeof(float), hipMemcpyDeviceToHost, ms)) != hipSuccess) return g->hipfail(e, "loss copy");
		if ((e = hipStreamSynchronize(ms)) != hipSuccess) return g->hipfail(e, "loss sync");
		*loss_out = s / (3.0F * static_cast<float>(n)); }
	return 0; }

int rgl_mlp_train(rgl_gpu* g, const float* dev_samples, long long n, float* dev_params, float* dev_opt, int steps, int batch,
                  float lr, uint32_t seed, int step0, float* loss_out, void* hip_stream) {
	if (check(g)) return RGL_E_ARG;
	if (!dev_samples || !dev_params || !dev_opt || n <= 0 || steps < 0 || batch <= 0 || !(lr > 0.0F) || step0 < 0)
		return g->fail(RGL_E_ARG, "bad training arguments");
	const hipStream_t ms = static_cast<hipStream_t>(hip_stream);
	int r = g->AuxBegin(ms);
	if (r != 0) return r;
	TrainBufs tb;
	if ((r = g->TrainBuffers(batch, tb)) != 0) return r;
	for (int s = 0; s < steps; ++s) {
		r = g->TrainPass(ms, tb, dev_samples, n, batch, dev_params, dev_opt, lr, 0.9F, 0.99F, step0 + s + 1,
		                 seed == 0xffffffffu ? 0u : seed, nullptr);
		if (r != 0) return r; }
	if ((r = g->AuxEnd(ms)) != 0) return r;
	if (loss_out) {
		hipError_t e;
		float s = 0.0F;
		if ((e = hipMemcpyAsync(&s, g->tLoss.p, sizeof(float), hipMemcpyDeviceToHost, ms)) != hipSuccess) return g->hipfail(e, "loss copy");
		if ((e = hipStreamSynchronize(ms)) != hipSuccess) return g->hipfail(e, "loss sync");
		*loss_out = s / (3.0F * static_cast<float>(batch)); }
	return 0; }

/* ---- no allocation (no reference counterpart) ---- */
int rgl_mlp_train_noalloc(rgl_gpu* g, const float* dev_samples, long long n, float* dev_params, float* dev_opt, int steps,
                          int batch, float lr, uint32_t seed, int step0, float* dev_loss_out, void* hip_stream) {
	if (check(g)) return RGL_E_ARG;
	if (!dev_samples || !dev_params || !dev_opt || n <= 0 || steps < 0 || batch <= 0 || !(lr > 0.0F) || step0 < 0)
		return g->fail(RGL_E_ARG, "bad training arguments");
	if (batch > kNoAllocMaxBatch) return g->fail(RGL_E_CAPACITY, "rgl_mlp_train_noalloc: batch exceeds its scratch limit of 4096 samples");
	if (!g->scratchDev) return g->fail(RGL_E_STATE, "no scratch address for this device");
	const hipStream_t ms = static_cast<hipStream_t>(hip_stream);
	int r = g->AuxBegin(ms);
	if (r != 0) return r;
	const TrainBufs tb = g->ScratchBuffers();
	for (int s = 0; s < steps; ++s) {
		r = g->TrainPass(ms, tb, dev_samples, n, batch, dev_params, dev_opt, lr, 0.9F, 0.99F, step0 + s + 1,
		                 seed == 0xffffffffu ? 0u : seed, nullptr);
		if (r != 0) return r; }
	if (dev_loss_out && steps > 0) k_scale_copy<<<1, 64, 0, ms>>>(tb.Loss, 1.0F / (3.0F * static_cast<float>(batch)), dev_loss_out);
	hipError_t e;
	if ((e = hipGetLastError()) != hipSuccess) return g->hipfail(e, "training launch");
	return g->AuxEnd(ms); }

}  // extern "C"
