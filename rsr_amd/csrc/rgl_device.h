// rgl_device.h — device-side arithmetic, frame descriptors and shader
// programs of the gfx950 rgl pipeline.
//
// Bit-exact parity with the reference's SSE code rests on three rules that
// every function here keeps:
//   * float expressions are written in the reference's evaluation order and
//     compiled with -ffp-contract=off (no FMA is ever formed);
//   * division and sqrt are IEEE correctly rounded
//     (-fhip-fp32-correctly-rounded-divide-sqrt), denormals are kept;
//   * the x86 instructions whose results are not IEEE-defined are emulated
//     exactly: cvttps2dq (0x80000000 on overflow/NaN), rcpps (captured
//     table), maxps/minps (NaN returns the second operand), and 32-bit
//     wrapping integer products.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rgl {

#include "rcpps_table.inc"
#include "rsqrtps_table.inc"
#include "srgb8_table.inc"

constexpr int kMaxDraws = 256;      // draws per batch
constexpr int kMaxFrames = 64;      // frames (render targets) per batch
constexpr int kMaxVary = 16;       // varyings per vertex (the largest program's 15, padded to 16 bytes)
constexpr int kClipVerts = 8;      // 3 + one per clipping plane
constexpr int kFanMax = 6;         // kClipVerts - 2
constexpr int kSlotsPerTri = 1 + kFanMax;  // inline slot + fan slots

// ---------------------------------------------------------------------------
// SSE-faithful scalar arithmetic (reference src/rml/rmlv/rmlv_mvec4.hxx)
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
__host__ __device__ inline float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// _mm_cvttps_epi32 / cvttss2si (rmlv_mvec4.hxx:552)
__host__ __device__ inline int32_t cvtt(float f) {
	return (f >= -2147483648.0F && f < 2147483648.0F) ? static_cast<int32_t>(f) : INT32_MIN; }

__host__ __device__ inline float itof(int32_t i) { return static_cast<float>(i); }

// rcpps on the Intel host the reference output is defined on:
// result mantissa from an 11-bit table, exponent 253 - e.
__device__ inline float rcpps(float x) {
	const uint32_t in = f2u(x);
	const uint32_t e = (in >> 23) & 0xffu, s = in & 0x80000000u, m = in & 0x7fffffu;
	if (e == 0xffu) return u2f(m ? (in | 0x400000u) : s);
	if (e == 0u) return u2f(s | 0x7f800000u);
	const int re = 253 - static_cast<int>(e);
	if (re <= 0) return u2f(s);
	return u2f(s | (static_cast<uint32_t>(re) << 23) | kRcpMantissa[m >> 12]); }

// rsqrtps on the same host (rmlv::rsqrt, rmlv_mvec4.hxx:499-500): the result
// of a table entry chosen by the exponent's parity and the top 11 mantissa
// bits, its exponent shifted by half the input's
__device__ inline float rsqrtps(float x) {
	const uint32_t in = f2u(x);
	const uint32_t e = (in >> 23) & 0xffu, s = in & 0x80000000u, m = in & 0x7fffffu;
	if (e == 0xffu) return u2f(m ? (in | 0x400000u) : (s ? 0xffc00000u : 0u));
	if (e == 0u) return u2f(s | 0x7f800000u);
	if (s) return u2f(0xffc00000u);
	const bool even = (e & 1u) == 0u;
	const uint32_t t = kRsqrtTab[(even ? 2048u : 0u) + (m >> 12)];
	const int re = static_cast<int>((t >> 23) & 0xffu) - (even ? (static_cast<int>(e) - 128) / 2 : (static_cast<int>(e) - 127) / 2);
	return u2f((static_cast<uint32_t>(re) << 23) | (t & 0x7fffffu)); }

// rmlv::oneover: rcpps + one Newton step (rmlv_mvec4.hxx:645-649)
__device__ inline float oneover(float a) {
	const float r = rcpps(a);
	const float muls = a * (r * r);
	return (r + r) - muls; }

// the same with the mantissa table in LDS
__device__ inline float oneover_tab(float a, const uint32_t* tab) {
	const uint32_t in = f2u(a);
	const uint32_t e = (in >> 23) & 0xffu, s = in & 0x80000000u, m = in & 0x7fffffu;
	float r;
	if (e == 0xffu) r = u2f(m ? (in | 0x400000u) : s);
	else if (e == 0u) r = u2f(s | 0x7f800000u);
	else {
		const int re = 253 - static_cast<int>(e);
		r = re <= 0 ? u2f(s) : u2f(s | (static_cast<uint32_t>(re) << 23) | tab[m >> 12]); }
	const float muls = a * (r * r);
	return (r + r) - muls; }

__host__ __device__ inline float sse_max(float a, float b) { return a > b ? a : b; }
__host__ __device__ inline float sse_min(float a, float b) { return a < b ? a : b; }

__host__ __device__ inline int32_t wadd(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b)); }
__host__ __device__ inline int32_t wsub(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) - static_cast<uint32_t>(b)); }
__host__ __device__ inline int32_t wmul(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) * static_cast<uint32_t>(b)); }
__host__ __device__ inline int32_t shl(int32_t a, int n) { return static_cast<int32_t>(static_cast<uint32_t>(a) << n); }
__host__ __device__ inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
__host__ __device__ inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

// ryg float_to_srgb8_var2_SSE2 (3rdparty/ryg-srgb/ryg-srgb.h:183-219)
__device__ inline uint32_t to_srgb8(float f) {
	const float clamp1 = sse_max(f, u2f((127u - 13u) << 23));
	const float clamp2 = sse_min(clamp1, u2f(0x3f7fffffu));
	const uint32_t bits = f2u(clamp2);
	const uint32_t tab = kSrgb8Tab[(bits >> 20) - (127u - 13u) * 8u];
	const uint32_t t = (bits >> 12) & 0xffu;
	return ((tab & 0xffffu) * t + (tab >> 16) * 512u) >> 16; }

__device__ inline uint32_t to_srgb8_tab(float f, const uint32_t* tab4) {
	const float clamp1 = sse_max(f, u2f((127u - 13u) << 23));
	const float clamp2 = sse_min(clamp1, u2f(0x3f7fffffu));
	const uint32_t bits = f2u(clamp2);
	const uint32_t tab = tab4[(bits >> 20) - (127u - 13u) * 8u];
	const uint32_t t = (bits >> 12) & 0xffu;
	return ((tab & 0xffffu) * t + (tab >> 16) * 512u) >> 16; }

// to_tc_basic (rglr_canvas_util.hxx:40-48)
__device__ inline uint32_t to_linear8(float a) {
	float r = sse_min(a, 1.0F);
	r = sse_max(r, 0.0F);
	return static_cast<uint32_t>(cvtt(r * 255.0F)); }

// ---- the IQ post program's arithmetic (viewer/shaders.hxx:56-67) ----
// _mm_cvtps_epi32 under the default rounding mode: nearest, ties to even
__device__ inline int32_t cvtn(float f) {
	if (f >= -2147483648.0F && f < 2147483648.0F) return static_cast<int32_t>(__builtin_rintf(f));
	return INT32_MIN; }

// exp2 and log2 of J. R. Fonseca's SSE2 pow (3rdparty/sse-pow/sse_pow.h:
// 19-87 with EXP_POLY_DEGREE 3, LOG_POLY_DEGREE 5): the method's minimax
// coefficients, evaluated in Horner order as its POLYn macros nest
__device__ inline float sse_exp2(float x) {
	x = sse_min(x, 129.00000F);
	x = sse_max(x, -126.99999F);
	const int32_t ip = cvtn(x - 0.5F);
	const float fp = x - static_cast<float>(ip);
	const float scale = u2f(static_cast<uint32_t>(shl(wadd(ip, 127), 23)));
	float p = 7.8024521e-2F;
	p = p * fp + 2.2606716e-1F;
	p = p * fp + 6.9583356e-1F;
	p = p * fp + 9.9992520e-1F;
	return scale * p; }

__device__ inline float sse_log2(float x) {
	const uint32_t i = f2u(x);
	const float e = static_cast<float>(static_cast<int32_t>((i & 0x7F800000u) >> 23) - 127);
	const float m = u2f((i & 0x007FFFFFu) | 0x3F800000u);
	float p = static_cast<float>(0.0596515482674574969533);
	p = p * m + static_cast<float>(-0.465725644288844778798);
	p = p * m + static_cast<float>(1.48116647521213171641);
	p = p * m + static_cast<float>(-2.52074962577807006663);
	p = p * m + static_cast<float>(2.8882704548164776201);
	return p * (m - 1.0F) + e; }

// rmlv pow (rmlv_mvec4.hxx:652-654)
__device__ inline float sse_pow(float x, float y) { return sse_exp2(sse_log2(x) * y); }

// rmlv sin (rmlv_mvec4.hxx:730-746): x / pi wrapped to [-1, 1] (sign
// flipped on odd periods, wrap1), then y = M - M |M|, y (3.1 + 3.6 |y|)
__device__ inline float approx_sin(float x) {
	float m = x * static_cast<float>(1.0 / 3.14159265358979323846);
	const int32_t whole = cvtt(m);
	m = m - static_cast<float>(whole);
	m = u2f(f2u(m) ^ static_cast<uint32_t>(shl(whole, 31)));
	const float y = m - m * __builtin_fabsf(m);
	return y * (3.1F + 3.6F * __builtin_fabsf(y)); }

// rmlv fract (rmlv_mvec4.hxx:576-580): a - trunc(a) through cvttps2dq
__device__ inline float fract_t(float a) { return a - static_cast<float>(cvtt(a)); }

// IQPostProgram::ShadeCanvas of one pixel at canvas coordinate q:
// per-channel gamma, vignette, and 1/255 of iq's hash3 of q.x + 13 q.y
// (rglv_fragment.hxx:34-36)
__device__ inline void iq_post(float c[3], float qx, float qy) {
	c[0] = sse_pow(c[0], 0.45F); c[1] = sse_pow(c[1], 0.5F); c[2] = sse_pow(c[2], 0.55F);
	const float t = (((16.0F * qx) * qy) * (1.0F - qx)) * (1.0F - qy);
	const float v = 0.2F + 0.8F * sse_pow(t, 0.2F);
	const float n = qx + 13.0F * qy;
	const float k = static_cast<float>(1.0 / 255.0);
#pragma unroll
	for (int ch = 0; ch < 3; ++ch) {
		const float h = fract_t(approx_sin(n + static_cast<float>(ch)) * 43758.5453123F);
		c[ch] = c[ch] * v;
		c[ch] = c[ch] + k * h; }}

// FilterTile's canvas coordinate of pixel (x, y) in the tile whose corner
// is (left, top) (rglr_algorithm.hxx:36-62): the reference steps it per
// 2 x 2 quad, +2/W along a row and -2/H per row pair from the tile corner,
// so the same float additions are repeated here
__device__ inline void filter_q(int W, int H, int left, int top, int x, int y, float& qx, float& qy) {
	const float iw = 1.0F / static_cast<float>(W), ih = 1.0F / static_cast<float>(H);
	const float dx = iw * 2.0F, dy = -ih * 2.0F;
	qx = (static_cast<float>(left) + 0.5F) / static_cast<float>(W);
	if ((x - left) & 1) qx = qx + iw;
	for (int k = (x - left) >> 1; k > 0; --k) qx = qx + dx;
	qy = (static_cast<float>(H - top) - 0.5F) / static_cast<float>(H);
	if ((y - top) & 1) qy = qy - ih;
	for (int k = (y - top) >> 1; k > 0; --k) qy = qy + dy; }

// StoreTrueColor's ShadeCanvas of a pixel: post program 1 (identity),
// 2 (exposure) or 3 (IQ) (viewer/shaders.hxx:32-68)
__device__ inline void post_shade(int post, float exposure, float c[3], int W, int H, int left, int top, int x, int y) {
	if (post == 2) { c[0] = c[0] * exposure; c[1] = c[1] * exposure; c[2] = c[2] * exposure; }
	else if (post == 3) {
		float qx, qy;
		filter_q(W, H, left, top, x, y, qx, qy);
		iq_post(c, qx, qy); }}

// qmat4 * qfloat4 (rmlm_soa.hxx:58-64), column-major f
__device__ inline void xform4(const float* f, const float b[4], float o[4]) {
	o[0] = ((f[0]*b[0] + f[4]*b[1]) + f[ 8]*b[2]) + f[12]*b[3];
	o[1] = ((f[1]*b[0] + f[5]*b[1]) + f[ 9]*b[2]) + f[13]*b[3];
	o[2] = ((f[2]*b[0] + f[6]*b[1]) + f[10]*b[2]) + f[14]*b[3];
	o[3] = ((f[3]*b[0] + f[7]*b[1]) + f[11]*b[2]) + f[15]*b[3]; }

// mul_w0 (rmlm_soa.hxx:42-47)
__device__ inline void xform3w0(const float* f, const float b[3], float o[3]) {
	o[0] = (f[0]*b[0] + f[4]*b[1]) + f[ 8]*b[2];
	o[1] = (f[1]*b[0] + f[5]*b[1]) + f[ 9]*b[2];
	o[2] = (f[2]*b[0] + f[6]*b[1]) + f[10]*b[2]; }

__device__ inline void normalize3(float& x, float& y, float& z) {
	const float len2 = (x*x + y*y) + z*z;
	const float inv = 1.0F / __builtin_sqrtf(len2);
	x = x * inv; y = y * inv; z = z * inv; }

// depth ordering key: float order as unsigned, -0 == +0
__device__ inline uint32_t depth_order(float d) {
	uint32_t u = f2u(d);
	if (u == 0x80000000u) u = 0u;
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u); }

// ---------------------------------------------------------------------------
// frame descriptors (host fills, device reads)
// ---------------------------------------------------------------------------
struct SamplerDesc {
	const float* buf;
	int kind;      // -1 none, 0 nearest-mip, 1 linear-mip, 2 non-power-of-two nearest
	int power, width, height, stride;
	int texels;    // number of RGBA texels addressable (bounds for the UB cases)
};

struct DrawDesc {
	int program;
	int instances;           // I
	int tris;                // T per instance
	int verts;               // V: vertex-cache stride per instance
	int nvary;               // varyings per vertex for this program
	int elements;            // 1: indexed
	long long triBase;       // first global triangle index of this draw
	long long vBase;         // first vertex-cache entry
	uint32_t slotBase;       // first visibility slot: 1 + 7 * triBase
	uint32_t primBase;       // first primitive id (== triBase)
	const uint16_t* indices;
	const float* buffers[16];
	float vpm[16];           // projection * view (MakeMatrices, rglv_gpu_impl.hxx:44-51)
	float vm[16];            // view
	float nm[16];            // transpose(inverse(view))
	float DS[2], DO[2];      // GPU::DSDO (rglv_gpu.hxx:265-270)
	float gbFactor;          // CalcGuardBandFactor (rglv_view_frustum.hxx:36-39)
	int scissor[4];          // x0, y0, x1, y1 (GPU::ScissorRect)
	int keepBacks, keepFronts;
	int cullingEnabled, cullFace;
	// fragment state of the in-order tile path (FragmentStateKey,
	// rglv_gl.hxx:136-146): depth test and function, depth and colour write
	// masks, alpha blending, scissor (scissor[] then clips the raster rect)
	int depthTest, depthFunc, depthWrite, colorWrite, blend, scissorOn;
	float uniforms[32];      // UniformsSD (<= 128 bytes)
	SamplerDesc tu0, tu1;
	const float* tu3;        // DepthTextureUnit (rglr_texture_sampler.hxx:33-79): dim x dim floats
	int tu3dim;
	// the frame this draw renders into and its tile grid
	int frame;
	int tileW, tileH, tilesX, tilesY, tileBase;
	int fW, fH;    // the frame's size
};

// one render target of a batch (a GPU::Reset'd canvas with its clear,
// depth function and stores)
struct FrameDesc {
	int W, H, tileW, tileH, tilesX, tilesY, nTiles;
	int tileBase;            // first batch tile of this frame
	long long keyBase;       // first per-pixel key of this frame (row-major, stride W)
	int blockBase;           // first resolve block of this frame
	int blocksX;             // resolve blocks across (kResolveW pixels each)
	int depthFunc;           // GL_LESS or GL_LEQUAL for the whole frame
	int drawEnabled;         // 0 when the clear depth makes every test fail
	float clearColor[3];
	float clearDepth;
	int doClearColor, doClearDepth;
	uint32_t* outColor; int gamma; int postProgram; float exposure;
	float* outDepth;
	uint32_t* outPrim;
	// 1: drawn by the in-order tile path (its commands opBase .. opBase +
	// nOps of the batch's frame-command table apply in submission order)
	int inOrder, opBase, nOps;
};

// a batch of frames drawn by one launch sequence: triangles, tiles and
// resolve blocks are numbered across the batch
struct BatchDesc {
	int nFrames;
	int nTiles;              // sum of the frames' tiles
	int nDraws;
	int units;               // binning units (unitTris triangles each)
	int unitTris;            // triangles per binning unit (kUnitTris)
	long long totalTris;     // S
	int nBlocks;             // resolve blocks, all frames
	int segPairs;            // (triangle, tile) pairs per visibility work item
	int splitQuads;          // quad work per quadrant from which it is cut in 2, and from twice this in 4 (0: never)
	int unitWindow;          // most tiles one binning unit's frames span (its LDS tile table)
	const FrameDesc* frames; // device table, nFrames entries
};

// last k in [0, n) with base(k) <= v, for a nondecreasing table
__device__ inline int find_base(const int* base, int n, int v) {
	int lo = 0, hi = n - 1;
	while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (base[mid] <= v) lo = mid; else hi = mid - 1; }
	return lo; }

// ---------------------------------------------------------------------------
// texture units (rglr_texture_sampler.cxx)
// ---------------------------------------------------------------------------
// LevelOfDetail (rglr_texture_sampler.cxx:25-42): coarse, from lanes 0 and 1
__device__ inline int level_of_detail(float u0, float u1, float v0, float v1) {
	const float dux = u1 - u0;
	const float dvx = v1 - v0;
	const float sqd = dux*dux + dvx*dvx;
	return (static_cast<int32_t>(f2u(sqd)) - (127 << 23)) >> 24; }

__device__ inline int clamp_texel(int32_t ofs4, int texels) {
	// offsets outside the buffer are undefined behaviour in the reference
	// (negative wrap coordinates); clamp so the device never faults
	const int32_t t = ofs4 >> 2;
	return (t < 0 || t >= texels) ? 0 : ofs4; }

// texture LOD of a quad from the base-level coordinates of its lanes 0
// and 1, clamped to the mip chain (LevelOfDetail and its clamp,
// rglr_texture_sampler.cxx:25-42, :81, :188)
__device__ inline int texture_lod(const SamplerDesc& s, float u0, float u1, float v0, float v1) {
	if (s.kind < 0 || s.kind == 2) return 0;
	const float baseDim = static_cast<float>(1 << s.power);
	const int lod = level_of_detail(u0 * baseDim, u1 * baseDim, v0 * baseDim, v1 * baseDim);
	return lod < 0 ? 0 : (lod > s.power ? s.power : lod); }

// one lane of TextureUnit::sample (rglr_texture_sampler.cxx:63-311) at a
// given LOD; every lane of a quad uses the quad's LOD
__device__ inline float4 sample1(const SamplerDesc& s, float u, float v, int lod) {
	if (s.kind < 0 || s.buf == nullptr) return make_float4(0.0F, 0.0F, 0.0F, 0.0F);
	const float almostone = u2f(0x3f7fffffu);
	if (s.kind == 2) {  // TextureUnitRGBAF32_NM_ONEMAP_WRAP_NEAREST (:289-311)
		const float fu = u + 100.0F, fv = v + 100.0F;
		const float fru = fu - itof(cvtt(fu)), frv = fv - itof(cvtt(fv));
		const int32_t pu = cvtt(fru * static_cast<float>(s.width));
		const int32_t pv = cvtt((almostone - frv) * static_cast<float>(s.height));
		const int ofs = clamp_texel(shl(wadd(wmul(pv, s.stride), pu), 2), s.texels);
		return *reinterpret_cast<const float4*>(s.buf + ofs); }
	const int power = s.power;
	if (s.kind == 0) {  // ..._P2_MIPMAP_WRAP_NEAREST (:63-109)
		const float levelDim = static_cast<float>(1 << (power - lod));
		const int32_t levelBeginRow = static_cast<int32_t>((0xfffffffeu << (power - lod)) & ((1u << (power + 1)) - 1u));
		const float fu = u + 10.0F, fv = v + 10.0F;
		const float fru = fu - itof(cvtt(fu)), frv = fv - itof(cvtt(fv));
		const int32_t lx = cvtt(fru * levelDim);
		const int32_t ly = cvtt((almostone - frv) * levelDim);
		const int ofs = clamp_texel(shl(wadd(shl(wadd(levelBeginRow, ly), power), lx), 2), s.texels);
		return *reinterpret_cast<const float4*>(s.buf + ofs); }
	// ..._P2_MIPMAP_WRAP_LINEAR (:172-286)
	const int levelDim = 1 << (power - lod);
	const int32_t wrapMask = levelDim - 1;
	const int32_t levelLastRow = static_cast<int32_t>((0xffffffffu << (power - lod)) & ((1u << (power + 1)) - 1u)) - 1;
	const float lx = u * static_cast<float>(levelDim);
	const float ly = v * static_cast<float>(levelDim);
	int32_t tx0 = cvtt(lx - 0.5F), ty0 = cvtt(ly - 0.5F);
	int32_t tx1 = wadd(tx0, 1), ty1 = wadd(ty0, 1);
	const float fx = (lx - itof(tx0)) - 0.5F;
	const float fy = (ly - itof(ty0)) - 0.5F;
	const float fx1 = 1.0F - fx, fy1 = 1.0F - fy;
	const float w00 = fx1 * fy1, w10 = fx * fy1, w01 = fx1 * fy, w11 = fx * fy;
	tx0 &= wrapMask; ty0 &= wrapMask; tx1 &= wrapMask; ty1 &= wrapMask;
	const int32_t by0 = levelLastRow - ty0, by1 = levelLastRow - ty1;
	const float4 p00 = *reinterpret_cast<const float4*>(s.buf + ((((by0 << power) + tx0)) << 2));
	const float4 p10 = *reinterpret_cast<const float4*>(s.buf + ((((by0 << power) + tx1)) << 2));
	const float4 p01 = *reinterpret_cast<const float4*>(s.buf + ((((by1 << power) + tx0)) << 2));
	const float4 p11 = *reinterpret_cast<const float4*>(s.buf + ((((by1 << power) + tx1)) << 2));
	return make_float4(((p00.x * w00 + p10.x * w10) + p01.x * w01) + p11.x * w11,
	                   ((p00.y * w00 + p10.y * w10) + p01.y * w01) + p11.y * w11,
	                   ((p00.z * w00 + p10.z * w10) + p01.z * w01) + p11.z * w11,
	                   ((p00.w * w00 + p10.w * w10) + p01.w * w01) + p11.w * w11); }

}  // namespace rgl
