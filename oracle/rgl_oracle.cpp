/*
 * rgl_oracle — CPU restatement of the rgl tile rasteriser (usrlocalben/rsr,
 * src/rgl/rglv + src/rgl/rglr + src/rml), used ONLY as test infrastructure:
 * the parity checker for the HIP pipeline and the timed CPU baseline.
 * Nothing in the product links, loads or calls this file.
 *
 * It follows the reference's algorithm step by step, lane by lane, with
 * the SSE arithmetic restated in scalar C++:
 *   - cvttps2dq truncation with 0x80000000 for out-of-range/NaN;
 *   - rcpps from a captured 2048-entry table (rcpps_table.inc), then the
 *     one Newton step of rmlv::oneover (rmlv_mvec4.hxx:630-650);
 *   - 32-bit wrapping integer products (sse2_mul32, rmlv_mmutil.hxx:22-34);
 *   - maxps/minps operand order (NaN returns the second operand);
 *   - every float expression in the reference's evaluation order, compiled
 *     with -ffp-contract=off so no FMA is formed.
 *
 * Differences that do not change results: per-tile command lists are
 * arrays of small structs instead of packed bytes, and every triangle entry
 * also carries its primitive id (instance * triangles_per_instance + index
 * within the draw + the draw's base), so a primitive-ID buffer can be
 * written next to colour and depth.  The reference has no such buffer.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <vector>

namespace {

#include "rcpps_table.inc"
#include "rsqrtps_table.inc"
#include "srgb8_table.inc"

// ------------------------------------------------------------------------
// SSE-faithful scalar arithmetic (rmlv_mvec4.hxx)
// ------------------------------------------------------------------------
inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// _mm_cvttps_epi32 (rmlv_mvec4.hxx:552)
inline int32_t cvtt(float f) {
	if (f >= -2147483648.0F && f < 2147483648.0F) return static_cast<int32_t>(f);
	return INT32_MIN; }

// _mm_cvtepi32_ps
inline float itof(int32_t i) { return static_cast<float>(i); }

// _mm_rcp_ps, Intel table model (see gen_rcpps_table.c)
inline float rcpps(float x) {
	const uint32_t in = f2u(x);
	const uint32_t e = (in >> 23) & 0xffu, s = in & 0x80000000u, m = in & 0x7fffffu;
	if (e == 0xffu) return u2f(m ? (in | 0x400000u) : s);
	if (e == 0u) return u2f(s | 0x7f800000u);
	const int re = 126 + 127 - static_cast<int>(e);
	if (re <= 0) return u2f(s);
	return u2f(s | (static_cast<uint32_t>(re) << 23) | kRcpMantissa[m >> 12]); }

// _mm_rsqrt_ps, Intel table model (see gen_rsqrtps_table.c)
inline float rsqrtps(float x) {
	const uint32_t in = f2u(x);
	const uint32_t e = (in >> 23) & 0xffu, s = in & 0x80000000u, m = in & 0x7fffffu;
	if (e == 0xffu) return u2f(m ? (in | 0x400000u) : (s ? 0xffc00000u : 0u));
	if (e == 0u) return u2f(s | 0x7f800000u);
	if (s) return u2f(0xffc00000u);
	const bool even = (e & 1u) == 0u;
	const uint32_t t = kRsqrtTab[(even ? 2048u : 0u) + (m >> 12)];
	const int re = static_cast<int>((t >> 23) & 0xffu) - (even ? (static_cast<int>(e) - 128) / 2 : (static_cast<int>(e) - 127) / 2);
	return u2f((static_cast<uint32_t>(re) << 23) | (t & 0x7fffffu)); }

// rmlv::oneover: rcp + one Newton-Raphson step (rmlv_mvec4.hxx:645-649)
inline float oneover(float a) {
	const float r = rcpps(a);
	const float muls = a * (r * r);
	return (r + r) - muls; }

inline float sse_max(float a, float b) { return a > b ? a : b; }  // maxps
inline float sse_min(float a, float b) { return a < b ? a : b; }  // minps

inline int32_t wadd(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b)); }
inline int32_t wsub(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) - static_cast<uint32_t>(b)); }
inline int32_t wmul(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) * static_cast<uint32_t>(b)); }
inline int32_t sar(int32_t a, int n) { return a >> n; }  // arithmetic on this target
inline int32_t shl(int32_t a, int n) { return static_cast<int32_t>(static_cast<uint32_t>(a) << n); }

// ryg float_to_srgb8_var2_SSE2 (3rdparty/ryg-srgb/ryg-srgb.h:183-219)
inline uint32_t to_srgb8(float f) {
	const float clamp1 = sse_max(f, u2f((127u - 13u) << 23));
	const float clamp2 = sse_min(clamp1, u2f(0x3f7fffffu));
	const uint32_t bits = f2u(clamp2);
	const uint32_t tab = kSrgb8Tab[(bits >> 20) - (127u - 13u) * 8u];
	const uint32_t t = (bits >> 12) & 0xffu;
	const uint32_t prod = (tab & 0xffffu) * t + (tab >> 16) * 512u;  // pmaddwd of positive halves
	return prod >> 16; }

// to_tc_basic (rglr_canvas_util.hxx:40-48)
inline uint32_t to_linear8(float a) {
	float r0 = sse_min(a, 1.0F);
	r0 = sse_max(r0, 0.0F);
	r0 = r0 * 255.0F;
	return static_cast<uint32_t>(cvtt(r0)); }

// ------------------------------------------------------------------------
// mat4, column-major storage (rmlm_mat4.hxx:157-207)
// ------------------------------------------------------------------------
struct mat4 { float ff[16]; };

inline mat4 mul(const mat4& l, const mat4& r) {
	mat4 out;
	for (int row = 0; row < 4; ++row) {
		for (int col = 0; col < 4; ++col) {
			float ax = l.ff[0*4 + row] * r.ff[col*4 + 0];
			ax += l.ff[1*4 + row] * r.ff[col*4 + 1];
			ax += l.ff[2*4 + row] * r.ff[col*4 + 2];
			ax += l.ff[3*4 + row] * r.ff[col*4 + 3];
			out.ff[col*4 + row] = ax; }}
	return out; }

// qmat4 * qfloat4 (rmlm_soa.hxx:58-64)
inline void xform4(const float* f, const float b[4], float o[4]) {
	o[0] = ((f[0]*b[0] + f[4]*b[1]) + f[ 8]*b[2]) + f[12]*b[3];
	o[1] = ((f[1]*b[0] + f[5]*b[1]) + f[ 9]*b[2]) + f[13]*b[3];
	o[2] = ((f[2]*b[0] + f[6]*b[1]) + f[10]*b[2]) + f[14]*b[3];
	o[3] = ((f[3]*b[0] + f[7]*b[1]) + f[11]*b[2]) + f[15]*b[3]; }

// mul_w0 (rmlm_soa.hxx:42-47)
inline void xform3w0(const float* f, const float b[3], float o[3]) {
	o[0] = (f[0]*b[0] + f[4]*b[1]) + f[ 8]*b[2];
	o[1] = (f[1]*b[0] + f[5]*b[1]) + f[ 9]*b[2];
	o[2] = (f[2]*b[0] + f[6]*b[1]) + f[10]*b[2]; }

// rmlm::inverse (rmlm_mat4.cxx:23-167): each element a signed sum of six
// triple products of the column-major elements, summed left to right, then
// scaled by 1 / det with det = m0 inv0 + m1 inv4 + m2 inv8 + m3 inv12
struct InvTerm { signed char s; unsigned char a, b, c; };
const InvTerm kInvTerms[16][6] = {
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

inline void mat4_inverse(const float* m, float* inv) {
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
inline void mat4_transpose(const float* m, float* out) {
	for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) out[c * 4 + r] = m[r * 4 + c]; }

// ------------------------------------------------------------------------
// protocol and state (rglv_gpu_protocol.hxx, rglv_gl.hxx)
// ------------------------------------------------------------------------
constexpr int GL_CULL_FACE = 1, GL_SCISSOR_TEST = 2, GL_BLEND = 3, GL_DEPTH_TEST = 4;
constexpr int GL_FRONT = 1, GL_BACK = 2;
constexpr int GL_LESS = 0, GL_LEQUAL = 1, GL_EQUAL = 2;
constexpr int GL_COLOR_BUFFER_BIT = 1, GL_DEPTH_BUFFER_BIT = 2;
constexpr int GL_DEPTH_ATTACHMENT = 0, GL_COLOR_ATTACHMENT0 = 2;
constexpr int RB_COLOR_DEPTH = 0, RB_RGBF32 = 1, RB_RGBAF32 = 2;
constexpr int UNIFORM_BUFFER_SIZE = 32;

struct TextureState { const float* ptr; int width, height, stride, filter; };

struct GLState {
	float clearColor[4];
	float clearDepth;
	bool cullingEnabled;
	int cullFace;
	bool scissorEnabled;
	int scissorOrigin[2], scissorSize[2];
	int viewportOrigin[2];
	bool hasViewportSize;
	int viewportSize[2];
	bool blendingEnabled, colorWriteMask, depthWriteMask, depthTestEnabled;
	int depthFunc;
	int programId;
	int uniformsOfs;
	int color0AttachmentType, depthAttachmentType;
	const float* buffers[16];
	TextureState tus[2];
	const float* tu3ptr; int tu3dim;   // GL::BindTexture3 (rglv_gl.hxx:275-278)
	mat4 viewMatrix, normalMatrix, projectionMatrix;

	void reset() {  // GLState::reset (rglv_gl.hxx:152-179)
		std::memset(this, 0, sizeof(*this));
		clearColor[3] = 1.0F;
		clearDepth = 1.0F;
		cullFace = GL_BACK;
		colorWriteMask = true;
		depthWriteMask = true;
		depthTestEnabled = true;
		depthFunc = GL_LESS;
		uniformsOfs = -1;
		for (int i = 0; i < 2; ++i) { tus[i].ptr = nullptr; tus[i].width = tus[i].height = tus[i].stride = 8; tus[i].filter = 0; }
		tu3ptr = nullptr; tu3dim = 256; }};

struct Cmd {
	enum Kind { STATE, CLEAR, DRAW_ELEMENTS, DRAW_ARRAYS, DRAW_ELEMENTS_INSTANCED, DRAW_ARRAYS_INSTANCED,
	            STORE_TC, STORE_DEPTH, STORE_PRIM, STORE_FP } kind;
	int a, b, c;
	const void* ptr; };

// ------------------------------------------------------------------------
// texture units (rglr_texture_sampler.cxx)
// ------------------------------------------------------------------------
inline int ilog2(unsigned x) { int p = 0; while (x != 0U) { x >>= 1; ++p; } return p - 1; }

// LevelOfDetail (rglr_texture_sampler.cxx:25-42), lanes 0 and 1 only
inline int level_of_detail(const float u[4], const float v[4]) {
	const float dux = u[1] - u[0];
	const float dvx = v[1] - v[0];
	const float sqd = dux*dux + dvx*dvx;
	return (static_cast<int32_t>(f2u(sqd)) - (127 << 23)) >> 24; }

struct Sampler {
	int kind = -1;  // 0 nearest-mip, 1 linear-mip, 2 npot nearest
	int power = 0, width = 0, height = 0, stride = 0;
	const float* buf = nullptr;

	static Sampler Make(const TextureState& t, std::string& err) {
		Sampler s;
		s.buf = t.ptr; s.width = t.width; s.height = t.height; s.stride = t.stride;
		const int power = ilog2(static_cast<unsigned>(t.width));
		const bool isPow2 = (1 << power) == t.width && t.width == t.height && t.stride == t.width;
		if (!isPow2) { s.kind = 2; return s; }
		if (t.width < 4 || t.width > 1024) { err = "can't make TextureUnit for pow2 size"; return s; }
		s.power = power;
		s.kind = t.filter ? 1 : 0;
		return s; }

	// one quad, 4 lanes -> rgba[4][4] (channel, lane)
	void sample(const float u[4], const float v[4], float out[4][4]) const {
		if (buf == nullptr) { for (int c = 0; c < 4; ++c) for (int l = 0; l < 4; ++l) out[c][l] = 0.0F; return; }
		const float almostone = u2f(0x3f7fffffu);
		if (kind == 2) {  // TextureUnitRGBAF32_NM_ONEMAP_WRAP_NEAREST (:289-311)
			for (int l = 0; l < 4; ++l) {
				const float fu = u[l] + 100.0F, fv = v[l] + 100.0F;
				const float fru = fu - itof(cvtt(fu)), frv = fv - itof(cvtt(fv));
				const int32_t pu = cvtt(fru * static_cast<float>(width));
				const int32_t pv = cvtt((almostone - frv) * static_cast<float>(height));
				const int32_t ofs = shl(wadd(wmul(pv, stride), pu), 2);
				for (int c = 0; c < 4; ++c) out[c][l] = buf[ofs + c]; }
			return; }
		const float baseDim = static_cast<float>(1 << power);
		float bu[4], bv[4];
		for (int l = 0; l < 4; ++l) { bu[l] = u[l] * baseDim; bv[l] = v[l] * baseDim; }
		const int lod = std::clamp(level_of_detail(bu, bv), 0, power);
		if (kind == 0) {  // ..._P2_MIPMAP_WRAP_NEAREST (:63-109)
			const float levelDim = static_cast<float>(1 << (power - lod));
			const int32_t levelBeginRow = static_cast<int32_t>((0xfffffffeu << (power - lod)) & ((1u << (power + 1)) - 1u));
			for (int l = 0; l < 4; ++l) {
				const float fu = u[l] + 10.0F, fv = v[l] + 10.0F;
				const float fru = fu - itof(cvtt(fu)), frv = fv - itof(cvtt(fv));
				const int32_t lx = cvtt(fru * levelDim);
				const int32_t ly = cvtt((almostone - frv) * levelDim);
				const int32_t by = wadd(levelBeginRow, ly);
				const int32_t ofs = shl(wadd(shl(by, power), lx), 2);
				for (int c = 0; c < 4; ++c) out[c][l] = buf[ofs + c]; }
			return; }
		// ..._P2_MIPMAP_WRAP_LINEAR (:172-286)
		const int levelDim = 1 << (power - lod);
		const int32_t wrapMask = levelDim - 1;
		const int32_t levelLastRow = static_cast<int32_t>((0xffffffffu << (power - lod)) & ((1u << (power + 1)) - 1u)) - 1;
		for (int l = 0; l < 4; ++l) {
			const float lx = u[l] * static_cast<float>(levelDim);
			const float ly = v[l] * static_cast<float>(levelDim);
			int32_t tx0 = cvtt(lx - 0.5F), ty0 = cvtt(ly - 0.5F);
			int32_t tx1 = wadd(tx0, 1), ty1 = wadd(ty0, 1);
			const float fx = (lx - itof(tx0)) - 0.5F;
			const float fy = (ly - itof(ty0)) - 0.5F;
			const float fx1 = 1.0F - fx, fy1 = 1.0F - fy;
			const float w00 = fx1 * fy1, w10 = fx * fy1, w01 = fx1 * fy, w11 = fx * fy;
			tx0 &= wrapMask; ty0 &= wrapMask; tx1 &= wrapMask; ty1 &= wrapMask;
			const int32_t by0 = wsub(levelLastRow, ty0), by1 = wsub(levelLastRow, ty1);
			const int32_t o00 = shl(wadd(shl(by0, power), tx0), 2);
			const int32_t o10 = shl(wadd(shl(by0, power), tx1), 2);
			const int32_t o01 = shl(wadd(shl(by1, power), tx0), 2);
			const int32_t o11 = shl(wadd(shl(by1, power), tx1), 2);
			for (int c = 0; c < 4; ++c) {
				float ax = buf[o00 + c] * w00;
				ax = ax + buf[o10 + c] * w10;
				ax = ax + buf[o01 + c] * w01;
				ax = ax + buf[o11 + c] * w11;
				out[c][l] = ax; }}}};

// ------------------------------------------------------------------------
// shader programs (rglv_gpu_shaders.hxx concept)
// ------------------------------------------------------------------------
constexpr int kMaxVary = 16;
struct VOut { float f[kMaxVary]; };

// DepthTextureUnit (rglr_texture_sampler.hxx:33-79): nearest, border -1
// outside [0, 1)^2, offset wrapped by dim^2 - 1
struct DepthUnit {
	const float* buf = nullptr;
	int dim = 256;
	void sample(const float x[4], const float y[4], float out[4]) const {
		const float dimf = static_cast<float>(dim);
		for (int l = 0; l < 4; ++l) {
			const bool hit = (x[l] >= 0.0F) && (x[l] < 1.0F) && (y[l] >= 0.0F) && (y[l] < 1.0F);
			const int32_t px = cvtt(x[l] * dimf);
			const int32_t py = cvtt((1.0F - y[l]) * dimf);
			const int32_t ofs = wadd(wmul(py, dim), px) & wsub(wmul(dim, dim), 1);
			const float c = buf ? buf[ofs] : 0.0F;
			out[l] = hit ? c : -1.0F; }}};

struct DrawEnv {
	const GLState* st;
	const float* uniforms;  // UniformsSD, may be null
	mat4 vpm, vm, nm;       // MakeMatrices (rglv_gpu_impl.hxx:44-51)
	Sampler tu0, tu1;
	DepthUnit tu3; };

inline void normalize3(float& x, float& y, float& z) {
	const float len2 = (x*x + y*y) + z*z;
	const float inv = 1.0F / std::sqrt(len2);
	x = x * inv; y = y * inv; z = z * inv; }

inline float interp(const float bp[3], float a, float b, float c) {
	return (bp[0]*a + bp[1]*b) + bp[2]*c; }

struct Program {
	int id = 0;
	bool earlyZ = true;
	int nvary = 0;
	bool instanced = false;
	virtual ~Program() = default;
	virtual void ShadeVertex(const DrawEnv& e, int iid, int idx, float pos[4], VOut& o) const = 0;
	// one quad; vary[k][lane], colour[c][lane]; mask lanes may be cleared
	virtual void ShadeFragment(const DrawEnv& e, const float bs[3][4], const float bp[3][4], const float vary[kMaxVary][4],
	                           const float fragCoord[2][4], const float fragDepth[4], float color[4][4], int32_t mask[4]) const = 0;
};

inline float buf_at(const float* p, int idx, float dflt) { return p ? p[idx] : dflt; }

inline void load_pos(const DrawEnv& e, int idx, float p[4]) {
	const auto& b = e.st->buffers;
	p[0] = buf_at(b[0], idx, 0.0F); p[1] = buf_at(b[1], idx, 0.0F); p[2] = buf_at(b[2], idx, 0.0F); p[3] = 1.0F; }

// AmyProgram (viewer/shaders.hxx:69-161): uv varying, texture colour
struct AmyProgram final : Program {
	AmyProgram() { id = 4; nvary = 2; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		xform4(e.vpm.ff, p, pos);
		o.f[0] = buf_at(e.st->buffers[9], idx, 0.0F);
		o.f[1] = buf_at(e.st->buffers[10], idx, 0.0F); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		e.tu0.sample(vary[0], vary[1], color); }};

// ManyProgram (viewer/shaders.hxx:482-583): instanced, colour = (u, v, magic, 1)
struct ManyProgram final : Program {
	ManyProgram() { id = 6; nvary = 2; instanced = true; }
	void ShadeVertex(const DrawEnv& e, int iid, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		const float* imat = e.st->buffers[15] + 16 * iid;
		float p1[4]; xform4(imat, p, p1);
		xform4(e.vpm.ff, p1, pos);
		o.f[0] = buf_at(e.st->buffers[9], idx, 0.0F);
		o.f[1] = buf_at(e.st->buffers[10], idx, 0.0F); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		const float magic = e.uniforms ? e.uniforms[0] : 0.0F;
		for (int l = 0; l < 4; ++l) { color[0][l] = vary[0][l]; color[1][l] = vary[1][l]; color[2][l] = magic; color[3][l] = 1.0F; }}};

// LambertTexProgram (new): textured, one directional light.
// uniforms: [0..2] unit direction toward the light, [4] ambient, [5] diffuse
struct LambertTexProgram final : Program {
	LambertTexProgram() { id = 80; nvary = 5; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		xform4(e.vpm.ff, p, pos);
		const auto& b = e.st->buffers;
		o.f[0] = buf_at(b[3], idx, 0.0F); o.f[1] = buf_at(b[4], idx, 0.0F); o.f[2] = buf_at(b[5], idx, 1.0F);
		o.f[3] = buf_at(b[9], idx, 0.0F); o.f[4] = buf_at(b[10], idx, 0.0F); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		const float* u = e.uniforms;
		float tex[4][4];
		e.tu0.sample(vary[3], vary[4], tex);
		for (int l = 0; l < 4; ++l) {
			float nx = vary[0][l], ny = vary[1][l], nz = vary[2][l];
			normalize3(nx, ny, nz);
			float ndl = (nx*u[0] + ny*u[1]) + nz*u[2];
			ndl = sse_max(ndl, 0.0F);
			const float k = u[4] + u[5] * ndl;
			color[0][l] = tex[0][l] * k; color[1][l] = tex[1][l] * k; color[2][l] = tex[2][l] * k;
			color[3][l] = 1.0F; }}};

// ManyBlinnPhongProgram (new): instanced, textured, one point light, Blinn-Phong.
// uniforms: [0..2] eye, [4..6] light position, [8] ka, [9] kd, [10] ks
struct ManyBlinnPhongProgram final : Program {
	ManyBlinnPhongProgram() { id = 81; nvary = 8; instanced = true; }
	void ShadeVertex(const DrawEnv& e, int iid, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		const auto& b = e.st->buffers;
		const float* imat = b[15] + 16 * iid;
		float p1[4]; xform4(imat, p, p1);
		xform4(e.vpm.ff, p1, pos);
		const float n[3] = { buf_at(b[3], idx, 0.0F), buf_at(b[4], idx, 0.0F), buf_at(b[5], idx, 1.0F) };
		float wn[3]; xform3w0(imat, n, wn);
		o.f[0] = p1[0]; o.f[1] = p1[1]; o.f[2] = p1[2];
		o.f[3] = wn[0]; o.f[4] = wn[1]; o.f[5] = wn[2];
		o.f[6] = buf_at(b[9], idx, 0.0F); o.f[7] = buf_at(b[10], idx, 0.0F); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		const float* u = e.uniforms;
		float tex[4][4];
		e.tu0.sample(vary[6], vary[7], tex);
		for (int l = 0; l < 4; ++l) {
			const float px = vary[0][l], py = vary[1][l], pz = vary[2][l];
			float nx = vary[3][l], ny = vary[4][l], nz = vary[5][l];
			normalize3(nx, ny, nz);
			float vx = u[0] - px, vy = u[1] - py, vz = u[2] - pz;
			normalize3(vx, vy, vz);
			float lx = u[4] - px, ly = u[5] - py, lz = u[6] - pz;
			normalize3(lx, ly, lz);
			float hx = lx + vx, hy = ly + vy, hz = lz + vz;
			normalize3(hx, hy, hz);
			const float ndl = sse_max((nx*lx + ny*ly) + nz*lz, 0.0F);
			const float ndh = sse_max((nx*hx + ny*hy) + nz*hz, 0.0F);
			float s = ndh * ndh; s = s * s; s = s * s; s = s * s; s = s * s;  // ndh^32
			const float kdiff = u[8] + u[9] * ndl;
			const float kspec = u[10] * s;
			color[0][l] = tex[0][l] * kdiff + kspec;
			color[1][l] = tex[1][l] * kdiff + kspec;
			color[2][l] = tex[2][l] * kdiff + kspec;
			color[3][l] = 1.0F; }}};

// GGXEnvProgram (new, the heavy shader): ManyBlinnPhong's vertex stage
// plus per-instance roughness; 16 point lights with Cook-Torrance GGX
// (Smith-Schlick G, Schlick F, F0 = 0.04) and an 8-tap environment cube
// lookup around the reflection vector.  Only + - * / sqrt in a fixed order.
//   buffers: 11 environment cube (6 faces of 16x16 RGBA floats), 12 taps
//            (8 x {dx, dy, dz, weight}), 13 lights (16 x {x, y, z, 0, r, g,
//            b, 0}), 14 per-instance roughness, 15 instance matrices
//   uniforms: [0..2] eye, [3] ambient, [4] env specular, [5] env diffuse
constexpr int kGGXLights = 16, kEnvTaps = 8, kEnvDim = 16;
constexpr float kPi = 3.14159265F, kInvPi = 0.318309886F;

struct ShadeInput { float n[3], v[3], p[3], albedo[3], rough; };

void env_sample(const float* env, float dx, float dy, float dz, float out[3]) {
	out[0] = 0.0F; out[1] = 0.0F; out[2] = 0.0F;
	if (env == nullptr) return;
	const float ax = dx < 0.0F ? -dx : dx, ay = dy < 0.0F ? -dy : dy, az = dz < 0.0F ? -dz : dz;
	int face; float ma, sc, tc;
	if (ax >= ay && ax >= az) { face = dx > 0.0F ? 0 : 1; ma = ax; sc = dx > 0.0F ? -dz : dz; tc = -dy; }
	else if (ay >= az) { face = dy > 0.0F ? 2 : 3; ma = ay; sc = dx; tc = dy > 0.0F ? dz : -dz; }
	else { face = dz > 0.0F ? 4 : 5; ma = az; sc = dz > 0.0F ? dx : -dx; tc = -dy; }
	if (!(ma > 0.0F)) return;
	const float E = static_cast<float>(kEnvDim);
	float fs = ((sc / ma + 1.0F) * 0.5F) * E - 0.5F;
	float ft = ((tc / ma + 1.0F) * 0.5F) * E - 0.5F;
	fs = fs >= 0.0F ? fs : 0.0F; fs = fs <= E - 1.0F ? fs : E - 1.0F;
	ft = ft >= 0.0F ? ft : 0.0F; ft = ft <= E - 1.0F ? ft : E - 1.0F;
	const int s0 = static_cast<int>(fs), t0 = static_cast<int>(ft);
	const int s1 = std::min(s0 + 1, kEnvDim - 1), t1 = std::min(t0 + 1, kEnvDim - 1);
	const float fx = fs - static_cast<float>(s0), fy = ft - static_cast<float>(t0);
	const float fx1 = 1.0F - fx, fy1 = 1.0F - fy;
	const float w00 = fx1 * fy1, w10 = fx * fy1, w01 = fx1 * fy, w11 = fx * fy;
	const float* f0 = env + face * kEnvDim * kEnvDim * 4;
	const float* p00 = f0 + (t0 * kEnvDim + s0) * 4;
	const float* p10 = f0 + (t0 * kEnvDim + s1) * 4;
	const float* p01 = f0 + (t1 * kEnvDim + s0) * 4;
	const float* p11 = f0 + (t1 * kEnvDim + s1) * 4;
	for (int c = 0; c < 3; ++c) out[c] = ((p00[c] * w00 + p10[c] * w10) + p01[c] * w01) + p11[c] * w11; }

void ggx_shade(const DrawEnv& e, const ShadeInput& s, float color[3]) {
	const float* u = e.uniforms;
	const auto& b = e.st->buffers;
	const float* lights = b[13];
	const float* taps = b[12];
	const float rough = s.rough;
	const float a = rough * rough, a2 = a * a;
	const float k = ((rough + 1.0F) * (rough + 1.0F)) * 0.125F;
	const float ndvRaw = (s.n[0]*s.v[0] + s.n[1]*s.v[1]) + s.n[2]*s.v[2];
	const float ndv = ndvRaw > 1e-4F ? ndvRaw : 1e-4F;
	const float gv = ndv / (ndv * (1.0F - k) + k);
	float acc[3] = { 0.0F, 0.0F, 0.0F };
	for (int i = 0; lights != nullptr && i < kGGXLights; ++i) {
		const float* L = lights + 8 * i;
		float lx = L[0] - s.p[0], ly = L[1] - s.p[1], lz = L[2] - s.p[2];
		const float d2 = (lx*lx + ly*ly) + lz*lz;
		const float inv = 1.0F / std::sqrt(d2);
		lx = lx * inv; ly = ly * inv; lz = lz * inv;
		const float ndl = (s.n[0]*lx + s.n[1]*ly) + s.n[2]*lz;
		if (!(ndl > 0.0F)) continue;
		float hx = lx + s.v[0], hy = ly + s.v[1], hz = lz + s.v[2];
		normalize3(hx, hy, hz);
		float ndh = (s.n[0]*hx + s.n[1]*hy) + s.n[2]*hz;
		ndh = ndh > 0.0F ? ndh : 0.0F;
		float vdh = (s.v[0]*hx + s.v[1]*hy) + s.v[2]*hz;
		vdh = vdh > 0.0F ? vdh : 0.0F;
		const float t = (ndh * ndh) * (a2 - 1.0F) + 1.0F;
		const float D = a2 / ((kPi * t) * t);
		const float gl = ndl / (ndl * (1.0F - k) + k);
		const float f = 1.0F - vdh, f2 = f * f, f5 = (f2 * f2) * f;
		const float F = 0.04F + 0.96F * f5;
		const float spec = (((D * gl) * gv) * F) / ((4.0F * ndl) * ndv);
		const float kd = (1.0F - F) * kInvPi;
		const float id2 = 1.0F / d2;
		for (int c = 0; c < 3; ++c) acc[c] = acc[c] + ((s.albedo[c] * kd + spec) * (L[4 + c] * id2)) * ndl; }
	const float rx = (2.0F * ndvRaw) * s.n[0] - s.v[0];
	const float ry = (2.0F * ndvRaw) * s.n[1] - s.v[1];
	const float rz = (2.0F * ndvRaw) * s.n[2] - s.v[2];
	float ev[3] = { 0.0F, 0.0F, 0.0F };
	for (int j = 0; taps != nullptr && j < kEnvTaps; ++j) {
		const float* T = taps + 4 * j;
		float smp[3];
		env_sample(b[11], rx + T[0] * rough, ry + T[1] * rough, rz + T[2] * rough, smp);
		for (int c = 0; c < 3; ++c) ev[c] = ev[c] + T[3] * smp[c]; }
	const float fe = 1.0F - ndv, fe2 = fe * fe, fe5 = (fe2 * fe2) * fe;
	const float Fe = 0.04F + 0.96F * fe5;
	const float ks = (u[4] * Fe) * (1.0F - rough);
	for (int c = 0; c < 3; ++c) color[c] = (acc[c] + u[3] * s.albedo[c]) + ev[c] * (ks + u[5] * s.albedo[c]); }

struct GGXEnvProgram final : Program {
	GGXEnvProgram() { id = 82; nvary = 9; instanced = true; }
	void ShadeVertex(const DrawEnv& e, int iid, int idx, float pos[4], VOut& o) const override {
		ManyBlinnPhongProgram().ShadeVertex(e, iid, idx, pos, o);
		const float* r = e.st->buffers[14];
		o.f[8] = r ? r[iid] : 0.5F; }
	static void Inputs(const DrawEnv& e, const float vary[kMaxVary][4], const float tex[4][4], int l, ShadeInput& s) {
		const float* u = e.uniforms;
		s.albedo[0] = tex[0][l]; s.albedo[1] = tex[1][l]; s.albedo[2] = tex[2][l];
		s.p[0] = vary[0][l]; s.p[1] = vary[1][l]; s.p[2] = vary[2][l];
		float nx = vary[3][l], ny = vary[4][l], nz = vary[5][l];
		normalize3(nx, ny, nz);
		s.n[0] = nx; s.n[1] = ny; s.n[2] = nz;
		float vx = u[0] - vary[0][l], vy = u[1] - vary[1][l], vz = u[2] - vary[2][l];
		normalize3(vx, vy, vz);
		s.v[0] = vx; s.v[1] = vy; s.v[2] = vz;
		s.rough = vary[8][l]; }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		float tex[4][4];
		e.tu0.sample(vary[6], vary[7], tex);
		for (int l = 0; l < 4; ++l) {
			ShadeInput s;
			Inputs(e, vary, tex, l, s);
			float c[3];
			ggx_shade(e, s, c);
			color[0][l] = c[0]; color[1][l] = c[1]; color[2][l] = c[2]; color[3][l] = 1.0F; }}};

// rglv::BaseProgram (rglv_gpu_shaders.hxx:21-95): position only, colour 1
struct BaseProgram final : Program {
	BaseProgram() { id = 0; nvary = 0; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut&) const override {
		float p[4]; load_pos(e, idx, p);
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv&, const float[3][4], const float[3][4], const float[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		for (int c = 0; c < 4; ++c) for (int l = 0; l < 4; ++l) color[c][l] = 1.0F; }};

// ---- the viewer's other programs (viewer/shaders.hxx) ----
inline float dot4(const float a[4], const float b[4]) { return ((a[0]*b[0] + a[1]*b[1]) + a[2]*b[2]) + a[3]*b[3]; }

// AlphaTextureProgram (viewer/shaders.hxx:164-230): late depth test,
// gl_triMask &= alpha > 0
struct AlphaTextureProgram final : Program {
	AlphaTextureProgram() { id = 65; nvary = 2; earlyZ = false; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		o.f[0] = buf_at(e.st->buffers[9], idx, 0.0F);
		o.f[1] = buf_at(e.st->buffers[10], idx, 0.0F);
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t mask[4]) const override {
		e.tu0.sample(vary[0], vary[1], color);
		for (int l = 0; l < 4; ++l) mask[l] &= (color[3][l] > 0.0F) ? -1 : 0; }};

// TextProgram (viewer/shaders.hxx:232-318): texture times vertex colour
// (slots 6-8, default white), late depth test, alpha discard
struct TextProgram final : Program {
	TextProgram() { id = 26; nvary = 5; earlyZ = false; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		const auto& b = e.st->buffers;
		if (b[6]) { o.f[0] = b[6][idx]; o.f[1] = b[7][idx]; o.f[2] = b[8][idx]; }
		else { o.f[0] = o.f[1] = o.f[2] = 1.0F; }
		o.f[3] = buf_at(b[9], idx, 0.0F);
		o.f[4] = buf_at(b[10], idx, 0.0F);
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t mask[4]) const override {
		e.tu0.sample(vary[3], vary[4], color);
		for (int l = 0; l < 4; ++l) {
			color[0][l] *= vary[0][l]; color[1][l] *= vary[1][l]; color[2][l] *= vary[2][l];
			mask[l] &= (color[3][l] > 0.0F) ? -1 : 0; }}};

// DepthProgram (viewer/shaders.hxx:321-420): tu3 at the uv, linearised
struct DepthProgram final : Program {
	DepthProgram() { id = 5; nvary = 2; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		o.f[0] = buf_at(e.st->buffers[9], idx, 0.0F);
		o.f[1] = buf_at(e.st->buffers[10], idx, 0.0F);
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		const float zNear = 10.0F, zFar = 1000.0F;
		float tmp[4];
		e.tu3.sample(vary[0], vary[1], tmp);
		for (int l = 0; l < 4; ++l) {
			const float t = (2.0F * zNear) / ((zFar + zNear) - tmp[l] * (zFar - zNear));
			color[0][l] = t; color[1][l] = t; color[2][l] = t; color[3][l] = 1.0F; }}};

// PatternProgram (viewer/shaders.hxx:422-480): tu0 at gl_FragCoord / dim.y
// + offset.xy; uniforms offset (vec4), dim (vec4)
struct PatternProgram final : Program {
	PatternProgram() { id = 41; nvary = 1; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		o.f[0] = 0.0F;
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float[kMaxVary][4], const float fc[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		const float* u = e.uniforms;
		float uu[4], vv[4];
		for (int l = 0; l < 4; ++l) { uu[l] = fc[0][l] / u[5]; vv[l] = fc[1][l] / u[5]; uu[l] = uu[l] + u[0]; vv[l] = vv[l] + u[1]; }
		e.tu0.sample(uu, vv, color); }};

inline void obj_normal(const DrawEnv& e, int idx, float n[3]) {
	const auto& b = e.st->buffers;
	if (b[3]) { n[0] = b[3][idx]; n[1] = b[4][idx]; n[2] = b[5][idx]; } else { n[0] = 0.0F; n[1] = 0.0F; n[2] = 1.0F; }}
inline void obj_kd(const DrawEnv& e, int idx, float kd[3]) {
	const auto& b = e.st->buffers;
	if (b[6]) { kd[0] = b[6][idx]; kd[1] = b[7][idx]; kd[2] = b[8][idx]; } else { kd[0] = kd[1] = kd[2] = 1.0F; }}

// the eye-light diffuse of OBJ1 (per vertex) and OBJ2 (per fragment): the
// qfloat4 difference lightPos - p has w = lightPos.w - p.z
// (rmlv_soa.hxx:163), length uses sqrt, normalize rsqrtps
inline float obj_diffuse(const float p[4], const float n[4]) {
	const float lightPos[4] = { 0.0F, 0.0F, 0.0F, 1.0F };
	float d[4] = { lightPos[0] - p[0], lightPos[1] - p[1], lightPos[2] - p[2], lightPos[3] - p[2] };
	const float distance = std::sqrt(dot4(d, d));
	const float scale = rsqrtps(dot4(d, d));
	const float lv[4] = { d[0] * scale, d[1] * scale, d[2] * scale, d[3] * scale };
	float diffuse = sse_max(dot4(n, lv), 0.1F);
	diffuse = diffuse * (100.0F / distance);
	return diffuse; }

// OBJ1Program (viewer/shaders.hxx:584-681)
struct OBJ1Program final : Program {
	OBJ1Program() { id = 7; nvary = 3; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		float n[3], kd[3]; obj_normal(e, idx, n); obj_kd(e, idx, kd);
		float mvv[4], mvn[4];
		xform4(e.vm.ff, p, mvv);
		const float n4[4] = { n[0], n[1], n[2], 0.0F };
		xform4(e.vm.ff, n4, mvn);
		const float diffuse = obj_diffuse(mvv, mvn);
		o.f[0] = kd[0] * diffuse; o.f[1] = kd[1] * diffuse; o.f[2] = kd[2] * diffuse;
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv&, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		for (int l = 0; l < 4; ++l) { color[0][l] = vary[0][l]; color[1][l] = vary[1][l]; color[2][l] = vary[2][l]; color[3][l] = 1.0F; }}};

// OBJ2Program (viewer/shaders.hxx:701-818): varyings sp (4), sn (4), kd (3)
struct OBJ2Program final : Program {
	OBJ2Program() { id = 8; nvary = 11; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		float n[3]; obj_normal(e, idx, n);
		xform4(e.vm.ff, p, o.f);
		const float n4[4] = { n[0], n[1], n[2], 0.0F };
		xform4(e.vm.ff, n4, o.f + 4);
		obj_kd(e, idx, o.f + 8);
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv&, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		for (int l = 0; l < 4; ++l) {
			const float sp[4] = { vary[0][l], vary[1][l], vary[2][l], vary[3][l] };
			const float sn[4] = { vary[4][l], vary[5][l], vary[6][l], vary[7][l] };
			const float diffuse = obj_diffuse(sp, sn);
			color[0][l] = vary[8][l] * diffuse; color[1][l] = vary[9][l] * diffuse; color[2][l] = vary[10][l] * diffuse;
			color[3][l] = 1.0F; }}};

// OBJ2SProgram (viewer/shaders.hxx:821-978): varyings sp (4), sn (4), kd
// (3), lp (4); uniforms modelToShadow (mat4), lpos, ldir (vec3), lcos
struct OBJ2SProgram final : Program {
	OBJ2SProgram() { id = 9; nvary = 15; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		float n[3]; obj_normal(e, idx, n);
		xform4(e.vm.ff, p, o.f);
		const float n4[4] = { n[0], n[1], n[2], 1.0F };
		xform4(e.nm.ff, n4, o.f + 4);
		obj_kd(e, idx, o.f + 8);
		xform4(e.uniforms, p, o.f + 11);
		xform4(e.vpm.ff, p, pos); }
	static void normalize3r(float v[3]) {
		const float s = rsqrtps((v[0]*v[0] + v[1]*v[1]) + v[2]*v[2]);
		v[0] = v[0] * s; v[1] = v[1] * s; v[2] = v[2] * s; }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		const float* u = e.uniforms;
		float att[4], diff[4], lx[4], ly[4], lz[4];
		for (int l = 0; l < 4; ++l) {
			float stl[3] = { u[16] - vary[0][l], u[17] - vary[1][l], u[18] - vary[2][l] };
			const float dist = std::sqrt((stl[0]*stl[0] + stl[1]*stl[1]) + stl[2]*stl[2]);
			normalize3r(stl);
			att[l] = 100.0F / dist;
			float ld[3] = { u[19], u[20], u[21] };
			normalize3r(ld);
			const float ang = ((-stl[0])*ld[0] + (-stl[1])*ld[1]) + (-stl[2])*ld[2];
			if (ang < u[22]) att[l] = 0.111F;
			float sn[3] = { vary[4][l], vary[5][l], vary[6][l] };
			normalize3r(sn);
			diff[l] = sse_max(0.0F, (sn[0]*stl[0] + sn[1]*stl[1]) + sn[2]*stl[2]);
			lx[l] = vary[11][l] / vary[14][l]; ly[l] = vary[12][l] / vary[14][l]; lz[l] = vary[13][l] / vary[14][l]; }
		float sx[4], sy[4], depth[4];
		for (int l = 0; l < 4; ++l) { sx[l] = lx[l] * 0.5F + 0.5F; sy[l] = ly[l] * 0.5F + 0.5F; }
		e.tu3.sample(sx, sy, depth);
		for (int l = 0; l < 4; ++l) {
			if (lz[l] - 0.0005F > depth[l]) att[l] = 0.111F;
			color[0][l] = (att[l] * vary[8][l]) * diff[l];
			color[1][l] = (att[l] * vary[9][l]) * diff[l];
			color[2][l] = (att[l] * vary[10][l]) * diff[l];
			color[3][l] = 1.0F; }}};

// EnvmapProgram (viewer/shaders_envmap.hxx:26-127): uv of the sphere map
// at reflect(e, n), e the unit view-space position, n the unit normal
// matrix image of mix(smooth, face, 0)
struct EnvmapProgram final : Program {
	EnvmapProgram() { id = 10; nvary = 2; }
	static void normalize3r(float v[3]) {
		const float sc = rsqrtps((v[0]*v[0] + v[1]*v[1]) + v[2]*v[2]);
		v[0] = v[0] * sc; v[1] = v[1] * sc; v[2] = v[2] * sc; }
	void ShadeVertex(const DrawEnv& e, int, int idx, float pos[4], VOut& o) const override {
		float p[4]; load_pos(e, idx, p);
		float smooth[3]; obj_normal(e, idx, smooth);
		const auto& b = e.st->buffers;
		float face[3] = { 1.0F, 1.0F, 1.0F };
		if (b[6]) { face[0] = b[6][idx]; face[1] = b[7][idx]; face[2] = b[8][idx]; }
		float position[4];
		xform4(e.vm.ff, p, position);
		float ev[3] = { position[0], position[1], position[2] };
		normalize3r(ev);
		float vn[3];
		for (int k = 0; k < 3; ++k) { const float diff = face[k] - smooth[k]; vn[k] = smooth[k] + 0.0F * diff; }
		const float* nm = e.nm.ff;
		float n[3] = { (nm[0]*vn[0] + nm[4]*vn[1]) + nm[8]*vn[2],
		               (nm[1]*vn[0] + nm[5]*vn[1]) + nm[9]*vn[2],
		               (nm[2]*vn[0] + nm[6]*vn[1]) + nm[10]*vn[2] };
		normalize3r(n);
		const float d = (n[0]*ev[0] + n[1]*ev[1]) + n[2]*ev[2];
		const float twod = 2.0F * d;
		float r[3];
		for (int k = 0; k < 3; ++k) r[k] = ev[k] - twod * n[k];
		const float zp = r[2] + 1.0F;
		const float m = 2.0F * std::sqrt(((r[0]*r[0]) + (r[1]*r[1])) + (zp*zp));
		o.f[0] = r[0] / m + 0.5F;
		o.f[1] = r[1] / m + 0.5F;
		xform4(e.vpm.ff, p, pos); }
	void ShadeFragment(const DrawEnv& e, const float[3][4], const float[3][4], const float vary[kMaxVary][4], const float[2][4], const float[4],
	                   float color[4][4], int32_t[4]) const override {
		e.tu0.sample(vary[0], vary[1], color);
		for (int l = 0; l < 4; ++l) color[3][l] = 0.5F; }};

const Program* find_program(int id) {
	static const AmyProgram amy; static const ManyProgram many;
	static const LambertTexProgram lam; static const ManyBlinnPhongProgram mbp;
	static const GGXEnvProgram ggx;
	static const AlphaTextureProgram alpha; static const TextProgram text; static const DepthProgram depth;
	static const PatternProgram pattern; static const OBJ1Program obj1; static const OBJ2Program obj2; static const OBJ2SProgram obj2s;
	static const EnvmapProgram envmap; static const BaseProgram base;
	switch (id) {
	case 0: return &base; case 4: return &amy; case 6: return &many; case 80: return &lam; case 81: return &mbp; case 82: return &ggx;
	case 65: return &alpha; case 26: return &text; case 5: return &depth; case 41: return &pattern;
	case 7: return &obj1; case 8: return &obj2; case 9: return &obj2s; case 10: return &envmap;
	default: return nullptr; }}

// post programs for StoreTrueColor (ShadeCanvas); 1 = Default, 2 = Exposure,
// 3 = IQ (viewer/shaders.hxx:32-68)
bool known_post_program(int id) { return id == 1 || id == 2 || id == 3; }

// _mm_cvtps_epi32 under the default rounding mode (nearest, ties to even)
inline int32_t cvtn(float f) {
	if (f >= -2147483648.0F && f < 2147483648.0F) return static_cast<int32_t>(std::nearbyint(f));
	return INT32_MIN; }

// J. R. Fonseca's SSE2 exp2 / log2 (3rdparty/sse-pow/sse_pow.h:19-87, as
// the reference configures it: EXP_POLY_DEGREE 3, LOG_POLY_DEGREE 5),
// one lane; the polynomial coefficients are that method's minimax fits
inline float sse_exp2(float x) {
	x = sse_min(x, 129.00000F);
	x = sse_max(x, -126.99999F);
	const int32_t ipart = cvtn(x - 0.5F);
	const float fpart = x - itof(ipart);
	const float expipart = u2f(static_cast<uint32_t>(shl(wadd(ipart, 127), 23)));
	float p = 7.8024521e-2F;
	p = p * fpart + 2.2606716e-1F;
	p = p * fpart + 6.9583356e-1F;
	p = p * fpart + 9.9992520e-1F;
	return expipart * p; }

inline float sse_log2(float x) {
	const uint32_t i = f2u(x);
	const float e = itof(static_cast<int32_t>((i & 0x7F800000u) >> 23) - 127);
	const float m = u2f((i & 0x007FFFFFu) | f2u(1.0F));
	float p = static_cast<float>(0.0596515482674574969533);
	p = p * m + static_cast<float>(-0.465725644288844778798);
	p = p * m + static_cast<float>(1.48116647521213171641);
	p = p * m + static_cast<float>(-2.52074962577807006663);
	p = p * m + static_cast<float>(2.8882704548164776201);
	p = p * (m - 1.0F);
	return p + e; }

// rmlv pow (rmlv_mvec4.hxx:652-654)
inline float sse_pow(float x, float y) { return sse_exp2(sse_log2(x) * y); }

// rmlv sin (rmlv_mvec4.hxx:700-760): scale by 1/pi, wrap to [-1, 1] with
// the sign flipped for odd periods, then the parabola y (3.1 + 3.6 |y|)
inline float approx_sin(float x) {
	float M = x * static_cast<float>(1.0 / 3.14159265358979323846);
	const int32_t whole = cvtt(M);
	M = M - itof(whole);
	M = u2f(f2u(M) ^ static_cast<uint32_t>(shl(whole, 31)));
	const float y = M - (M * std::fabs(M));
	return y * (3.1F + 3.6F * std::fabs(y)); }

// rmlv fract (rmlv_mvec4.hxx:576-580)
inline float fract_t(float a) { return a - itof(cvtt(a)); }

// IQPostProgram::ShadeCanvas for one pixel at q (viewer/shaders.hxx:56-67;
// rglv_fragment.hxx:34-36 hash3)
inline void iq_post(float c[3], float qx, float qy) {
	c[0] = sse_pow(c[0], 0.45F); c[1] = sse_pow(c[1], 0.5F); c[2] = sse_pow(c[2], 0.55F);
	float t = 16.0F * qx;
	t = t * qy;
	t = t * (1.0F - qx);
	t = t * (1.0F - qy);
	const float v = 0.2F + 0.8F * sse_pow(t, 0.2F);
	const float n = qx + 13.0F * qy;
	const float k = static_cast<float>(1.0 / 255.0);
	for (int ch = 0; ch < 3; ++ch) {
		c[ch] = c[ch] * v;
		const float h = fract_t(approx_sin(n + static_cast<float>(ch)) * 43758.5453123F);
		c[ch] = c[ch] + k * h; }}

// ------------------------------------------------------------------------
// per-tile command lists
// ------------------------------------------------------------------------
struct TileOp {
	enum Op : uint8_t { STATE, CLEAR, TRI, CLIPPED_TRI, STORE_TC, STORE_DEPTH, STORE_PRIM, STORE_FP } op;
	int32_t a;      // state index | clear bits | iid | gamma
	uint16_t i0, i1, i2;
	uint32_t prim;
	const void* ptr; };

struct ClippedVertex { float coord[4]; VOut data; };

struct Rect { int x0, y0, x1, y1; };

struct Frustum {  // ViewFrustum (rglv_view_frustum.hxx:36-73)
	float factor;
	// fixedBand: the factor is raised to at least 1 (a guard band no smaller
	// than the viewport); the reference's formula falls below 1 above 2048
	// pixels and then culls on-screen triangles
	explicit Frustum(int w, int h, bool fixedBand = false) {
		const int longSide = std::max(w, h);
		const int half = longSide / 2;
		factor = (2048.0F - static_cast<float>(half)) / static_cast<float>(half);
		if (fixedBand && factor < 1.0F) factor = 1.0F; }
	uint32_t Test(const float p[4]) const {  // SIMD version, cmple
		const float w = p[3] * factor;
		uint32_t f = 0;
		if (w + p[0] <= 0.0F) f |= 1u;
		if (w + p[1] <= 0.0F) f |= 2u;
		if (p[3] + p[2] <= 0.0F) f |= 4u;
		if (w - p[0] <= 0.0F) f |= 8u;
		if (w - p[1] <= 0.0F) f |= 16u;
		return f; }};

// clipping planes in reference order: Left, Bottom, Near, Right, Top (:245-251)
inline float plane_dist(int plane, const float p[4]) {
	switch (plane) {
	case 0: return p[3] + p[0];
	case 1: return p[3] + p[1];
	case 2: return p[3] + p[2];
	case 3: return p[3] - p[0];
	default: return p[3] - p[1]; }}

// Sutherland-Hodgman over L, B, N, R, T against the unscaled planes, each
// new vertex interpolated from the inside vertex toward the outside one
// (ClipTriangles phase 2, rglv_gpu_impl.hxx:727-762; IsInside and Distance,
// rglv_view_frustum.hxx:75-94; mix, rmlv_math.hxx:87-90).  A is replaced by
// the clipped polygon (empty when nothing is left).
inline void mix_vout(const VOut& a, const VOut& b, float t, VOut& o, int nvary) {
	for (int k = 0; k < nvary; ++k) o.f[k] = (1.0F - t) * a.f[k] + t * b.f[k]; }

void clip_polygon(std::vector<ClippedVertex>& A, int nvary) {
	std::vector<ClippedVertex> B;
	for (int plane = 0; plane < 5; ++plane) {
		bool hereIsInside = plane_dist(plane, A[0].coord) >= 0.0F;
		B.clear();
		for (size_t h = 0; h < A.size(); ++h) {
			const size_t n = (h + 1) % A.size();
			const bool nextIsInside = plane_dist(plane, A[n].coord) >= 0.0F;
			if (hereIsInside) B.push_back(A[h]);
			if (hereIsInside != nextIsInside) {
				const ClippedVertex& from = hereIsInside ? A[h] : A[n];
				const ClippedVertex& to = hereIsInside ? A[n] : A[h];
				const float fa = plane_dist(plane, from.coord), fb = plane_dist(plane, to.coord);
				const float d = fa / (fa - fb);
				ClippedVertex nv{};
				for (int q = 0; q < 4; ++q) nv.coord[q] = (1.0F - d) * from.coord[q] + d * to.coord[q];
				mix_vout(from.data, to.data, d, nv.data, nvary);
				B.push_back(nv);
				hereIsInside = !hereIsInside; }}
		std::swap(A, B);
		if (A.empty()) break; }}

}  // namespace

// ------------------------------------------------------------------------
// the oracle GPU
// ------------------------------------------------------------------------
struct rglo_gpu {
	// recorder (rglv::GL)
	GLState cs{};
	bool dirty = true;
	std::vector<float> ubuf;
	std::deque<GLState> states;
	std::vector<Cmd> commands;

	// render target
	int W = 0, H = 0, tileW = 64, tileH = 64, tilesX = 0, tilesY = 0;
	bool fixedGuardBand = false;
	std::vector<std::vector<TileOp>> tiles;
	std::vector<ClippedVertex> clipped;
	std::vector<uint32_t> installed;
	std::string err;
	int threads = 1;
	// statistics
	long long statSubmitted = 0, statClipped = 0, statBinned = 0, statFragments = 0, statClippedPairs = 0;

	rglo_gpu() { cs.reset(); }

	void MaybeUpdateState() {  // GL::MaybeUpdateState (rglv_gl.cxx:100-106)
		if (!dirty) return;
		states.push_back(cs);
		commands.push_back({ Cmd::STATE, static_cast<int>(states.size()) - 1, 0, 0, nullptr });
		dirty = false; }

	const float* Uniforms(const GLState& s) const {
		return s.uniformsOfs < 0 ? nullptr : ubuf.data() + s.uniformsOfs; }

	Rect TileRect(int t) const {  // GPU::TileRect (rglv_gpu.hxx:291-299)
		const int ty = t / tilesX, tx = t % tilesX;
		return { tx * tileW, ty * tileH, std::min((tx + 1) * tileW, W), std::min((ty + 1) * tileH, H) }; }

	// GPU::ScissorRect (rglv_gpu.hxx:258-276), clamped to the target: a rect
	// reaching past it makes the reference bin into tiles outside its grid
	Rect ScissorRect(const GLState& s) const {
		if (!s.scissorEnabled) return { 0, 0, W, H };
		const int left = s.scissorOrigin[0], right = left + s.scissorSize[0];
		const int bottom = H - s.scissorOrigin[1] - 1, top = bottom - s.scissorSize[1];
		auto cl = [](int v, int hi) { return std::min(std::max(v, 0), hi); };
		return { cl(left, W), cl(top, H), cl(right, W), cl(bottom, H) }; }

	// the rect a tile rasterises a draw in: the tile rect, intersected with
	// the scissor rect when the scissor test is on (rglv_gpu_impl.hxx:848-849)
	Rect RasterRect(const Rect& tr, const GLState& s) const {
		if (!s.scissorEnabled) return tr;
		const Rect sr = ScissorRect(s);
		return { std::max(tr.x0, sr.x0), std::max(tr.y0, sr.y0), std::min(tr.x1, sr.x1), std::min(tr.y1, sr.y1) }; }

	void DSDO(const GLState& s, float DS[2], float DO[2]) const {  // (rglv_gpu.hxx:265-270)
		const int vx = s.hasViewportSize ? s.viewportSize[0] : W;
		const int vy = s.hasViewportSize ? s.viewportSize[1] : H;
		DS[0] = static_cast<float>(vx / 2); DS[1] = static_cast<float>(-vy / 2);
		DO[0] = static_cast<float>(vx / 2 + s.viewportOrigin[0]);
		DO[1] = static_cast<float>(H - (vy / 2 + s.viewportOrigin[1])); }

	bool KeyInstalled(uint32_t key) const {
		return std::find(installed.begin(), installed.end(), key) != installed.end(); }

	static uint32_t FragmentStateKey(const GLState& s) {  // (rglv_gl.hxx:136-146)
		uint32_t key = static_cast<uint32_t>(s.programId) << 24;
		key |= static_cast<uint32_t>(s.scissorEnabled);
		key |= static_cast<uint32_t>(s.depthTestEnabled) << 1;
		if (s.depthTestEnabled) key |= static_cast<uint32_t>(s.depthFunc) << 2;
		key |= static_cast<uint32_t>(s.blendingEnabled) << 4;
		key |= static_cast<uint32_t>(s.depthWriteMask) << 5;
		key |= static_cast<uint32_t>(s.colorWriteMask) << 6;
		key |= static_cast<uint32_t>(s.color0AttachmentType) << 7;
		key |= static_cast<uint32_t>(s.depthAttachmentType) << 9;
		return key; }

	bool MakeEnv(const GLState& s, DrawEnv& e) {
		e.st = &s;
		static const float kNoUniforms[UNIFORM_BUFFER_SIZE] = {};  // a draw without uniforms reads zeros
		e.uniforms = Uniforms(s) ? Uniforms(s) : kNoUniforms;
		e.vpm = mul(s.projectionMatrix, s.viewMatrix);  // MakeMatrices (rglv_gpu_impl.hxx:44-51)
		e.vm = s.viewMatrix;
		mat4 inv;
		mat4_inverse(s.viewMatrix.ff, inv.ff);
		mat4_transpose(inv.ff, e.nm.ff);
		e.tu3.buf = s.tu3ptr; e.tu3.dim = s.tu3dim;
		std::string terr;
		e.tu0 = Sampler::Make(s.tus[0], terr);
		e.tu1 = Sampler::Make(s.tus[1], terr);
		if (!terr.empty()) { err = terr; return false; }
		return true; }

	// ---- binning (GPU::BinImpl, rglv_gpu.cxx:119-260) ----
	bool Bin() {
		tiles.resize(static_cast<size_t>(tilesX) * tilesY);
		for (auto& t : tiles) t.clear();  // (keeps the lists' capacity from the previous frame)
		clipped.clear();
		const GLState* binState = nullptr;
		int binStateIdx = -1;
		uint32_t primBase = 0;
		for (const Cmd& c : commands) {
			switch (c.kind) {
			case Cmd::STATE:
				binStateIdx = c.a; binState = &states[c.a];
				for (auto& t : tiles) t.push_back({ TileOp::STATE, c.a, 0, 0, 0, 0, nullptr });
				break;
			case Cmd::CLEAR:
				for (auto& t : tiles) t.push_back({ TileOp::CLEAR, c.a, 0, 0, 0, 0, nullptr });
				break;
			case Cmd::STORE_TC:
				if (binState == nullptr || !KeyInstalled((static_cast<uint32_t>(binState->programId) << 24) | 0x40000000u)) {
					err = "STORE_COLOR: no blt dispatch entry"; return false; }
				for (auto& t : tiles) t.push_back({ TileOp::STORE_TC, c.a, 0, 0, 0, 0, c.ptr });
				break;
			case Cmd::STORE_DEPTH:
				for (auto& t : tiles) t.push_back({ TileOp::STORE_DEPTH, 0, 0, 0, 0, 0, c.ptr });
				break;
			case Cmd::STORE_PRIM:
				for (auto& t : tiles) t.push_back({ TileOp::STORE_PRIM, 0, 0, 0, 0, 0, c.ptr });
				break;
			case Cmd::STORE_FP:  // a = the canvas stride, i0 = 0 full, 1 half (downsample), 2 quads
				for (auto& t : tiles) t.push_back({ TileOp::STORE_FP, c.b, static_cast<uint16_t>(c.a), 0, 0, 0, c.ptr });
				break;
			default: {
				if (binState == nullptr) { err = "draw before state"; return false; }
				const uint32_t vkey = static_cast<uint32_t>(binState->programId) << 24;
				if (!KeyInstalled(vkey | 0x80000000u)) { err = "BIN: no dispatch entry"; return false; }
				if (!KeyInstalled(FragmentStateKey(*binState))) { err = "DRAW: no dispatch entry"; return false; }
				const bool elements = c.kind == Cmd::DRAW_ELEMENTS || c.kind == Cmd::DRAW_ELEMENTS_INSTANCED;
				const bool inst = c.kind == Cmd::DRAW_ELEMENTS_INSTANCED || c.kind == Cmd::DRAW_ARRAYS_INSTANCED;
				const int count = c.a;
				const int instanceCnt = inst ? c.b : 1;
				const auto* indices = static_cast<const uint16_t*>(c.ptr);
				if (!BinTriangles(*binState, binStateIdx, count, elements ? indices : nullptr, instanceCnt, inst, primBase)) return false;
				primBase += static_cast<uint32_t>(count / 3) * static_cast<uint32_t>(instanceCnt);
				break; }}}
		return true; }

	// BinTriangles1P (rglv_gpu_impl.hxx:517-675); 2P gives the same lists
	bool BinTriangles(const GLState& state, int stateIdx, int count, const uint16_t* indices, int instanceCnt,
	                  bool instanced, uint32_t primBase) {
		const Program* pgm = find_program(state.programId);
		if (!pgm) { err = "unknown program"; return false; }
		DrawEnv env;
		if (!MakeEnv(state, env)) return false;
		const Frustum frustum(W, H, fixedGuardBand);
		const Rect sr = ScissorRect(state);
		float DS[2], DO[2]; DSDO(state, DS, DO);
		const int tlx = sr.x0, tly = sr.y0, brx = sr.x1 - 1, bry = sr.y1 - 1;
		const bool keepBacks = !(state.cullingEnabled && state.cullFace == GL_BACK);
		const bool keepFronts = !(state.cullingEnabled && state.cullFace == GL_FRONT);
		const int numPrims = count / 3;
		std::vector<int> touched(tiles.size(), 0);
		for (auto& t : tiles) t.push_back({ TileOp::TRI, -1, 0, 0, 0, 0, nullptr });  // draw marker
		struct ClipItem { int iid; uint16_t i[3]; uint32_t prim; };
		std::vector<ClipItem> clipQueue;
		// each vertex of an instance shaded once (BinTriangles2P's vertex
		// phase, rglv_gpu_impl.hxx:314-420): clip flags and device coordinates
		int nVerts = 0;
		for (int k = 0; k < numPrims * 3; ++k) nVerts = std::max(nVerts, (indices ? indices[k] : k) + 1);
		std::vector<uint32_t> vcf(static_cast<size_t>(nVerts));
		std::vector<float> vdx(static_cast<size_t>(nVerts)), vdy(static_cast<size_t>(nVerts));
		for (int iid = 0; iid < instanceCnt; ++iid) {
			for (int v = 0; v < nVerts; ++v) {
				float pos[4]; VOut vo;
				pgm->ShadeVertex(env, iid, v, pos, vo);
				vcf[v] = frustum.Test(pos);
				const float iw = oneover(pos[3]);  // pdiv (rglv_math.hxx:19-23)
				vdx[v] = (pos[0] * iw) * DS[0] + DO[0];
				vdy[v] = (pos[1] * iw) * DS[1] + DO[1]; }
			for (int pi = 0; pi < numPrims; ++pi) {
				++statSubmitted;
				uint16_t ii[3];
				for (int k = 0; k < 3; ++k) ii[k] = indices ? indices[pi*3 + k] : static_cast<uint16_t>(pi*3 + k);
				uint32_t cf[3]; float dcx[3], dcy[3];
				for (int k = 0; k < 3; ++k) { cf[k] = vcf[ii[k]]; dcx[k] = vdx[ii[k]]; dcy[k] = vdy[ii[k]]; }
				const bool pointsOutside = (cf[0] | cf[1] | cf[2]) > 0;
				const bool primsOutside = (cf[0] & cf[1] & cf[2]) > 0;
				const uint32_t prim = primBase + static_cast<uint32_t>(iid) * numPrims + pi;
				if (primsOutside) continue;
				if (pointsOutside) clipQueue.push_back({ iid, { ii[0], ii[1], ii[2] }, prim });
				// rmlg::Area (rmlg_triangle.hxx:17-21)
				const float d31x = dcx[2] - dcx[0], d31y = dcy[2] - dcy[0];
				const float d21x = dcx[1] - dcx[0], d21y = dcy[1] - dcy[0];
				const float area = d31x*d21y - d31y*d21x;
				const bool front = area > 0.0F;
				const int32_t ix0 = cvtt(dcx[0]), iy0 = cvtt(dcy[0]);
				const int32_t ix1 = cvtt(dcx[1]), iy1 = cvtt(dcy[1]);
				const int32_t ix2 = cvtt(dcx[2]), iy2 = cvtt(dcy[2]);
				const int32_t vminx = std::max(std::min(ix0, std::min(ix1, ix2)), tlx);
				const int32_t vminy = std::max(std::min(iy0, std::min(iy1, iy2)), tly);
				const int32_t vmaxx = std::min(wadd(std::max(ix0, std::max(ix1, ix2)), 1), brx);
				const int32_t vmaxy = std::min(wadd(std::max(iy0, std::max(iy1, iy2)), 1), bry);
				bool notCulled = (!front && keepBacks) || (front && keepFronts);
				notCulled = notCulled && !pointsOutside;
				const bool accept = notCulled && vmaxx > vminx && vmaxy > vminy;
				if (!accept) continue;
				uint16_t n0 = front ? ii[0] : ii[2];
				const uint16_t n1 = ii[1];
				const uint16_t n2 = front ? ii[2] : ii[0];
				if (!front) n0 |= 0x8000;
				const int tx0 = vminx / tileW, ty0 = vminy / tileH, tx1 = vmaxx / tileW, ty1 = vmaxy / tileH;
				for (int ty = ty0; ty <= ty1; ++ty) {
					for (int tx = tx0; tx <= tx1; ++tx) {
						const int t = ty * tilesX + tx;
						touched[t] = 1;
						++statBinned;
						tiles[t].push_back({ TileOp::TRI, iid, n0, n1, n2, prim, nullptr }); }}}}
		for (size_t t = 0; t < tiles.size(); ++t) {
			if (!touched[t]) {
				tiles[t].pop_back(); }  // untouched: remove the marker (Unappend)
			else {
				tiles[t].push_back({ TileOp::TRI, -2, 0, 0, 0, 0, nullptr }); }}  // 0xffff terminator
		statClipped += static_cast<long long>(clipQueue.size());
		// ClipTriangles (rglv_gpu_impl.hxx:677-793)
		for (const ClipItem& ci : clipQueue) {
			std::vector<ClippedVertex> A;
			for (int k = 0; k < 3; ++k) {
				ClippedVertex cv{};
				pgm->ShadeVertex(env, ci.iid, ci.i[k], cv.coord, cv.data);
				A.push_back(cv); }
			clip_polygon(A, pgm->nvary);
			if (A.empty()) continue;
			for (auto& v : A) {
				const float iw = oneover(v.coord[3]);  // scalar pdiv (rglv_math.hxx:30-39)
				v.coord[0] = v.coord[0] * iw; v.coord[1] = v.coord[1] * iw; v.coord[2] = v.coord[2] * iw; v.coord[3] = iw;
				v.coord[0] = v.coord[0] * DS[0] + DO[0];
				v.coord[1] = v.coord[1] * DS[1] + DO[1]; }
			const float d31x = A[2].coord[0] - A[0].coord[0], d31y = A[2].coord[1] - A[0].coord[1];
			const float d21x = A[1].coord[0] - A[0].coord[0], d21y = A[1].coord[1] - A[0].coord[1];
			const bool backfacing = (d31x*d21y - d31y*d21x) < 0.0F;
			bool willCull = true;
			if (backfacing) { if (!state.cullingEnabled || (state.cullFace & GL_BACK) == 0) willCull = false; }
			else { if (!state.cullingEnabled || (state.cullFace & GL_FRONT) == 0) willCull = false; }
			if (willCull) continue;
			if (backfacing) std::reverse(A.begin(), A.end());
			const int bi = static_cast<int>(clipped.size());
			clipped.insert(clipped.end(), A.begin(), A.end());
			for (int vi = 1; vi < static_cast<int>(A.size()) - 1; ++vi) {
				const int i0 = 0, i1 = vi, i2 = vi + 1;
				// ForEachCoveredTile (:795-817)
				int idx[3] = { cvtt(A[i0].coord[0]), cvtt(A[i1].coord[0]), cvtt(A[i2].coord[0]) };
				int idy[3] = { cvtt(A[i0].coord[1]), cvtt(A[i1].coord[1]), cvtt(A[i2].coord[1]) };
				const int vminx = std::max(std::min({ idx[0], idx[1], idx[2] }), sr.x0);
				const int vminy = std::max(std::min({ idy[0], idy[1], idy[2] }), sr.y0);
				const int vmaxx = std::min(wadd(std::max({ idx[0], idx[1], idx[2] }), 1), sr.x1 - 1);
				const int vmaxy = std::min(wadd(std::max({ idy[0], idy[1], idy[2] }), 1), sr.y1 - 1);
				const int tx0 = vminx / tileW, ty0 = vminy / tileH, tx1 = vmaxx / tileW, ty1 = vmaxy / tileH;
				for (int ty = ty0; ty <= ty1; ++ty) {
					for (int tx = tx0; tx <= tx1; ++tx) {
						if (tx < 0 || ty < 0 || tx >= tilesX || ty >= tilesY) { err = "clipped triangle outside tile grid"; return false; }
						++statClippedPairs;
						tiles[ty * tilesX + tx].push_back({ TileOp::CLIPPED_TRI, stateIdx,
							static_cast<uint16_t>((bi + i0) | (backfacing ? 0x8000 : 0)),
							static_cast<uint16_t>(bi + i1), static_cast<uint16_t>(bi + i2), ci.prim, nullptr }); }}}}
		return true; }

	// ---- per-tile drawing (GPU::DrawImpl, rglv_gpu.cxx:263-432) ----
	struct TileBuf {
		std::vector<float> r, g, b, d, a;   // a: RB_RGBAF32's fourth channel (set by clears only)
		std::vector<uint32_t> prim;
		long long fragments = 0;
		long long quads = 0; };
	std::vector<long long> tileQuads, tilePairs;

	struct TriCtx {
		const Program* pgm;
		const DrawEnv* env;
		float w[3], z[3];
		VOut vo[3];
		uint32_t prim;
		int depthFunc;
		bool depthTest, depthWrite, colorWrite, blend; };

	// TriangleProgram::Render (rglv_gpu_impl.hxx:166-222) for one quad
	void Render(TileBuf& tb, const Rect& tr, const TriCtx& tc, int qx, int qy, const float fc[2][4],
	            const int32_t trimask[4], const float bs[3][4]) const {
		int pix[4];
		for (int l = 0; l < 4; ++l) {
			const int px = qx + (l & 1), py = qy + (l >> 1);
			pix[l] = (py - tr.y0) * tileW + (px - tr.x0); }
		float fragDepth[4];
		for (int l = 0; l < 4; ++l) fragDepth[l] = (bs[0][l]*tc.z[0] + bs[1][l]*tc.z[1]) + bs[2][l]*tc.z[2];
		for (int l = 0; l < 4; ++l) if (!trimask[l]) ++tb.fragments;  // covered lanes reaching the depth test
		int32_t mask[4];
		auto depthPass = [&](int l) {
			const float dst = tb.d[pix[l]];
			switch (tc.depthFunc) {
			case GL_LESS: return fragDepth[l] < dst;
			case GL_LEQUAL: return fragDepth[l] <= dst;
			default: return fragDepth[l] == dst; }};
		const bool earlyZ = tc.pgm->earlyZ;
		if (tc.depthTest && earlyZ) {
			bool any = false;
			for (int l = 0; l < 4; ++l) { mask[l] = (!trimask[l] && depthPass(l)) ? -1 : 0; any |= mask[l] != 0; }
			if (!any) return; }
		else {
			for (int l = 0; l < 4; ++l) mask[l] = trimask[l] ? 0 : -1; }
		if (tc.depthWrite && earlyZ) {
			for (int l = 0; l < 4; ++l) if (mask[l]) tb.d[pix[l]] = fragDepth[l]; }
		// perspective-correct barycentrics
		float bp[3][4];
		for (int l = 0; l < 4; ++l) {
			const float fragW = oneover((bs[0][l]*tc.w[0] + bs[1][l]*tc.w[1]) + bs[2][l]*tc.w[2]);
			bp[0][l] = (tc.w[0] * bs[0][l]) * fragW;
			bp[2][l] = (tc.w[2] * bs[2][l]) * fragW;
			bp[1][l] = (1.0F - bp[0][l]) - bp[2][l]; }
		float vary[kMaxVary][4];
		for (int k = 0; k < tc.pgm->nvary; ++k) {
			for (int l = 0; l < 4; ++l) {
				const float b3[3] = { bp[0][l], bp[1][l], bp[2][l] };
				vary[k][l] = interp(b3, tc.vo[0].f[k], tc.vo[1].f[k], tc.vo[2].f[k]); }}
		float color[4][4];
		tc.pgm->ShadeFragment(*tc.env, bs, bp, vary, fc, fragDepth, color, mask);
		if (tc.depthTest && !earlyZ) {
			bool any = false;
			for (int l = 0; l < 4; ++l) { if (mask[l] && !depthPass(l)) mask[l] = 0; any |= mask[l] != 0; }
			if (!any) return; }
		if (tc.depthWrite && !earlyZ) {
			for (int l = 0; l < 4; ++l) if (mask[l]) tb.d[pix[l]] = fragDepth[l]; }
		if (tc.colorWrite) {
			for (int l = 0; l < 4; ++l) {
				if (!mask[l]) continue;
				const int p = pix[l];
				if (tc.blend) {  // BlendAlpha (:80-84)
					const float a = color[3][l], oma = 1.0F - a;
					tb.r[p] = color[0][l]*a + tb.r[p]*oma;
					tb.g[p] = color[1][l]*a + tb.g[p]*oma;
					tb.b[p] = color[2][l]*a + tb.b[p]*oma; }
				else {
					tb.r[p] = color[0][l]; tb.g[p] = color[1][l]; tb.b[p] = color[2][l]; }}}
		for (int l = 0; l < 4; ++l) if (mask[l]) tb.prim[pix[l]] = tc.prim; }

	// ScissorFlags (rglv_triangle.hxx:20-29): the raster rect's odd edges
	static int ScissorFlags(const Rect& r) {
		return ((r.y0 & 1) ? 1 : 0) | ((r.y1 & 1) ? 2 : 0) | ((r.x0 & 1) ? 4 : 0) | ((r.x1 & 1) ? 8 : 0); }

	// VTriangleRasterizer::Draw (rglv_triangle.hxx:193-304), one lane; tr is
	// the tile (buffer addressing), rr the raster rect, scissor the
	// SCISSOR_TEST template argument (lanes of quads on the rect's odd edges
	// masked, :270-300)
	void RasterVector(TileBuf& tb, const Rect& tr, const Rect& rr, bool scissor, const TriCtx& tc, const int32_t X[3],
	                  const int32_t Y[3]) const {
		const int32_t x1 = X[0], x2 = X[1], x3 = X[2], y1 = Y[0], y2 = Y[1], y3 = Y[2];
		int32_t minx = std::max(sar(std::min(std::min(x1, x2), x3), 4), rr.x0);
		const int32_t maxx = std::min(sar(wadd(std::max(std::max(x1, x2), x3), 15), 4), rr.x1);
		int32_t miny = std::max(sar(std::min(std::min(y1, y2), y3), 4), rr.y0);
		const int32_t maxy = std::min(sar(wadd(std::max(std::max(y1, y2), y3), 15), 4), rr.y1);
		const int sflags = scissor ? ScissorFlags(rr) : 0;
		minx &= ~1; miny &= ~1;
		const int32_t dx12 = wsub(x1, x2), dy12 = wsub(y2, y1);
		const int32_t dx23 = wsub(x2, x3), dy23 = wsub(y3, y2);
		const int32_t dx31 = wsub(x3, x1), dy31 = wsub(y1, y3);
		const int32_t sx = wadd(shl(minx, 4), 8), sy = wadd(shl(miny, 4), 8);
		int32_t c1 = wadd(wmul(dy12, wsub(sx, x1)), wmul(dx12, wsub(sy, y1)));
		int32_t c2 = wadd(wmul(dy23, wsub(sx, x2)), wmul(dx23, wsub(sy, y2)));
		int32_t c3 = wadd(wmul(dy31, wsub(sx, x3)), wmul(dx31, wsub(sy, y3)));
		auto tl = [](int32_t dy, int32_t dx) { return (dy > 0 || (dy == 0 && dx > 0)) ? 1 : 0; };
		c1 = sar(wsub(wadd(c1, tl(dy12, dx12)), 1), 4);
		c2 = sar(wsub(wadd(c2, tl(dy23, dx23)), 1), 4);
		c3 = sar(wsub(wadd(c3, tl(dy31, dx31)), 1), 4);
		const float scale = 1.0F / itof(wadd(wadd(c1, c2), c3));
		const float topRow = static_cast<float>(H) - 0.5F;
		int32_t r1 = c1, r2 = c2, r3 = c3;
		for (int y = miny; y < maxy; y += 2, r1 = wadd(r1, wmul(dx12, 2)), r2 = wadd(r2, wmul(dx23, 2)), r3 = wadd(r3, wmul(dx31, 2))) {
			const int32_t sTop = (y == (rr.y0 & ~1) && (sflags & 1)) ? -1 : 0;
			const int32_t sBot = (y >= (rr.y1 & ~1) && (sflags & 2)) ? -1 : 0;
			int32_t q1 = r1, q2 = r2, q3 = r3;
			for (int x = minx; x < maxx; x += 2, q1 = wadd(q1, wmul(dy12, 2)), q2 = wadd(q2, wmul(dy23, 2)), q3 = wadd(q3, wmul(dy31, 2))) {
				int32_t e1[4] = { q1, wadd(q1, dy12), wadd(q1, dx12), wadd(wadd(q1, dx12), dy12) };
				int32_t e2[4] = { q2, wadd(q2, dy23), wadd(q2, dx23), wadd(wadd(q2, dx23), dy23) };
				int32_t e3[4] = { q3, wadd(q3, dy31), wadd(q3, dx31), wadd(wadd(q3, dx31), dy31) };
				int32_t trimask[4]; int neg = 0;
				for (int l = 0; l < 4; ++l) { trimask[l] = (e1[l] | e2[l] | e3[l]) >> 31; neg += trimask[l] ? 1 : 0; }
				++tb.quads;
				if (neg == 4) continue;
				if (scissor) {
					const int32_t sLeft = (x == (rr.x0 & ~1) && (sflags & 4)) ? -1 : 0;
					const int32_t sRight = (x >= (rr.x1 & ~1) && (sflags & 8)) ? -1 : 0;
					trimask[0] |= sLeft | sTop; trimask[1] |= sRight | sTop; trimask[2] |= sLeft | sBot; trimask[3] |= sRight | sBot; }
				float fc[2][4], bs[3][4];
				const float fx = static_cast<float>(x) + 0.5F;
				const float fy = topRow - static_cast<float>(y);
				for (int l = 0; l < 4; ++l) {
					fc[0][l] = fx + static_cast<float>(l & 1);
					fc[1][l] = fy - static_cast<float>(l >> 1);
					bs[0][l] = itof(e2[l]) * scale;
					bs[2][l] = itof(e1[l]) * scale;
					bs[1][l] = (1.0F - bs[0][l]) - bs[2][l]; }
				Render(tb, tr, tc, x, y, fc, trimask, bs); }}}

	// TriangleRasterizer::Draw (rglv_triangle.hxx:79-167), clipped triangles;
	// tr, rr and scissor as for RasterVector
	void RasterScalar(TileBuf& tb, const Rect& tr, const Rect& rr, bool scissor, const TriCtx& tc, const int32_t X[3],
	                  const int32_t Y[3]) const {
		const int x1 = X[0], x2 = X[1], x3 = X[2], y1 = Y[0], y2 = Y[1], y3 = Y[2];
		int minx = std::max(std::min({ x1, x2, x3 }) >> 4, rr.x0);
		const int endx = std::min((std::max({ x1, x2, x3 }) + 15) >> 4, rr.x1);
		int miny = std::max(std::min({ y1, y2, y3 }) >> 4, rr.y0);
		const int endy = std::min((std::max({ y1, y2, y3 }) + 15) >> 4, rr.y1);
		const int sflags = scissor ? ScissorFlags(rr) : 0;
		minx &= ~1; miny &= ~1;
		const int64_t dx12 = x1 - x2, dy12 = y2 - y1;
		const int64_t dx23 = x2 - x3, dy23 = y3 - y2;
		const int64_t dx31 = x3 - x1, dy31 = y1 - y3;
		const int64_t sx = (static_cast<int64_t>(minx) << 4) + 8, sy = (static_cast<int64_t>(miny) << 4) + 8;
		int64_t c1 = dy12*(sx - x1) + dx12*(sy - y1);
		int64_t c2 = dy23*(sx - x2) + dx23*(sy - y2);
		int64_t c3 = dy31*(sx - x3) + dx31*(sy - y3);
		if (dy12 > 0 || (dy12 == 0 && dx12 > 0)) { c1++; }
		--c1;
		if (dy23 > 0 || (dy23 == 0 && dx23 > 0)) { c2++; }
		--c2;
		if (dy31 > 0 || (dy31 == 0 && dx31 > 0)) { c3++; }
		--c3;
		c1 >>= 4; c2 >>= 4; c3 >>= 4;
		int32_t cb1[4] = { static_cast<int32_t>(c1), static_cast<int32_t>(c1 + dy12), static_cast<int32_t>(c1 + dx12), static_cast<int32_t>(c1 + dx12 + dy12) };
		int32_t cb2[4] = { static_cast<int32_t>(c2), static_cast<int32_t>(c2 + dy23), static_cast<int32_t>(c2 + dx23), static_cast<int32_t>(c2 + dx23 + dy23) };
		int32_t cb3[4] = { static_cast<int32_t>(c3), static_cast<int32_t>(c3 + dy31), static_cast<int32_t>(c3 + dx31), static_cast<int32_t>(c3 + dx31 + dy31) };
		const int32_t s1x = wmul(static_cast<int32_t>(dy12), 2), s1y = wmul(static_cast<int32_t>(dx12), 2);
		const int32_t s2x = wmul(static_cast<int32_t>(dy23), 2), s2y = wmul(static_cast<int32_t>(dx23), 2);
		const int32_t s3x = wmul(static_cast<int32_t>(dy31), 2), s3y = wmul(static_cast<int32_t>(dx31), 2);
		const float scale = 1.0F / static_cast<float>(c1 + c2 + c3);
		for (int y = miny; y < endy; y += 2) {
			const int32_t sTop = (y == (rr.y0 & ~1) && (sflags & 1)) ? -1 : 0;
			const int32_t sBot = (y >= (rr.y1 & ~1) && (sflags & 2)) ? -1 : 0;
			int32_t q1[4], q2[4], q3[4];
			std::memcpy(q1, cb1, sizeof q1); std::memcpy(q2, cb2, sizeof q2); std::memcpy(q3, cb3, sizeof q3);
			for (int x = minx; x < endx; x += 2) {
				int32_t trimask[4]; int neg = 0;
				for (int l = 0; l < 4; ++l) { trimask[l] = (q1[l] | q2[l] | q3[l]) >> 31; neg += trimask[l] ? 1 : 0; }
				if (neg != 4 && scissor) {
					const int32_t sLeft = (x == (rr.x0 & ~1) && (sflags & 4)) ? -1 : 0;
					const int32_t sRight = (x >= (rr.x1 & ~1) && (sflags & 8)) ? -1 : 0;
					trimask[0] |= sLeft | sTop; trimask[1] |= sRight | sTop; trimask[2] |= sLeft | sBot; trimask[3] |= sRight | sBot; }
				if (neg != 4) {
					float fc[2][4], bs[3][4];
					const float fx = static_cast<float>(x) + 0.5F;
					const float fy = static_cast<float>(H - y) - 0.5F;
					for (int l = 0; l < 4; ++l) {
						fc[0][l] = fx + static_cast<float>(l & 1);
						fc[1][l] = fy + static_cast<float>(l < 2 ? 1 : 0);
						bs[0][l] = itof(q2[l]) * scale;
						bs[2][l] = itof(q1[l]) * scale;
						bs[1][l] = (1.0F - bs[0][l]) - bs[2][l]; }
					Render(tb, tr, tc, x, y, fc, trimask, bs); }
				for (int l = 0; l < 4; ++l) { q1[l] = wadd(q1[l], s1x); q2[l] = wadd(q2[l], s2x); q3[l] = wadd(q3[l], s3x); }}
			for (int l = 0; l < 4; ++l) { cb1[l] = wadd(cb1[l], s1y); cb2[l] = wadd(cb2[l], s2y); cb3[l] = wadd(cb3[l], s3y); }}}

	bool DrawTile(int t, TileBuf& tb) {
		const Rect tr = TileRect(t);
		tb.quads = 0;
		long long pairs = 0;
		for (const TileOp& op : tiles[t]) pairs += (op.op == TileOp::TRI && op.a >= 0) || op.op == TileOp::CLIPPED_TRI;
		struct Rec { rglo_gpu* g; int t; TileBuf& tb; long long pairs;
			~Rec() { g->tileQuads[t] = tb.quads; g->tilePairs[t] = pairs; } } rec{ this, t, tb, pairs };
		const size_t area = static_cast<size_t>(tileW) * tileH;
		tb.r.assign(area, 0.0F); tb.g.assign(area, 0.0F); tb.b.assign(area, 0.0F);
		tb.d.assign(area, 0.0F); tb.a.assign(area, 0.0F); tb.prim.assign(area, 0xffffffffu);
		const GLState* st = nullptr;
		DrawEnv env{};
		const Program* pgm = nullptr;
		bool inDraw = false;
		for (const TileOp& op : tiles[t]) {
			switch (op.op) {
			case TileOp::STATE:
				st = &states[op.a];
				if (!MakeEnv(*st, env)) return false;
				pgm = find_program(st->programId);
				break;
			case TileOp::CLEAR:
				if (op.a & GL_COLOR_BUFFER_BIT) {
					std::fill(tb.r.begin(), tb.r.end(), st->clearColor[0]);
					std::fill(tb.g.begin(), tb.g.end(), st->clearColor[1]);
					std::fill(tb.b.begin(), tb.b.end(), st->clearColor[2]);
					std::fill(tb.a.begin(), tb.a.end(), st->clearColor[3]);
					std::fill(tb.prim.begin(), tb.prim.end(), 0xffffffffu); }
				if (op.a & GL_DEPTH_BUFFER_BIT) std::fill(tb.d.begin(), tb.d.end(), st->clearDepth);
				break;
			case TileOp::TRI: {
				if (op.a == -1) { inDraw = true; break; }
				if (op.a == -2) { inDraw = false; break; }
				if (!inDraw || !pgm) { err = "triangle outside a draw"; return false; }
				TriCtx tc{ pgm, &env, {}, {}, {}, op.prim, st->depthFunc, st->depthTestEnabled,
				           st->depthWriteMask, st->colorWriteMask, st->blendingEnabled };
				float DS[2], DO[2]; DSDO(*st, DS, DO);
				const uint16_t idx[3] = { static_cast<uint16_t>(op.i0 & 0x7fff), op.i1, op.i2 };
				int32_t X[3], Y[3];
				for (int k = 0; k < 3; ++k) {  // GPUTileImpl::DrawTriangles flush (:880-886)
					float pos[4];
					pgm->ShadeVertex(env, op.a, idx[k], pos, tc.vo[k]);
					const float iw = oneover(pos[3]);
					const float dx = pos[0] * iw, dy = pos[1] * iw, dz = pos[2] * iw;
					X[k] = cvtt(16.0F * (dx * DS[0] + DO[0]));
					Y[k] = cvtt(16.0F * (dy * DS[1] + DO[1]));
					tc.w[k] = iw; tc.z[k] = dz; }
				RasterVector(tb, tr, RasterRect(tr, *st), st->scissorEnabled, tc, X, Y);
				break; }
			case TileOp::CLIPPED_TRI: {  // DrawClipped (:949-998)
				if (!pgm) { err = "clipped triangle outside a draw"; return false; }
				TriCtx tc{ pgm, &env, {}, {}, {}, op.prim, st->depthFunc, st->depthTestEnabled,
				           st->depthWriteMask, st->colorWriteMask, st->blendingEnabled };
				const int idx[3] = { op.i0 & 0x7fff, op.i1, op.i2 };
				int32_t X[3], Y[3];
				for (int k = 0; k < 3; ++k) {
					const ClippedVertex& cv = clipped[idx[k]];
					X[k] = static_cast<int32_t>(cvtt(cv.coord[0] * 16.0F));
					Y[k] = static_cast<int32_t>(cvtt(cv.coord[1] * 16.0F));
					tc.w[k] = cv.coord[3]; tc.z[k] = cv.coord[2]; tc.vo[k] = cv.data; }
				RasterScalar(tb, tr, RasterRect(tr, *st), st->scissorEnabled, tc, X, Y);
				break; }
			case TileOp::STORE_TC: {  // GPUBltImpl::StoreTrueColor -> FilterTile (rglr_algorithm.hxx:31-104)
				if (!known_post_program(st->programId)) { err = "no blt dispatch entry"; return false; }
				const float* u = Uniforms(*st);
				const float exposure = (st->programId == 2 && u) ? u[0] : 1.0F;
				auto* out = static_cast<uint32_t*>(const_cast<void*>(op.ptr));
				// the canvas coordinate q as FilterTile steps it: per 2 x 2
				// quad, fcx += 2 / W along the row, fcy += -2 / H per row pair,
				// both restarting at the tile's corner
				const float iw = 1.0F / static_cast<float>(W), ih = 1.0F / static_cast<float>(H);
				const float fcdx = iw * 2.0F, fcdy = -ih * 2.0F;
				const float fcxLeft = (static_cast<float>(tr.x0) + 0.5F) / static_cast<float>(W);
				float fcy[4];
				for (int l = 0; l < 4; ++l) {
					fcy[l] = (static_cast<float>(H - tr.y0) - 0.5F) / static_cast<float>(H);
					if (l >= 2) fcy[l] = fcy[l] - ih; }
				for (int y = tr.y0; y < tr.y1; y += 2) {
					float fcx[4];
					for (int l = 0; l < 4; ++l) fcx[l] = (l & 1) ? fcxLeft + iw : fcxLeft;
					for (int x = tr.x0; x < tr.x1; x += 2) {
						for (int l = 0; l < 4; ++l) {
							const int px_ = x + (l & 1), py_ = y + (l >> 1);
							if (px_ >= tr.x1 || py_ >= tr.y1) continue;
							const int p = (py_ - tr.y0) * tileW + (px_ - tr.x0);
							float c[3] = { tb.r[p], tb.g[p], tb.b[p] };
							if (st->programId == 2) { c[0] = c[0] * exposure; c[1] = c[1] * exposure; c[2] = c[2] * exposure; }
							if (st->programId == 3) iq_post(c, fcx[l], fcy[l]);
							uint32_t px;
							if (op.a) px = (to_srgb8(c[0]) << 16) | (to_srgb8(c[1]) << 8) | to_srgb8(c[2]);
							else px = (to_linear8(c[0]) << 16) | (to_linear8(c[1]) << 8) | to_linear8(c[2]);
							out[static_cast<size_t>(py_) * W + px_] = px; }
						for (int l = 0; l < 4; ++l) fcx[l] = fcx[l] + fcdx; }
					for (int l = 0; l < 4; ++l) fcy[l] = fcy[l] + fcdy; }
				break; }
			case TileOp::STORE_DEPTH: {
				auto* out = static_cast<float*>(const_cast<void*>(op.ptr));
				for (int y = tr.y0; y < tr.y1; ++y)
					for (int x = tr.x0; x < tr.x1; ++x)
						out[static_cast<size_t>(y) * W + x] = tb.d[(y - tr.y0) * tileW + (x - tr.x0)];
				break; }
			case TileOp::STORE_PRIM: {
				auto* out = static_cast<uint32_t*>(const_cast<void*>(op.ptr));
				for (int y = tr.y0; y < tr.y1; ++y)
					for (int x = tr.x0; x < tr.x1; ++x)
						out[static_cast<size_t>(y) * W + x] = tb.prim[(y - tr.y0) * tileW + (x - tr.x0)];
				break; }
			case TileOp::STORE_FP: {  // CMD_STORE_COLOR_{FULL,HALF}_LINEAR_FP, _FULL_QUADS_FP (rglv_gpu.cxx:345-384)
				const int ct = st->color0AttachmentType;
				if (ct != RB_COLOR_DEPTH && ct != RB_RGBF32 && ct != RB_RGBAF32) { err = "float store: attachment type"; return false; }
				auto* out = static_cast<float*>(const_cast<void*>(op.ptr));
				const long long stride = op.a;
				auto px = [&](int x, int y) { return (y - tr.y0) * tileW + (x - tr.x0); };
				if (op.i0 == 0) {  // Copy to a FloatingPointCanvas (rglr_algorithm.cxx:194-280): R, G, B, 0
					for (int y = tr.y0; y < tr.y1; ++y)
						for (int x = tr.x0; x < tr.x1; ++x) {
							float* o = out + 4 * (static_cast<long long>(y) * stride + x);
							const int p = px(x, y);
							o[0] = tb.r[p]; o[1] = tb.g[p]; o[2] = tb.b[p]; o[3] = 0.0F; }}
				else {
					// per 2 x 2 quad with its top-left pixel in the tile rect; the
					// quad's pixels in lane order (x, y), (x+1, y), (x, y+1),
					// (x+1, y+1), read from the whole tile buffer
					for (int y = tr.y0; y < tr.y1; y += 2)
						for (int x = tr.x0; x < tr.x1; x += 2) {
							int p[4];
							for (int l = 0; l < 4; ++l) p[l] = px(x + (l & 1), y + (l >> 1));
							const long long q = static_cast<long long>(y / 2) * stride + x / 2;
							if (op.i0 == 1) {  // Downsample (rglr_algorithm.cxx:118-165)
								float* o = out + 4 * q;
								const std::vector<float>* ch[3] = { &tb.r, &tb.g, &tb.b };
								for (int c = 0; c < 3; ++c) {
									const auto& v = *ch[c];
									o[c] = (((v[p[0]] + v[p[1]]) + v[p[2]]) + v[p[3]]) * 0.25F; }
								o[3] = (((0.0F + 0.0F) + 0.0F) + 0.0F) * 0.25F; }
							else {  // Copy to a QFloat4Canvas (rglr_algorithm.cxx:320-368)
								float* o = out + 16 * q;
								for (int l = 0; l < 4; ++l) {
									o[l] = tb.r[p[l]]; o[4 + l] = tb.g[p[l]]; o[8 + l] = tb.b[p[l]];
									o[12 + l] = ct == RB_RGBF32 ? 1.0F : ct == RB_COLOR_DEPTH ? tb.d[p[l]] : tb.a[p[l]]; }}}}
				break; }}}
		return true; }

	double lastBinSec = 0.0, lastDrawSec = 0.0;  // the last Run's serial bin and parallel draw

	bool Run() {
		err.clear();
		const auto t0 = std::chrono::steady_clock::now();
		if (!Bin()) return false;
		const auto t1 = std::chrono::steady_clock::now();
		lastBinSec = std::chrono::duration<double>(t1 - t0).count();
		tileQuads.assign(tiles.size(), 0);
		tilePairs.assign(tiles.size(), 0);
		// cost-sorted tile order (GPU::RunImpl, rglv_gpu.cxx:100-104)
		std::vector<std::pair<size_t, int>> order;
		for (int t = 0; t < static_cast<int>(tiles.size()); ++t) order.push_back({ tiles[t].size(), t });
		std::stable_sort(order.begin(), order.end(), [](auto a, auto b) { return a.first > b.first; });
		std::atomic<int> next{ 0 };
		std::atomic<bool> ok{ true };
		std::atomic<long long> frags{ 0 };
		std::vector<std::string> errs(static_cast<size_t>(std::max(threads, 1)));
		auto worker = [&](int wid) {
			TileBuf tb;
			for (;;) {
				const int i = next.fetch_add(1);
				if (i >= static_cast<int>(order.size())) break;
				if (!DrawTileSafe(order[i].second, tb, errs[wid])) { ok = false; break; }}
			frags += tb.fragments; };
		const int nt = std::max(threads, 1);
		if (nt == 1) worker(0);
		else {
			std::vector<std::thread> pool;
			for (int i = 0; i < nt; ++i) pool.emplace_back(worker, i);
			for (auto& th : pool) th.join(); }
		lastDrawSec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
		statFragments += frags.load();
		if (!ok) { for (auto& e : errs) if (!e.empty()) err = e; return false; }
		return true; }

	bool DrawTileSafe(int t, TileBuf& tb, std::string& e) {
		const bool r = DrawTile(t, tb);
		if (!r) e = err.empty() ? std::string("draw failed") : err;
		return r; }

	void SetSize(int w, int h, int tbx, int tby) {  // GPU::Reset / Retile (rglv_gpu.cxx:50-87)
		W = w; H = h; tileW = tbx * 8; tileH = tby * 8;
		tilesX = (W + tileW - 1) / tileW; tilesY = (H + tileH - 1) / tileH; }

	void ResetGL() {  // GL::Reset (rglv_gl.cxx:440-445)
		dirty = true; states.clear(); cs.reset(); ubuf.clear(); commands.clear(); }
};

// ------------------------------------------------------------------------
// C ABI, same shape as the product's rgl_* entry points
// ------------------------------------------------------------------------
extern "C" {

rglo_gpu* rglo_gpu_create(int threads) { auto* g = new rglo_gpu(); g->threads = threads; return g; }
void rglo_gpu_destroy(rglo_gpu* g) { delete g; }
void rglo_gpu_set_threads(rglo_gpu* g, int threads) { g->threads = threads; }
void rglo_gpu_set_guard_band(rglo_gpu* g, int fixed) { g->fixedGuardBand = fixed != 0; }
const char* rglo_gpu_last_error(rglo_gpu* g) { return g->err.c_str(); }
int rglo_gpu_install(rglo_gpu* g, int programId, uint32_t stateKey, int kind) {
	// kind 0 bin, 1 draw, 2 blt.  Bin entries are tagged with the top bit.
	uint32_t key = (static_cast<uint32_t>(programId) << 24) | stateKey;
	if (kind == 0) key |= 0x80000000u;
	if (kind == 2) key |= 0x40000000u;
	g->installed.push_back(key);
	return 0; }
int rglo_gpu_reset(rglo_gpu* g, int w, int h, int tbx, int tby) {
	if (w <= 0 || h <= 0 || tbx <= 0 || tby <= 0 || tbx > 16 || tby > 16) { g->err = "bad size"; return -1; }
	g->SetSize(w, h, tbx, tby); g->ResetGL(); return 0; }
int rglo_gpu_run(rglo_gpu* g) { return g->Run() ? 0 : -1; }
// the last run's wall time: out[0] bin (one thread), out[1] draw + stores
void rglo_gpu_times(rglo_gpu* g, double out[2]) { out[0] = g->lastBinSec; out[1] = g->lastDrawSec; }

void rglo_gpu_stats(rglo_gpu* g, long long out[5]) {
	out[0] = g->statSubmitted; out[1] = g->statClipped; out[2] = g->statBinned; out[3] = g->statFragments; out[4] = g->statClippedPairs;
	g->statSubmitted = g->statClipped = g->statBinned = g->statFragments = g->statClippedPairs = 0; }

void rglo_gl_enable(rglo_gpu* g, int v) {
	g->dirty = true;
	if (v == GL_CULL_FACE) g->cs.cullingEnabled = true;
	else if (v == GL_SCISSOR_TEST) g->cs.scissorEnabled = true;
	else if (v == GL_BLEND) g->cs.blendingEnabled = true;
	else if (v == GL_DEPTH_TEST) g->cs.depthTestEnabled = true; }
void rglo_gl_disable(rglo_gpu* g, int v) {
	g->dirty = true;
	if (v == GL_CULL_FACE) g->cs.cullingEnabled = false;
	else if (v == GL_SCISSOR_TEST) g->cs.scissorEnabled = false;
	else if (v == GL_BLEND) g->cs.blendingEnabled = false;
	else if (v == GL_DEPTH_TEST) g->cs.depthTestEnabled = false; }
void rglo_gl_depth_func(rglo_gpu* g, int v) { g->dirty = true; g->cs.depthFunc = v; }
void rglo_gl_depth_write_mask(rglo_gpu* g, int v) { g->dirty = true; g->cs.depthWriteMask = v != 0; }
void rglo_gl_color_write_mask(rglo_gpu* g, int v) { g->dirty = true; g->cs.colorWriteMask = v != 0; }
void rglo_gl_cull_face(rglo_gpu* g, int v) { g->dirty = true; g->cs.cullFace = v; }
void rglo_gl_scissor(rglo_gpu* g, int x, int y, int w, int h) {
	g->dirty = true; g->cs.scissorOrigin[0] = x; g->cs.scissorOrigin[1] = y; g->cs.scissorSize[0] = w; g->cs.scissorSize[1] = h; }
void rglo_gl_viewport(rglo_gpu* g, int x, int y, int w, int h) {
	g->dirty = true; g->cs.viewportOrigin[0] = x; g->cs.viewportOrigin[1] = y;
	g->cs.hasViewportSize = true; g->cs.viewportSize[0] = w; g->cs.viewportSize[1] = h; }
void rglo_gl_use_program(rglo_gpu* g, int v) { g->dirty = true; g->cs.programId = v; }
void rglo_gl_use_uniforms(rglo_gpu* g, int ofs) { g->dirty = true; g->cs.uniformsOfs = ofs; }
int rglo_gl_alloc_uniform_buffer(rglo_gpu* g, const float* values, int n) {
	const int idx = static_cast<int>(g->ubuf.size());
	g->ubuf.resize(g->ubuf.size() + UNIFORM_BUFFER_SIZE, 0.0F);
	for (int i = 0; i < n && i < UNIFORM_BUFFER_SIZE; ++i) g->ubuf[idx + i] = values[i];
	return idx; }
void rglo_gl_renderbuffer_type(rglo_gpu* g, int attachment, int type) {
	g->dirty = true;
	if (attachment == GL_DEPTH_ATTACHMENT) g->cs.depthAttachmentType = type;
	else if (attachment == GL_COLOR_ATTACHMENT0) g->cs.color0AttachmentType = type; }
void rglo_gl_bind_texture(rglo_gpu* g, int unit, const float* ptr, int w, int h, int stride, int mode) {
	g->dirty = true; g->cs.tus[unit] = { ptr, w, h, stride, mode }; }
void rglo_gl_use_buffer(rglo_gpu* g, int idx, const float* ptr) { g->dirty = true; g->cs.buffers[idx] = ptr; }
void rglo_gl_bind_texture3(rglo_gpu* g, const float* ptr, int dim) { g->dirty = true; g->cs.tu3ptr = ptr; g->cs.tu3dim = dim; }
void rglo_gl_clear_color(rglo_gpu* g, float r, float gg, float b) {
	g->dirty = true; g->cs.clearColor[0] = r; g->cs.clearColor[1] = gg; g->cs.clearColor[2] = b; g->cs.clearColor[3] = 1.0F; }
void rglo_gl_clear_depth(rglo_gpu* g, float v) { g->dirty = true; g->cs.clearDepth = v; }
void rglo_gl_view_matrix(rglo_gpu* g, const float* m) { g->dirty = true; std::memcpy(g->cs.viewMatrix.ff, m, 64); }
void rglo_gl_projection_matrix(rglo_gpu* g, const float* m) { g->dirty = true; std::memcpy(g->cs.projectionMatrix.ff, m, 64); }
void rglo_gl_normal_matrix(rglo_gpu* g, const float* m) { g->dirty = true; std::memcpy(g->cs.normalMatrix.ff, m, 64); }
void rglo_gl_draw_elements(rglo_gpu* g, int, int count, int, const uint16_t* indices, int) {
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ELEMENTS, count, 1, 0, indices }); }
void rglo_gl_draw_elements_instanced(rglo_gpu* g, int, int count, int, const uint16_t* indices, int instanceCnt) {
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ELEMENTS_INSTANCED, count, instanceCnt, 0, indices }); }
void rglo_gl_draw_arrays(rglo_gpu* g, int, int, int count) {
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ARRAYS, count, 1, 0, nullptr }); }
void rglo_gl_draw_arrays_instanced(rglo_gpu* g, int, int, int count, int instanceCnt) {
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::DRAW_ARRAYS_INSTANCED, count, instanceCnt, 0, nullptr }); }
void rglo_gl_clear(rglo_gpu* g, int bits) { g->MaybeUpdateState(); g->commands.push_back({ Cmd::CLEAR, bits, 0, 0, nullptr }); }
void rglo_gl_store_color_tc(rglo_gpu* g, uint32_t* dst, int gamma) {
	g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_TC, gamma ? 1 : 0, 0, 0, dst }); }
void rglo_gl_store_depth(rglo_gpu* g, float* dst) { g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_DEPTH, 0, 0, 0, dst }); }
void rglo_gl_store_primitive_id(rglo_gpu* g, uint32_t* dst) { g->MaybeUpdateState(); g->commands.push_back({ Cmd::STORE_PRIM, 0, 0, 0, dst }); }
void rglo_gl_store_color_fp(rglo_gpu* g, float* dst, int stride, int downsample) {
	g->MaybeUpdateState();
	g->commands.push_back({ Cmd::STORE_FP, downsample ? 1 : 0, stride > 0 ? stride : downsample ? (g->W + 1) / 2 : g->W, 0, dst }); }
void rglo_gl_store_color_quads(rglo_gpu* g, float* dst, int stride) {
	g->MaybeUpdateState();
	g->commands.push_back({ Cmd::STORE_FP, 2, stride > 0 ? stride : (g->W + 1) / 2, 0, dst }); }
void rglo_gl_reset(rglo_gpu* g) { g->ResetGL(); }

// rglr::KawaseBlurFilter (rglr_kawase.cxx:21-85) over rows [0, height):
// clamped sampling near the edges, direct in the interior, as the reference
// splits it; RGBA float pixels, strides in pixels
void rglo_kawase_blur(const float* src, float* dst, int width, int height, int srcStride, int dstStride, int d) {
	const int ox[4] = { -d - 1, d, -d - 1, d }, oy[4] = { -d - 1, -d - 1, d, d };
	auto at = [&](int x, int y) { return src + 4 * (static_cast<long long>(y) * srcStride + x); };
	auto clampi = [](int a, int l, int h) { return std::min(std::max(a, l), h); };
	const int margin = d + 2;
	const int ySafeBegin = margin, ySafeEnd = height - margin, xSafeBegin = margin, xSafeEnd = width - margin;
	for (int y = 0; y < height; ++y) {
		for (int x = 0; x < width; ++x) {
			const bool safe = y < ySafeBegin || ySafeEnd < y || x < xSafeBegin || x >= xSafeEnd;
			float ax[4] = { 0.0F, 0.0F, 0.0F, 0.0F };
			for (int k = 0; k < 4; ++k) {
				const float* t[4];
				const int x0 = x + ox[k], y0 = y + oy[k];
				if (safe) {
					t[0] = at(clampi(x0, 0, width - 1), clampi(y0, 0, height - 1));
					t[1] = at(clampi(x0 + 1, 0, width - 1), clampi(y0, 0, height - 1));
					t[2] = at(clampi(x0, 0, width - 1), clampi(y0 + 1, 0, height - 1));
					t[3] = at(clampi(x0 + 1, 0, width - 1), clampi(y0 + 1, 0, height - 1)); }
				else {
					t[0] = at(x0, y0); t[1] = at(x0 + 1, y0); t[2] = at(x0, y0 + 1); t[3] = at(x0 + 1, y0 + 1); }
				for (int c = 0; c < 4; ++c) ax[c] = ax[c] + (((t[0][c] + t[1][c]) + t[2][c]) + t[3][c]); }
			float* o = dst + 4 * (static_cast<long long>(y) * dstStride + x);
			for (int c = 0; c < 4; ++c) o[c] = ax[c] * (1.0F / 16.0F); }}}

int rglo_gpu_tile_costs(rglo_gpu* g, long long* pairs, long long* quads, int n) {
	const int m = std::min(n, static_cast<int>(g->tileQuads.size()));
	for (int i = 0; i < m; ++i) { pairs[i] = g->tilePairs[i]; quads[i] = g->tileQuads[i]; }
	return m; }

// building blocks exposed for known-answer tests
float rglo_oneover(float x) { return oneover(x); }
float rglo_rcpps(float x) { return rcpps(x); }
int32_t rglo_cvtt(float x) { return cvtt(x); }
uint32_t rglo_to_srgb8(float x) { return to_srgb8(x); }
uint32_t rglo_to_linear8(float x) { return to_linear8(x); }
void rglo_mat4_mul(const float* l, const float* r, float* out) {
	mat4 a, b; std::memcpy(a.ff, l, 64); std::memcpy(b.ff, r, 64);
	const mat4 m = mul(a, b); std::memcpy(out, m.ff, 64); }
void rglo_xform4(const float* m, const float* v, float* out) { xform4(m, v, out); }

// the reference triangle test harness (rglv_triangle.t.cxx:35-131, 251-377):
// draws one triangle given in device coordinates with TriangleRasterizer into a
// w x h target, LEQUAL against depth 1.0, and returns per pixel
// [covered, BS.x, BS.y, BS.z, u, v] where (u, v) are the BP-interpolated
// vertex uvs.  Uncovered pixels are all zero.
int rglo_scalar_triangle(int w, int h, const float* pts /*3 x vec4*/, const float* uvs /*3 x vec2*/, float* out) {
	struct Capture : Program {
		Capture() { id = 1000; nvary = 2; earlyZ = true; }
		void ShadeVertex(const DrawEnv&, int, int, float*, VOut&) const override {}
		void ShadeFragment(const DrawEnv&, const float bs[3][4], const float[3][4], const float vary[kMaxVary][4],
		                   const float[2][4], const float[4], float color[4][4], int32_t[4]) const override {
			for (int l = 0; l < 4; ++l) { color[0][l] = bs[0][l]; color[1][l] = bs[1][l]; color[2][l] = bs[2][l];
			                              color[3][l] = vary[0][l]; }
			(void)vary; }};
	static const Capture cap;
	rglo_gpu g; g.SetSize(w, h, 1, 1);
	g.tileW = w; g.tileH = h;
	rglo_gpu::TileBuf tb;
	const size_t area = static_cast<size_t>(w) * h;
	tb.r.assign(area, 0.0F); tb.g.assign(area, 0.0F); tb.b.assign(area, 0.0F); tb.d.assign(area, 1.0F); tb.prim.assign(area, 0xffffffffu);
	std::vector<float> vv(area, 0.0F);
	DrawEnv env{};
	rglo_gpu::TriCtx tc{ &cap, &env, {}, {}, {}, 0u, GL_LEQUAL, true, true, true, false };
	int32_t X[3], Y[3];
	for (int k = 0; k < 3; ++k) {
		X[k] = static_cast<int32_t>(16.0F * pts[k*4 + 0]);
		Y[k] = static_cast<int32_t>(16.0F * pts[k*4 + 1]);
		tc.w[k] = pts[k*4 + 3]; tc.z[k] = pts[k*4 + 2];
		tc.vo[k].f[0] = uvs[k*2 + 0]; tc.vo[k].f[1] = uvs[k*2 + 1]; }
	const Rect r{ 0, 0, w, h };
	// pass 1: BS and u; pass 2: v (the capture program has three colour channels)
	g.RasterScalar(tb, r, r, false, tc, X, Y);
	std::vector<float> bsx = tb.r, bsy = tb.g, bsz = tb.b;
	std::vector<uint32_t> hit = tb.prim;
	struct CaptureUV : Program {
		CaptureUV() { id = 1001; nvary = 2; earlyZ = true; }
		void ShadeVertex(const DrawEnv&, int, int, float*, VOut&) const override {}
		void ShadeFragment(const DrawEnv&, const float[3][4], const float[3][4], const float vary[kMaxVary][4],
		                   const float[2][4], const float[4], float color[4][4], int32_t[4]) const override {
			for (int l = 0; l < 4; ++l) { color[0][l] = vary[0][l]; color[1][l] = vary[1][l]; color[2][l] = 0.0F; color[3][l] = 0.0F; }}};
	static const CaptureUV capuv;
	tb.r.assign(area, 0.0F); tb.g.assign(area, 0.0F); tb.b.assign(area, 0.0F); tb.d.assign(area, 1.0F); tb.prim.assign(area, 0xffffffffu);
	tc.pgm = &capuv;
	g.RasterScalar(tb, r, r, false, tc, X, Y);
	for (size_t p = 0; p < area; ++p) {
		const bool on = hit[p] != 0xffffffffu;
		out[p*6 + 0] = on ? 1.0F : 0.0F;
		out[p*6 + 1] = on ? bsx[p] : 0.0F;
		out[p*6 + 2] = on ? bsy[p] : 0.0F;
		out[p*6 + 3] = on ? bsz[p] : 0.0F;
		out[p*6 + 4] = on ? tb.r[p] : 0.0F;
		out[p*6 + 5] = on ? tb.g[p] : 0.0F; }
	return 0; }

// ---- building blocks for the pinning tests -------------------------------
// One triangle given by its fixed-point device vertices (1/16 pixel) drawn by
// the vector (mode 0, VTriangleRasterizer) or scalar (mode 1,
// TriangleRasterizer) rasteriser into the pixel rectangle [x0, x1) x [y0, y1)
// of a target h pixels high, depth test off; out[4 per pixel of the rect]
// = covered, BS.x, BS.y, BS.z.  Quad lanes past the rect's right or bottom
// edge (odd rects) are not reported.  x0 and y0 must be even, as every
// tile origin is.
int rglo_raster_tri(int mode, int x0, int y0, int x1, int y1, int h, const int32_t* X, const int32_t* Y, float* out) {
	struct CaptureBS : Program {
		CaptureBS() { id = 1002; nvary = 0; earlyZ = true; }
		void ShadeVertex(const DrawEnv&, int, int, float*, VOut&) const override {}
		void ShadeFragment(const DrawEnv&, const float bs[3][4], const float[3][4], const float[kMaxVary][4],
		                   const float[2][4], const float[4], float color[4][4], int32_t[4]) const override {
			for (int l = 0; l < 4; ++l) { color[0][l] = bs[0][l]; color[1][l] = bs[1][l]; color[2][l] = bs[2][l]; color[3][l] = 1.0F; }}};
	static const CaptureBS cap;
	if (x1 <= x0 || y1 <= y0 || (x0 & 1) || (y0 & 1)) return -1;   // tile origins are even
	const int rw = x1 - x0, rh = y1 - y0;
	rglo_gpu g; g.W = x1 + 2; g.H = h;
	g.tileW = rw + 2; g.tileH = rh + 2;   // the buffer stride: room for a quad's second lane past the rect
	rglo_gpu::TileBuf tb;
	const size_t area = static_cast<size_t>(g.tileW) * g.tileH;
	tb.r.assign(area, 0.0F); tb.g.assign(area, 0.0F); tb.b.assign(area, 0.0F); tb.d.assign(area, 1.0F); tb.prim.assign(area, 0xffffffffu);
	DrawEnv env{};
	rglo_gpu::TriCtx tc{ &cap, &env, {}, {}, {}, 0u, GL_LESS, false, false, true, false };
	for (int k = 0; k < 3; ++k) { tc.w[k] = 1.0F; tc.z[k] = 0.5F; }
	const Rect r{ x0, y0, x1, y1 };
	if (mode == 0) g.RasterVector(tb, r, r, false, tc, X, Y);
	else g.RasterScalar(tb, r, r, false, tc, X, Y);
	for (int y = 0; y < rh; ++y)
		for (int x = 0; x < rw; ++x) {
			const size_t p = static_cast<size_t>(y) * g.tileW + x;
			float* o = out + (static_cast<size_t>(y) * rw + x) * 4;
			const bool on = tb.prim[p] != 0xffffffffu;
			o[0] = on ? 1.0F : 0.0F; o[1] = on ? tb.r[p] : 0.0F; o[2] = on ? tb.g[p] : 0.0F; o[3] = on ? tb.b[p] : 0.0F; }
	return 0; }

// ClipTriangles' polygon clip of one triangle in clip coordinates with nv
// varyings per vertex: returns the clipped vertex count (0: nothing left),
// coordinates to out_coord (<= 8 x 4) and varyings to out_vary (<= 8 x nv).
int rglo_clip(const float* coord, const float* vary, int nv, float* out_coord, float* out_vary) {
	if (nv < 0 || nv > kMaxVary) return -1;
	std::vector<ClippedVertex> A(3);
	for (int k = 0; k < 3; ++k) {
		for (int q = 0; q < 4; ++q) A[k].coord[q] = coord[k * 4 + q];
		for (int j = 0; j < nv; ++j) A[k].data.f[j] = vary[k * nv + j]; }
	clip_polygon(A, nv);
	for (size_t k = 0; k < A.size(); ++k) {
		for (int q = 0; q < 4; ++q) out_coord[k * 4 + q] = A[k].coord[q];
		for (int j = 0; j < nv; ++j) out_vary[k * nv + j] = A[k].data.f[j]; }
	return static_cast<int>(A.size()); }

// TextureUnit::sample of one quad (4 lanes) on a float RGBA texture bound
// like GL::BindTexture: out[channel * 4 + lane]; returns -1 for a
// power-of-two size the reference cannot make a unit for.
int rglo_sample(const float* buf, int w, int h, int stride, int filter, const float* u, const float* v, float* out) {
	std::string err;
	const Sampler s = Sampler::Make(TextureState{ buf, w, h, stride, filter }, err);
	if (!err.empty()) return -1;
	float o[4][4];
	s.sample(u, v, o);
	for (int c = 0; c < 4; ++c) for (int l = 0; l < 4; ++l) out[c * 4 + l] = o[c][l];
	return s.kind; }

// LevelOfDetail from lanes 0 and 1 of base-level texel coordinates
// the rsqrtps model (rmlv_soa.hxx:243-248 normalize)
float rglo_rsqrtps(float x) { return rsqrtps(x); }

// the IQ post program's building blocks
float rglo_sse_pow(float x, float y) { return sse_pow(x, y); }
float rglo_approx_sin(float x) { return approx_sin(x); }
float rglo_fract(float x) { return fract_t(x); }
void rglo_iq_post(const float* in, float qx, float qy, float* out) {
	float c[3] = { in[0], in[1], in[2] };
	iq_post(c, qx, qy);
	for (int k = 0; k < 3; ++k) out[k] = c[k]; }

int rglo_level_of_detail(const float* u, const float* v) { return level_of_detail(u, v); }

// ShadeFragment of program pid for one quad: vary[k * 4 + lane] (k < the
// program's varyings), uniforms = 32 floats, buffers = 16 pointers (slots
// 11-14 for the heavy shader), texture unit 0 as GL::BindTexture; colour
// to out[channel * 4 + lane].
int rglo_shade(int pid, const float* uniforms, const float* const* buffers, const float* tex, int texDim, int filter,
               const float* vary, float* out) {
	const Program* pgm = find_program(pid);
	if (!pgm) return -1;
	GLState st; st.reset();
	for (int i = 0; i < 16; ++i) st.buffers[i] = buffers ? buffers[i] : nullptr;
	st.tus[0] = { tex, texDim, texDim, texDim, filter };
	DrawEnv env{};
	env.st = &st; env.uniforms = uniforms;
	std::string err;
	env.tu0 = Sampler::Make(st.tus[0], err);
	if (!err.empty()) return -1;
	float bs[3][4] = {}, bp[3][4] = {}, fc[2][4] = {}, fd[4] = {}, vv[kMaxVary][4] = {}, color[4][4] = {};
	for (int k = 0; k < pgm->nvary; ++k) for (int l = 0; l < 4; ++l) vv[k][l] = vary[k * 4 + l];
	int32_t mask[4] = { -1, -1, -1, -1 };
	pgm->ShadeFragment(env, bs, bp, vv, fc, fd, color, mask);
	for (int c = 0; c < 4; ++c) for (int l = 0; l < 4; ++l) out[c * 4 + l] = color[c][l];
	return 0; }

// the draw environment of the two calls below: view matrix vm (column-major,
// null = identity), the normal matrix from it as MakeMatrices forms it, the
// depth texture unit (tu3, dim)
static bool shade_env(int pid, const float* uniforms, const float* const* buffers, const float* tex, int texDim, int filter,
                      const float* vm, const float* tu3, int tu3dim, GLState& st, DrawEnv& env, const Program*& pgm) {
	pgm = find_program(pid);
	if (!pgm) return false;
	st.reset();
	for (int i = 0; i < 16; ++i) st.buffers[i] = buffers ? buffers[i] : nullptr;
	st.tus[0] = { tex, texDim, texDim, texDim, filter };
	env = DrawEnv{};
	static const float kNoUniforms[UNIFORM_BUFFER_SIZE] = {};
	env.st = &st; env.uniforms = uniforms ? uniforms : kNoUniforms;
	for (int i = 0; i < 16; ++i) env.vm.ff[i] = vm ? vm[i] : ((i % 5 == 0) ? 1.0F : 0.0F);
	mat4 inv;
	mat4_inverse(env.vm.ff, inv.ff);
	mat4_transpose(inv.ff, env.nm.ff);
	env.tu3.buf = tu3; env.tu3.dim = tu3dim;
	std::string err;
	if (tex) env.tu0 = Sampler::Make(st.tus[0], err);
	return err.empty(); }

// ShadeFragment with gl_FragCoord (fc[axis * 4 + lane], may be null), the
// view/normal matrices and the depth unit; colour to out[channel * 4 +
// lane], gl_triMask after the shader to mask_out[lane]
int rglo_shade_ex(int pid, const float* uniforms, const float* const* buffers, const float* tex, int texDim, int filter,
                  const float* vm, const float* tu3, int tu3dim, const float* fcIn, const float* vary, float* out,
                  int32_t* mask_out) {
	GLState st; DrawEnv env; const Program* pgm;
	if (!shade_env(pid, uniforms, buffers, tex, texDim, filter, vm, tu3, tu3dim, st, env, pgm)) return -1;
	float bs[3][4] = {}, bp[3][4] = {}, fc[2][4] = {}, fd[4] = {}, vv[kMaxVary][4] = {}, color[4][4] = {};
	for (int k = 0; k < pgm->nvary; ++k) for (int l = 0; l < 4; ++l) vv[k][l] = vary[k * 4 + l];
	if (fcIn) for (int a = 0; a < 2; ++a) for (int l = 0; l < 4; ++l) fc[a][l] = fcIn[a * 4 + l];
	int32_t mask[4] = { -1, -1, -1, -1 };
	pgm->ShadeFragment(env, bs, bp, vv, fc, fd, color, mask);
	for (int c = 0; c < 4; ++c) for (int l = 0; l < 4; ++l) out[c * 4 + l] = color[c][l];
	for (int l = 0; l < 4; ++l) mask_out[l] = mask[l];
	return pgm->nvary; }

// ShadeVertex of vertex idx (instance iid): clip position to pos[4],
// varyings to vary[kMaxVary]; vpm = projection * view (column-major)
int rglo_shade_vertex(int pid, const float* uniforms, const float* const* buffers, const float* vpm, const float* vm,
                      int iid, int idx, float* pos, float* vary) {
	GLState st; DrawEnv env; const Program* pgm;
	if (!shade_env(pid, uniforms, buffers, nullptr, 0, 0, vm, nullptr, 0, st, env, pgm)) return -1;
	for (int i = 0; i < 16; ++i) env.vpm.ff[i] = vpm[i];
	VOut o{};
	pgm->ShadeVertex(env, iid, idx, pos, o);
	for (int k = 0; k < kMaxVary; ++k) vary[k] = o.f[k];
	return pgm->nvary; }

}  // extern "C"
