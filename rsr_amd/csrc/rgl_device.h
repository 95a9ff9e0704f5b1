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
#include "srgb8_table.inc"

constexpr int kMaxDraws = 256;      // draws per batch
constexpr int kMaxFrames = 64;      // frames (render targets) per batch
constexpr int kMaxVary = 12;       // varyings per vertex (largest program, padded to 16 bytes)
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

// ---------------------------------------------------------------------------
// shader programs, in the reference's program concept
// (rglv_gpu_shaders.hxx:14-95): ShadeVertex per vertex; ShadeFragment
// evaluated here one lane at a time (lanes of a quad are independent except
// for the texture LOD, which the caller computes from lanes 0 and 1 of the
// quad); varyings are interpolated with the perspective-correct BP.
// kUV is the index of the first texture-coordinate varying (-1: none).
// ---------------------------------------------------------------------------
__device__ inline float buf_at(const float* p, int idx, float dflt) { return p ? p[idx] : dflt; }

__device__ inline void load_pos(const DrawDesc& d, int idx, float p[4]) {
	p[0] = buf_at(d.buffers[0], idx, 0.0F); p[1] = buf_at(d.buffers[1], idx, 0.0F);
	p[2] = buf_at(d.buffers[2], idx, 0.0F); p[3] = 1.0F; }

// AmyProgram (viewer/shaders.hxx:69-161)
struct ProgAmy {
	static constexpr int kId = 4, kVary = 2, kUV = 0;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		xform4(d.vpm, p, pos);
		vo[0] = buf_at(d.buffers[9], idx, 0.0F);
		vo[1] = buf_at(d.buffers[10], idx, 0.0F); }
	__device__ static void fragment1(const DrawDesc& d, const float* v, int lod, float color[4]) {
		const float4 t = sample1(d.tu0, v[0], v[1], lod);
		color[0] = t.x; color[1] = t.y; color[2] = t.z; color[3] = t.w; }};

// ManyProgram (viewer/shaders.hxx:482-583)
struct ProgMany {
	static constexpr int kId = 6, kVary = 2, kUV = -1;
	__device__ static void vertex(const DrawDesc& d, int iid, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		const float* imat = d.buffers[15] + 16 * iid;
		float m[16];
#pragma unroll
		for (int k = 0; k < 16; ++k) m[k] = imat[k];
		float p1[4]; xform4(m, p, p1);
		xform4(d.vpm, p1, pos);
		vo[0] = buf_at(d.buffers[9], idx, 0.0F);
		vo[1] = buf_at(d.buffers[10], idx, 0.0F); }
	__device__ static void fragment1(const DrawDesc& d, const float* v, int, float color[4]) {
		color[0] = v[0]; color[1] = v[1]; color[2] = d.uniforms[0]; color[3] = 1.0F; }};

// LambertTex (this project): textured, one directional light.
// uniforms: [0..2] unit direction toward the light, [4] ambient, [5] diffuse
struct ProgLambertTex {
	static constexpr int kId = 80, kVary = 5, kUV = 3;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		xform4(d.vpm, p, pos);
		vo[0] = buf_at(d.buffers[3], idx, 0.0F); vo[1] = buf_at(d.buffers[4], idx, 0.0F);
		vo[2] = buf_at(d.buffers[5], idx, 1.0F);
		vo[3] = buf_at(d.buffers[9], idx, 0.0F); vo[4] = buf_at(d.buffers[10], idx, 0.0F); }
	__device__ static void fragment1(const DrawDesc& d, const float* v, int lod, float color[4]) {
		const float* u = d.uniforms;
		const float4 tex = sample1(d.tu0, v[3], v[4], lod);
		color[3] = 1.0F;
		float nx = v[0], ny = v[1], nz = v[2];
		normalize3(nx, ny, nz);
		const float ndl = sse_max((nx*u[0] + ny*u[1]) + nz*u[2], 0.0F);
		const float k = u[4] + u[5] * ndl;
		color[0] = tex.x * k; color[1] = tex.y * k; color[2] = tex.z * k; }};

// ManyBlinnPhong (this project): instanced, textured, one point light.
// uniforms: [0..2] eye, [4..6] light position, [8] ka, [9] kd, [10] ks
struct ProgManyBlinnPhong {
	static constexpr int kId = 81, kVary = 8, kUV = 6;
	__device__ static void vertex(const DrawDesc& d, int iid, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		const float* imat = d.buffers[15] + 16 * iid;
		float m[16];
#pragma unroll
		for (int k = 0; k < 16; ++k) m[k] = imat[k];
		float p1[4]; xform4(m, p, p1);
		xform4(d.vpm, p1, pos);
		const float n[3] = { buf_at(d.buffers[3], idx, 0.0F), buf_at(d.buffers[4], idx, 0.0F), buf_at(d.buffers[5], idx, 1.0F) };
		float wn[3]; xform3w0(m, n, wn);
		vo[0] = p1[0]; vo[1] = p1[1]; vo[2] = p1[2];
		vo[3] = wn[0]; vo[4] = wn[1]; vo[5] = wn[2];
		vo[6] = buf_at(d.buffers[9], idx, 0.0F); vo[7] = buf_at(d.buffers[10], idx, 0.0F); }
	__device__ static void fragment1(const DrawDesc& d, const float* v, int lod, float color[4]) {
		const float* u = d.uniforms;
		const float4 tex = sample1(d.tu0, v[6], v[7], lod);
		color[3] = 1.0F;
		const float px = v[0], py = v[1], pz = v[2];
		float nx = v[3], ny = v[4], nz = v[5];
		normalize3(nx, ny, nz);
		float vx = u[0] - px, vy = u[1] - py, vz = u[2] - pz;
		normalize3(vx, vy, vz);
		float lx = u[4] - px, ly = u[5] - py, lz = u[6] - pz;
		normalize3(lx, ly, lz);
		float hx = lx + vx, hy = ly + vy, hz = lz + vz;
		normalize3(hx, hy, hz);
		const float ndl = sse_max((nx*lx + ny*ly) + nz*lz, 0.0F);
		const float ndh = sse_max((nx*hx + ny*hy) + nz*hz, 0.0F);
		float s = ndh * ndh; s = s * s; s = s * s; s = s * s; s = s * s;
		const float kdiff = u[8] + u[9] * ndl;
		const float kspec = u[10] * s;
		color[0] = tex.x * kdiff + kspec;
		color[1] = tex.y * kdiff + kspec;
		color[2] = tex.z * kdiff + kspec; }};

// GGXEnvProgram (new, the heavy shader): instanced and textured like
// ManyBlinnPhong; 16 point lights with Cook-Torrance GGX specular (Smith-
// Schlick G, Schlick F, F0 = 0.04) and per-instance roughness, plus an
// 8-tap environment lookup around the reflection vector.  Only + - * /
// and sqrt, in a fixed order, so the CPU oracle matches it bit for bit.
//   buffers: 11 environment cube (6 faces of kEnvDim^2 RGBA floats),
//            12 taps (8 x {dx, dy, dz, weight}), 13 lights (16 x {x, y, z,
//            0, r, g, b, 0}), 14 per-instance roughness, 15 instance mats
//   uniforms: [0..2] eye, [3] ambient, [4] env specular, [5] env diffuse
constexpr int kGGXLights = 16, kEnvTaps = 8, kEnvDim = 16;
constexpr float kPi = 3.14159265F, kInvPi = 0.318309886F;

// the shading inputs of one fragment: unit normal and view vector, world
// position, albedo, roughness (also the learned shader's 13 inputs)
struct ShadeInput { float n[3], v[3], p[3], albedo[3], rough; };

__device__ inline void env_sample(const float* env, float dx, float dy, float dz, float out[3]) {
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
	fs = fs >= 0.0F ? fs : 0.0F; fs = fs <= E - 1.0F ? fs : E - 1.0F;   // NaN -> 0
	ft = ft >= 0.0F ? ft : 0.0F; ft = ft <= E - 1.0F ? ft : E - 1.0F;
	const int s0 = static_cast<int>(fs), t0 = static_cast<int>(ft);
	const int s1 = s0 + 1 < kEnvDim ? s0 + 1 : kEnvDim - 1, t1 = t0 + 1 < kEnvDim ? t0 + 1 : kEnvDim - 1;
	const float fx = fs - static_cast<float>(s0), fy = ft - static_cast<float>(t0);
	const float fx1 = 1.0F - fx, fy1 = 1.0F - fy;
	const float w00 = fx1 * fy1, w10 = fx * fy1, w01 = fx1 * fy, w11 = fx * fy;
	const float* f0 = env + static_cast<long long>(face) * kEnvDim * kEnvDim * 4;
	const float* p00 = f0 + (t0 * kEnvDim + s0) * 4;
	const float* p10 = f0 + (t0 * kEnvDim + s1) * 4;
	const float* p01 = f0 + (t1 * kEnvDim + s0) * 4;
	const float* p11 = f0 + (t1 * kEnvDim + s1) * 4;
#pragma unroll
	for (int c = 0; c < 3; ++c) out[c] = ((p00[c] * w00 + p10[c] * w10) + p01[c] * w01) + p11[c] * w11; }

__device__ inline void ggx_shade(const DrawDesc& d, const ShadeInput& s, float color[3]) {
	const float* u = d.uniforms;
	const float* lights = d.buffers[13];
	const float* taps = d.buffers[12];
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
		const float inv = 1.0F / __builtin_sqrtf(d2);
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
#pragma unroll
		for (int c = 0; c < 3; ++c) acc[c] = acc[c] + ((s.albedo[c] * kd + spec) * (L[4 + c] * id2)) * ndl; }
	// environment: 8 taps spread around the reflection vector by roughness
	const float rx = (2.0F * ndvRaw) * s.n[0] - s.v[0];
	const float ry = (2.0F * ndvRaw) * s.n[1] - s.v[1];
	const float rz = (2.0F * ndvRaw) * s.n[2] - s.v[2];
	float e[3] = { 0.0F, 0.0F, 0.0F };
	for (int j = 0; taps != nullptr && j < kEnvTaps; ++j) {
		const float* T = taps + 4 * j;
		float smp[3];
		env_sample(d.buffers[11], rx + T[0] * rough, ry + T[1] * rough, rz + T[2] * rough, smp);
#pragma unroll
		for (int c = 0; c < 3; ++c) e[c] = e[c] + T[3] * smp[c]; }
	const float fe = 1.0F - ndv, fe2 = fe * fe, fe5 = (fe2 * fe2) * fe;
	const float Fe = 0.04F + 0.96F * fe5;
	const float ks = (u[4] * Fe) * (1.0F - rough);
#pragma unroll
	for (int c = 0; c < 3; ++c) color[c] = (acc[c] + u[3] * s.albedo[c]) + e[c] * (ks + u[5] * s.albedo[c]); }

struct ProgGGXEnv {
	static constexpr int kId = 82, kVary = 9, kUV = 6;
	__device__ static void vertex(const DrawDesc& d, int iid, int idx, float pos[4], float* vo) {
		ProgManyBlinnPhong::vertex(d, iid, idx, pos, vo);
		vo[8] = d.buffers[14] ? d.buffers[14][iid] : 0.5F; }
	// varyings -> shading inputs (the G-buffer of the learned shader)
	__device__ static void inputs(const DrawDesc& d, const float* v, int lod, ShadeInput& s) {
		const float4 tex = sample1(d.tu0, v[6], v[7], lod);
		s.albedo[0] = tex.x; s.albedo[1] = tex.y; s.albedo[2] = tex.z;
		s.p[0] = v[0]; s.p[1] = v[1]; s.p[2] = v[2];
		float nx = v[3], ny = v[4], nz = v[5];
		normalize3(nx, ny, nz);
		s.n[0] = nx; s.n[1] = ny; s.n[2] = nz;
		const float* u = d.uniforms;
		float vx = u[0] - v[0], vy = u[1] - v[1], vz = u[2] - v[2];
		normalize3(vx, vy, vz);
		s.v[0] = vx; s.v[1] = vy; s.v[2] = vz;
		s.rough = v[8]; }
	__device__ static void fragment1(const DrawDesc& d, const float* v, int lod, float color[4]) {
		ShadeInput s;
		inputs(d, v, lod, s);
		ggx_shade(d, s, color);
		color[3] = 1.0F; }};

__host__ __device__ inline int program_nvary(int pid) {
	switch (pid) {
	case ProgAmy::kId: return ProgAmy::kVary;
	case ProgMany::kId: return ProgMany::kVary;
	case ProgLambertTex::kId: return ProgLambertTex::kVary;
	case ProgManyBlinnPhong::kId: return ProgManyBlinnPhong::kVary;
	case ProgGGXEnv::kId: return ProgGGXEnv::kVary;
	default: return -1; }}

__host__ __device__ inline bool program_instanced(int pid) {
	return pid == ProgMany::kId || pid == ProgManyBlinnPhong::kId || pid == ProgGGXEnv::kId; }

}  // namespace rgl
