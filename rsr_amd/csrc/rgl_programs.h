// rgl_programs.h — shader programs of the gfx950 rgl pipeline, in the
// reference's program concept (rglv_gpu_shaders.hxx:14-95).
//
// A program is a struct with
//   kId        the program id (GL::UseProgram, GPU::Install)
//   kVary      varyings per vertex (<= kMaxVary)
//   kUV        the varying holding the texture coordinate whose lane-0/1
//              difference sets the quad's texture LOD (rglr_texture_sampler
//              LevelOfDetail), -1 for none, kUVFromFragCoord for a program
//              that samples at a coordinate made from gl_FragCoord (it then
//              defines texcoord())
//   kEarlyZ    the depth test runs before the fragment shader (the
//              reference's SHADER::earlyZ)
//   kDiscard   the fragment shader may clear lanes of gl_triMask
//   kInstanced the vertex shader reads the instance matrices (slot 15)
//   vertex(d, iid, idx, pos, vo)          ShadeVertex of vertex idx of
//                                         instance iid: clip position, varyings
//   fragment(d, v, lod, fc, color) -> bool ShadeFragment of one lane:
//                                         RGBA colour; false = discarded
// Every float operation is written in the reference's order and compiled
// without contraction, so the oracle's restatement matches bit for bit;
// rcpps and rsqrtps are emulated from the captured tables.
//
// The programs form one compile-time list, AllPrograms, from which every
// per-program launch and dispatch is generated.  A caller adds programs
// without editing this file: a header that defines
//     using UserPrograms = rgl::ProgramList<MyProgram, ...>;
// (programs written to the struct concept above) is named at build time
// (python -m rsr_amd.build --programs path/to/header.h, which passes
// -DRGL_USER_PROGRAMS="path").
#pragma once

#include "rgl_device.h"

namespace rgl {

constexpr int kUVFromFragCoord = 1000;

// gl_FragCoord (lower-left origin) and gl_FragDepth of one lane
struct FragCoord { float x, y, depth; };

__device__ inline float buf_at(const float* p, int idx, float dflt) { return p ? p[idx] : dflt; }

__device__ inline void load_pos(const DrawDesc& d, int idx, float p[4]) {
	p[0] = buf_at(d.buffers[0], idx, 0.0F); p[1] = buf_at(d.buffers[1], idx, 0.0F);
	p[2] = buf_at(d.buffers[2], idx, 0.0F); p[3] = 1.0F; }

// rmlv::normalize of a qfloat3 / qfloat4 lane: scaled by rsqrtps of the dot
// product (rmlv_soa.hxx:243-248)
__device__ inline void normalize3_rsqrt(float v[3]) {
	const float s = rsqrtps((v[0]*v[0] + v[1]*v[1]) + v[2]*v[2]);
	v[0] = v[0] * s; v[1] = v[1] * s; v[2] = v[2] * s; }
__device__ inline float dot4(const float a[4], const float b[4]) {
	return ((a[0]*b[0] + a[1]*b[1]) + a[2]*b[2]) + a[3]*b[3]; }

// DepthTextureUnit::sample (rglr_texture_sampler.hxx:70-79): nearest, the
// border value -1 outside [0, 1)^2, the offset wrapped by dim^2 - 1
__device__ inline float sample_depth(const float* buf, int dim, float x, float y) {
	const bool hit = x >= 0.0F && x < 1.0F && y >= 0.0F && y < 1.0F;
	const float dimf = static_cast<float>(dim);
	const int32_t px = cvtt(x * dimf);
	const int32_t py = cvtt((1.0F - y) * dimf);
	const int32_t ofs = wadd(wmul(py, dim), px) & wsub(wmul(dim, dim), 1);
	const float c = buf ? buf[ofs] : 0.0F;
	return hit ? c : -1.0F; }

// rglv::BaseProgram (rglv_gpu_shaders.hxx:21-95): the position alone, colour
// 1; the $layer shadow-map pass draws with it (node/gllayer.cxx:44-45).  It
// has no varyings; the one here is unused (the kernels' per-vertex arrays
// are sized by kVary)
struct ProgBase {
	static constexpr int kId = 0, kVary = 1, kUV = -1;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		xform4(d.vpm, p, pos);
		vo[0] = 0.0F; }
	__device__ static bool fragment(const DrawDesc&, const float*, int, const FragCoord&, float color[4]) {
		color[0] = 1.0F; color[1] = 1.0F; color[2] = 1.0F; color[3] = 1.0F;
		return true; }};

// AmyProgram (viewer/shaders.hxx:69-161)
struct ProgAmy {
	static constexpr int kId = 4, kVary = 2, kUV = 0;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		xform4(d.vpm, p, pos);
		vo[0] = buf_at(d.buffers[9], idx, 0.0F);
		vo[1] = buf_at(d.buffers[10], idx, 0.0F); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int lod, const FragCoord&, float color[4]) {
		const float4 t = sample1(d.tu0, v[0], v[1], lod);
		color[0] = t.x; color[1] = t.y; color[2] = t.z; color[3] = t.w;
		return true; }};

// ManyProgram (viewer/shaders.hxx:482-583)
struct ProgMany {
	static constexpr int kId = 6, kVary = 2, kUV = -1;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = true;
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
	__device__ static bool fragment(const DrawDesc& d, const float* v, int, const FragCoord&, float color[4]) {
		color[0] = v[0]; color[1] = v[1]; color[2] = d.uniforms[0]; color[3] = 1.0F;
		return true; }};

// LambertTex (this project): textured, one directional light.
// uniforms: [0..2] unit direction toward the light, [4] ambient, [5] diffuse
struct ProgLambertTex {
	static constexpr int kId = 80, kVary = 5, kUV = 3;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		xform4(d.vpm, p, pos);
		vo[0] = buf_at(d.buffers[3], idx, 0.0F); vo[1] = buf_at(d.buffers[4], idx, 0.0F);
		vo[2] = buf_at(d.buffers[5], idx, 1.0F);
		vo[3] = buf_at(d.buffers[9], idx, 0.0F); vo[4] = buf_at(d.buffers[10], idx, 0.0F); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int lod, const FragCoord&, float color[4]) {
		const float* u = d.uniforms;
		const float4 tex = sample1(d.tu0, v[3], v[4], lod);
		float nx = v[0], ny = v[1], nz = v[2];
		normalize3(nx, ny, nz);
		const float ndl = sse_max((nx*u[0] + ny*u[1]) + nz*u[2], 0.0F);
		const float k = u[4] + u[5] * ndl;
		color[0] = tex.x * k; color[1] = tex.y * k; color[2] = tex.z * k; color[3] = 1.0F;
		return true; }};

// ManyBlinnPhong (this project): instanced, textured, one point light.
// uniforms: [0..2] eye, [4..6] light position, [8] ka, [9] kd, [10] ks
struct ProgManyBlinnPhong {
	static constexpr int kId = 81, kVary = 8, kUV = 6;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = true;
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
	__device__ static bool fragment(const DrawDesc& d, const float* v, int lod, const FragCoord&, float color[4]) {
		const float* u = d.uniforms;
		const float4 tex = sample1(d.tu0, v[6], v[7], lod);
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
		color[2] = tex.z * kdiff + kspec;
		color[3] = 1.0F;
		return true; }};

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
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = true;
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
	__device__ static bool fragment(const DrawDesc& d, const float* v, int lod, const FragCoord&, float color[4]) {
		ShadeInput s;
		inputs(d, v, lod, s);
		ggx_shade(d, s, color);
		color[3] = 1.0F;
		return true; }};

// AlphaTextureProgram (viewer/shaders.hxx:164-230): textured, late depth
// test, lanes whose texel alpha is not > 0 discarded
struct ProgAlphaTexture {
	static constexpr int kId = 65, kVary = 2, kUV = 0;
	static constexpr bool kEarlyZ = false, kDiscard = true, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		xform4(d.vpm, p, pos);
		vo[0] = buf_at(d.buffers[9], idx, 0.0F);
		vo[1] = buf_at(d.buffers[10], idx, 0.0F); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int lod, const FragCoord&, float color[4]) {
		const float4 t = sample1(d.tu0, v[0], v[1], lod);
		color[0] = t.x; color[1] = t.y; color[2] = t.z; color[3] = t.w;
		return t.w > 0.0F; }};

// TextProgram (viewer/shaders.hxx:232-318): vertex colour (slots 6-8,
// default white) times the texture, late depth test, alpha discard
struct ProgText {
	static constexpr int kId = 26, kVary = 5, kUV = 3;
	static constexpr bool kEarlyZ = false, kDiscard = true, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		const bool c = d.buffers[6] != nullptr;
		vo[0] = c ? d.buffers[6][idx] : 1.0F; vo[1] = c ? d.buffers[7][idx] : 1.0F; vo[2] = c ? d.buffers[8][idx] : 1.0F;
		vo[3] = buf_at(d.buffers[9], idx, 0.0F);
		vo[4] = buf_at(d.buffers[10], idx, 0.0F);
		xform4(d.vpm, p, pos); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int lod, const FragCoord&, float color[4]) {
		const float4 t = sample1(d.tu0, v[3], v[4], lod);
		color[0] = t.x * v[0]; color[1] = t.y * v[1]; color[2] = t.z * v[2]; color[3] = t.w;
		return t.w > 0.0F; }};

// DepthProgram (viewer/shaders.hxx:321-420): the depth texture unit (tu3)
// at the vertex uv, linearised for zNear 10, zFar 1000
struct ProgDepth {
	static constexpr int kId = 5, kVary = 2, kUV = -1;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		vo[0] = buf_at(d.buffers[9], idx, 0.0F);
		vo[1] = buf_at(d.buffers[10], idx, 0.0F);
		xform4(d.vpm, p, pos); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int, const FragCoord&, float color[4]) {
		const float zNear = 10.0F, zFar = 1000.0F;
		float t = sample_depth(d.tu3, d.tu3dim, v[0], v[1]);
		t = (2.0F * zNear) / ((zFar + zNear) - t * (zFar - zNear));
		color[0] = t; color[1] = t; color[2] = t; color[3] = 1.0F;
		return true; }};

// PatternProgram (viewer/shaders.hxx:422-480): texture 0 at
// gl_FragCoord / dim.y + offset.xy; uniforms [0..3] offset, [4..7] dim
struct ProgPattern {
	static constexpr int kId = 41, kVary = 1, kUV = kUVFromFragCoord;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		vo[0] = 0.0F;
		xform4(d.vpm, p, pos); }
	__device__ static void texcoord(const DrawDesc& d, const FragCoord& fc, float& u, float& v) {
		u = fc.x / d.uniforms[5]; v = fc.y / d.uniforms[5];
		u = u + d.uniforms[0]; v = v + d.uniforms[1]; }
	__device__ static bool fragment(const DrawDesc& d, const float*, int lod, const FragCoord& fc, float color[4]) {
		float u, v;
		texcoord(d, fc, u, v);
		const float4 t = sample1(d.tu0, u, v, lod);
		color[0] = t.x; color[1] = t.y; color[2] = t.z; color[3] = t.w;
		return true; }};

// the point light at the eye of OBJ1/OBJ2 (viewer/shaders.hxx:660-668,
// 784-792): diffuse = max(n . l, 0.1) * 100 / distance, with l and the
// distance from lightPos - p as the reference's qfloat4 subtraction forms
// it: its w component is lightPos.w - p.z (rmlv_soa.hxx:163)
__device__ inline float obj_diffuse(const float p[4], const float n[4]) {
	const float lp[4] = { 0.0F, 0.0F, 0.0F, 1.0F };
	float l[4] = { lp[0] - p[0], lp[1] - p[1], lp[2] - p[2], lp[3] - p[2] };
	const float distance = __builtin_sqrtf(dot4(l, l));
	const float s = rsqrtps(dot4(l, l));
	l[0] = l[0] * s; l[1] = l[1] * s; l[2] = l[2] * s; l[3] = l[3] * s;
	float diffuse = sse_max(dot4(n, l), 0.1F);
	return diffuse * (100.0F / distance); }

__device__ inline void obj_kd(const DrawDesc& d, int idx, float kd[3]) {
	const bool c = d.buffers[6] != nullptr;
	kd[0] = c ? d.buffers[6][idx] : 1.0F; kd[1] = c ? d.buffers[7][idx] : 1.0F; kd[2] = c ? d.buffers[8][idx] : 1.0F; }

__device__ inline void obj_normal(const DrawDesc& d, int idx, float n[3]) {
	const bool c = d.buffers[3] != nullptr;
	n[0] = c ? d.buffers[3][idx] : 0.0F; n[1] = c ? d.buffers[4][idx] : 0.0F; n[2] = c ? d.buffers[5][idx] : 1.0F; }

// OBJ1Program (viewer/shaders.hxx:584-681): per-vertex diffuse
struct ProgOBJ1 {
	static constexpr int kId = 7, kVary = 3, kUV = -1;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		float nrm[3]; obj_normal(d, idx, nrm);
		float kd[3]; obj_kd(d, idx, kd);
		float mvv[4], mvn[4];
		xform4(d.vm, p, mvv);
		const float n4[4] = { nrm[0], nrm[1], nrm[2], 0.0F };
		xform4(d.vm, n4, mvn);
		const float diffuse = obj_diffuse(mvv, mvn);
		vo[0] = kd[0] * diffuse; vo[1] = kd[1] * diffuse; vo[2] = kd[2] * diffuse;
		xform4(d.vpm, p, pos); }
	__device__ static bool fragment(const DrawDesc&, const float* v, int, const FragCoord&, float color[4]) {
		color[0] = v[0]; color[1] = v[1]; color[2] = v[2]; color[3] = 1.0F;
		return true; }};

// OBJ2Program (viewer/shaders.hxx:701-818): per-fragment diffuse from the
// interpolated view-space position and normal
struct ProgOBJ2 {
	static constexpr int kId = 8, kVary = 11, kUV = -1;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		float nrm[3]; obj_normal(d, idx, nrm);
		xform4(d.vm, p, vo);
		const float n4[4] = { nrm[0], nrm[1], nrm[2], 0.0F };
		xform4(d.vm, n4, vo + 4);
		obj_kd(d, idx, vo + 8);
		xform4(d.vpm, p, pos); }
	__device__ static bool fragment(const DrawDesc&, const float* v, int, const FragCoord&, float color[4]) {
		const float diffuse = obj_diffuse(v, v + 4);
		color[0] = v[8] * diffuse; color[1] = v[9] * diffuse; color[2] = v[10] * diffuse; color[3] = 1.0F;
		return true; }};

// OBJ2SProgram (viewer/shaders.hxx:821-978): a spot light with a shadow
// map in the depth texture unit.  uniforms: [0..15] modelToShadow (mat4),
// [16..18] light position, [19..21] light direction, [22] cos of the cone
struct ProgOBJ2S {
	static constexpr int kId = 9, kVary = 15, kUV = -1;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		float nrm[3]; obj_normal(d, idx, nrm);
		xform4(d.vm, p, vo);                                   // sp
		const float n4[4] = { nrm[0], nrm[1], nrm[2], 1.0F };
		xform4(d.nm, n4, vo + 4);                              // sn
		obj_kd(d, idx, vo + 8);                                // kd
		xform4(d.uniforms, p, vo + 11);                        // lp = modelToShadow * p
		xform4(d.vpm, p, pos); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int, const FragCoord&, float color[4]) {
		const float* u = d.uniforms;
		float stl[3] = { u[16] - v[0], u[17] - v[1], u[18] - v[2] };
		const float dist = __builtin_sqrtf((stl[0]*stl[0] + stl[1]*stl[1]) + stl[2]*stl[2]);
		normalize3_rsqrt(stl);
		float attenuation = 100.0F / dist;
		float ld[3] = { u[19], u[20], u[21] };
		normalize3_rsqrt(ld);
		const float ang = ((-stl[0])*ld[0] + (-stl[1])*ld[1]) + (-stl[2])*ld[2];
		if (ang < u[22]) attenuation = 0.111F;
		float sn[3] = { v[4], v[5], v[6] };
		normalize3_rsqrt(sn);
		const float diffuse = sse_max(0.0F, (sn[0]*stl[0] + sn[1]*stl[1]) + sn[2]*stl[2]);
		const float lpx = v[11] / v[14], lpy = v[12] / v[14], lpz = v[13] / v[14];
		const float t = sample_depth(d.tu3, d.tu3dim, lpx * 0.5F + 0.5F, lpy * 0.5F + 0.5F);
		if (lpz - 0.0005F > t) attenuation = 0.111F;
		color[0] = (attenuation * v[8]) * diffuse; color[1] = (attenuation * v[9]) * diffuse;
		color[2] = (attenuation * v[10]) * diffuse; color[3] = 1.0F;
		return true; }};

// EnvmapProgram (viewer/shaders_envmap.hxx:26-127): a sphere-map lookup
// at the reflection of the view vector about the smooth normal (slots 3-5,
// default (0, 0, 1)); slots 6-8 hold the face normal (default (1, 1, 1)),
// mixed in with weight 0; alpha 0.5 for the blended pass
struct ProgEnvmap {
	static constexpr int kId = 10, kVary = 2, kUV = 0;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		float sn[3]; obj_normal(d, idx, sn);
		const bool c = d.buffers[6] != nullptr;
		const float fn[3] = { c ? d.buffers[6][idx] : 1.0F, c ? d.buffers[7][idx] : 1.0F, c ? d.buffers[8][idx] : 1.0F };
		float mv[4]; xform4(d.vm, p, mv);
		float e[3] = { mv[0], mv[1], mv[2] };
		normalize3_rsqrt(e);
		float vn[3];
#pragma unroll
		for (int k = 0; k < 3; ++k) vn[k] = sn[k] + 0.0F * (fn[k] - sn[k]);   // mix(a, b, 0) = a + t (b - a)
		float n[3]; xform3w0(d.nm, vn, n);
		normalize3_rsqrt(n);
		const float dn2 = 2.0F * ((n[0]*e[0] + n[1]*e[1]) + n[2]*e[2]);
		const float r[3] = { e[0] - dn2 * n[0], e[1] - dn2 * n[1], e[2] - dn2 * n[2] };   // reflect (rglv_math.hxx:45-46)
		const float rz1 = r[2] + 1.0F;
		const float m = 2.0F * __builtin_sqrtf(((r[0]*r[0]) + (r[1]*r[1])) + (rz1*rz1));
		vo[0] = r[0] / m + 0.5F;
		vo[1] = r[1] / m + 0.5F;
		xform4(d.vpm, p, pos); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int lod, const FragCoord&, float color[4]) {
		const float4 t = sample1(d.tu0, v[0], v[1], lod);
		color[0] = t.x; color[1] = t.y; color[2] = t.z; color[3] = 0.5F;
		return true; }};

// ---------------------------------------------------------------------------
// the program list and what is generated from it
// ---------------------------------------------------------------------------
template <class... Ps> struct ProgramList {};

template <class A, class B> struct ConcatPrograms;
template <class... As, class... Bs> struct ConcatPrograms<ProgramList<As...>, ProgramList<Bs...>> {
	using type = ProgramList<As..., Bs...>; };

using BuiltinPrograms = ProgramList<ProgBase, ProgAmy, ProgMany, ProgLambertTex, ProgManyBlinnPhong, ProgGGXEnv, ProgAlphaTexture,
                                    ProgText, ProgDepth, ProgPattern, ProgOBJ1, ProgOBJ2, ProgOBJ2S, ProgEnvmap>;

}  // namespace rgl

#ifdef RGL_USER_PROGRAMS
#include RGL_USER_PROGRAMS
namespace rgl { using AllPrograms = typename ConcatPrograms<BuiltinPrograms, UserPrograms>::type; }
#else
namespace rgl { using AllPrograms = BuiltinPrograms; }
#endif

namespace rgl {

// a program's properties for the host planner: -1 varyings = no such program
struct ProgramInfo { int nvary; bool instanced, earlyZ, discard; };

template <class... Ps>
__host__ __device__ inline ProgramInfo program_info_of(int pid, ProgramList<Ps...>) {
	ProgramInfo r{ -1, false, true, false };
	(void)((pid == Ps::kId ? (r = ProgramInfo{ Ps::kVary, Ps::kInstanced, Ps::kEarlyZ, Ps::kDiscard }, true) : false) || ...);
	return r; }

__host__ __device__ inline ProgramInfo program_info(int pid) { return program_info_of(pid, AllPrograms{}); }
__host__ __device__ inline int program_nvary(int pid) { return program_info(pid).nvary; }
__host__ __device__ inline bool program_instanced(int pid) { return program_info(pid).instanced; }

// calls f.template run<P>() for the program P with id pid; false if none
template <class F, class... Ps>
__host__ __device__ inline bool dispatch_program(int pid, F& f, ProgramList<Ps...>) {
	return ((pid == Ps::kId ? (f.template run<Ps>(), true) : false) || ...); }

// calls f.template run<P>() for every program of the list (host side)
template <class F, class... Ps>
inline void for_each_program(F& f, ProgramList<Ps...>) { (f.template run<Ps>(), ...); }

// the texture coordinate whose lane-0/1 difference sets the quad's LOD
template <class P>
__device__ inline bool lod_texcoord(const DrawDesc& d, const float* v, const FragCoord& fc, float& u, float& w) {
	if constexpr (P::kUV == kUVFromFragCoord) { P::texcoord(d, fc, u, w); return true; }
	else if constexpr (P::kUV >= 0) { u = v[P::kUV]; w = v[P::kUV + 1]; return true; }
	else { (void)d; (void)v; (void)fc; u = 0.0F; w = 0.0F; return false; }}

}  // namespace rgl
