// Learned shading on gfx950: a small MLP surrogate for the heavy shader
// (ProgGGXEnv), BF16 weights and activations with FP32 accumulation on
// MFMA (v_mfma_f32_16x16x32_bf16).
//
//   inputs   13 shading inputs of a fragment: unit normal, unit view
//            vector, world position (centred and scaled), albedo, roughness
//   encoding each input x -> x, sin(2^k pi x), cos(2^k pi x), k = 0..3
//            (117 values, padded with zeros to 128)
//   network  117 -> 64 -> 64 -> 64 -> 3, ReLU; output = sRGB-space colour
//
// Inference follows the resolve: a gather pass queues the shading inputs
// of covered pixels, a dense persistent pass runs them through the network,
// 64 per wave; the weights sit in LDS for the whole workgroup, and the
// encoded inputs and layer activations stay in registers, in the lanes the
// next product reads them from (the weight image's K columns are ordered
// to match).  Training uses a generic bf16 MFMA
// GEMM (operands rounded to bf16 when staged, FP32 accumulation) for the
// forward and backward passes, split-K weight gradients and an Adam update.
// No part of this has a counterpart in the reference, whose shaders are
// all analytic.
#pragma once
#include "rgl_kernels.h"

namespace rgl {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kMlpFeat = 13;
constexpr int kMlpBands = 4;
constexpr int kMlpEnc = kMlpFeat * (1 + 2 * kMlpBands);   // 117
constexpr int kMlpK1 = 128;                                // layer-1 inputs padded to the MFMA K step
constexpr int kMlpHid = 64, kMlpOut = 3;
constexpr int kMlpSample = 16;                             // floats per training sample: 13 inputs, 3 targets

// fp32 parameter block, offsets in floats.  W1 is [64][128] (columns of the
// padding stay zero), W2 and W3 [64][64], W4 [3][64]; kNorm holds the
// position centre (xyz) and 1 / half-extent, which are not trained.
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kMlpHid * kMlpK1;
constexpr int kW2 = kB1 + kMlpHid;
constexpr int kB2 = kW2 + kMlpHid * kMlpHid;
constexpr int kW3 = kB2 + kMlpHid;
constexpr int kB3 = kW3 + kMlpHid * kMlpHid;
constexpr int kW4 = kB3 + kMlpHid;
constexpr int kB4 = kW4 + kMlpOut * kMlpHid;
constexpr int kNorm = kB4 + kMlpOut;
constexpr int kMlpTrainable = kNorm;
constexpr int kMlpParams = kNorm + 4;

// bf16 image of the weights in the LDS layout the kernels read (rows padded
// by 8 elements so that 16 lanes reading 16 rows hit distinct banks),
// followed by the fp32 biases and the normalisation
constexpr int kLd1 = kMlpK1 + 8;      // 136
constexpr int kLdH = kMlpHid + 8;     // 72
constexpr int kImgW1 = 0;                          // uint16 offsets
constexpr int kImgW2 = kImgW1 + kMlpHid * kLd1;
constexpr int kImgW3 = kImgW2 + kMlpHid * kLdH;
constexpr int kImgW4 = kImgW3 + kMlpHid * kLdH;    // 16 rows, 3 used
constexpr int kImgEnd = kImgW4 + 16 * kLdH;        // 19072 uint16 = 38144 bytes
constexpr int kImgBiasF = kImgEnd / 2;             // float offset of b1 b2 b3 b4(4) norm(4)
constexpr int kImgFloats = kImgBiasF + 3 * kMlpHid + 4 + 4;
static_assert(kImgEnd % 8 == 0, "bias block must stay 16-byte aligned");

__device__ inline uint16_t bf16_bits(float v) {   // round to nearest even
	uint32_t u = f2u(v);
	u += 0x7fffu + ((u >> 16) & 1u);
	return static_cast<uint16_t>(u >> 16); }

__device__ inline float bf16_float(uint16_t b) { return u2f(static_cast<uint32_t>(b) << 16); }

// two values rounded to bf16 (nearest even) by one v_cvt_pk_bf16_f32, a in
// the low half; the same bits as bf16_bits for every finite value
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ inline uint32_t bf16_pack(float a, float b) {
	return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){ a, b }, bf16x2_t)); }

__device__ inline f32x4_t mfma_bf16(const bf16x8_t& a, const bf16x8_t& b, const f32x4_t& c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// ReLU of two packed bf16 values: as 16-bit integers, negative values
// (sign bit set, -0 included) are below 0 and positive ones keep their bits
typedef short i16x2_t __attribute__((ext_vector_type(2)));
__device__ inline uint32_t relu_bf16x2(uint32_t v) {
	return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2_t, v), (i16x2_t){ 0, 0 })); }

__device__ inline bf16x8_t lds_bf16x8(const uint16_t* p) { return *reinterpret_cast<const bf16x8_t*>(p); }

// K slot order of the weight image.  A 16x16x32 MFMA reads K slots
// 8 g .. 8 g + 7 of B from lane group g (= lane / 16), and leaves lane
// group g the result rows 4 g .. 4 g + 3.
// Layer 1: lane group g encodes features 3 g .. 3 g + 2 (slots q = 0..26
// of its 32, q = 8 c + t over the four K chunks c) and part of feature 12
// (group 0: its value and bands 0 and 1; group 1: bands 2 and 3).  Slot
// 32 c + 8 g + t therefore holds encoded input w1_slot_input (-1: none).
__device__ inline int w1_slot_input(int col) {
	const int c = col >> 5, g = (col >> 3) & 3, q = 8 * c + (col & 7);
	if (q < 27) return 27 * g + q;
	if (g == 0) return 108 + q - 27;
	if (g == 1 && q < 31) return 113 + q - 27;
	return -1; }
// Hidden layers: K chunk c of the next product takes, in lane group g, the
// rows 4 g .. 4 g + 3 of result tiles 2 c and 2 c + 1.
__device__ inline int hid_slot_input(int col) {
	const int c = col >> 5, g = (col >> 3) & 3, t = col & 7;
	return 32 * c + (t < 4 ? 4 * g + t : 16 + 4 * g + t - 4); }

// fp32 parameters -> bf16 LDS image, K columns in slot order
__global__ void k_mlp_pack(const float* __restrict__ prm, uint32_t* __restrict__ img) {
	uint16_t* h = reinterpret_cast<uint16_t*>(img);
	float* fl = reinterpret_cast<float*>(img);
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kImgFloats; i += gridDim.x * blockDim.x) {
		if (i < kImgBiasF) {
			for (int half = 0; half < 2; ++half) {
				const int e = 2 * i + half;
				float v = 0.0F;
				if (e < kImgW2) {
					const int r = e / kLd1, c = e - r * kLd1;
					const int src = c < kMlpK1 ? w1_slot_input(c) : -1;
					if (src >= 0) v = prm[kW1 + r * kMlpK1 + src]; }
				else if (e < kImgW3) { const int q = e - kImgW2, r = q / kLdH, c = q - r * kLdH; if (c < kMlpHid) v = prm[kW2 + r * kMlpHid + hid_slot_input(c)]; }
				else if (e < kImgW4) { const int q = e - kImgW3, r = q / kLdH, c = q - r * kLdH; if (c < kMlpHid) v = prm[kW3 + r * kMlpHid + hid_slot_input(c)]; }
				else {
					const int q = e - kImgW4, r = q / kLdH, c = q - r * kLdH;
					if (r < kMlpOut && c < kMlpHid) v = prm[kW4 + r * kMlpHid + hid_slot_input(c)]; }
				h[e] = bf16_bits(v); }}
		else {
			const int q = i - kImgBiasF;
			float v = 0.0F;
			if (q < 3 * kMlpHid) v = prm[kB1 + (q / kMlpHid) * (kMlpHid + kMlpHid * kMlpHid) + q % kMlpHid];
			else if (q < 3 * kMlpHid + 4) { const int c = q - 3 * kMlpHid; v = c < kMlpOut ? prm[kB4 + c] : 0.0F; }
			else v = prm[kNorm + (q - 3 * kMlpHid - 4)];
			fl[i] = v; }}}

// the 13 network inputs of a fragment
__device__ inline void mlp_inputs(const ShadeInput& s, const float* norm, float x[kMlpFeat]) {
	x[0] = s.n[0]; x[1] = s.n[1]; x[2] = s.n[2];
	x[3] = s.v[0]; x[4] = s.v[1]; x[5] = s.v[2];
	x[6] = (s.p[0] - norm[0]) * norm[3]; x[7] = (s.p[1] - norm[1]) * norm[3]; x[8] = (s.p[2] - norm[2]) * norm[3];
	x[9] = s.albedo[0]; x[10] = s.albedo[1]; x[11] = s.albedo[2];
	x[12] = s.rough; }

// frequency encoding, 9 values per input: x, then (sin, cos)(2^k pi x) for
// k = 0..3 by the double-angle recurrence from sin(pi x), cos(pi x)
__device__ inline float mlp_enc_value(const float x[kMlpFeat], int e) {
	if (e >= kMlpEnc) return 0.0F;
	const int i = e / 9, j = e - 9 * (e / 9);
	const float xi = x[i];
	if (j == 0) return xi;
	float sn = __builtin_amdgcn_sinf(0.5F * xi), cs = __builtin_amdgcn_cosf(0.5F * xi);   // revolutions
	const int kk = (j - 1) >> 1;
	for (int q = 0; q < kk; ++q) { const float s2 = 2.0F * sn * cs, c2 = (cs - sn) * (cs + sn); sn = s2; cs = c2; }
	return (j & 1) ? sn : cs; }

// sin and cos of 2^k pi v, k = 0..3, by the double-angle recurrence from
// sin(pi v), cos(pi v) (the operations of mlp_enc_value)
__device__ inline void enc_bands(float v, float sn[4], float cs[4]) {
	sn[0] = __builtin_amdgcn_sinf(0.5F * v); cs[0] = __builtin_amdgcn_cosf(0.5F * v);   // revolutions
#pragma unroll
	for (int k = 1; k < kMlpBands; ++k) { sn[k] = 2.0F * sn[k - 1] * cs[k - 1]; cs[k] = (cs[k - 1] - sn[k - 1]) * (cs[k - 1] + sn[k - 1]); }}

// one feature's 9 encoded values: x, then (sin, cos)(2^k pi x), k = 0..3,
// as bf16 into slot pairs of out starting at slot q (q + 9 <= 32)
__device__ inline void enc_feature(float x, float (&v)[32], int q) {
	float sn[4], cs[4];
	enc_bands(x, sn, cs);
	v[q] = x;
#pragma unroll
	for (int b = 0; b < kMlpBands; ++b) { v[q + 1 + 2 * b] = sn[b]; v[q + 2 + 2 * b] = cs[b]; }}

// lane group g's 32 layer-1 K slots of one sample (w1_slot_input's order)
// from the sample's inputs 3 g .. 3 g + 2 (fa, fb, fc) and 12 (fr), as 16
// bf16 pairs.  The inputs come as scalars picked by the caller: an array
// indexed by g would be kept in scratch memory.
__device__ inline void enc_group(int g, float fa, float fb, float fc, float fr, uint32_t out[16]) {
	float v[32];
	enc_feature(fa, v, 0);
	enc_feature(fb, v, 9);
	enc_feature(fc, v, 18);
	float sn[4], cs[4];
	enc_bands(fr, sn, cs);
	const bool g0 = g == 0, g1 = g == 1;
	v[27] = g0 ? fr : (g1 ? sn[2] : 0.0F);
	v[28] = g0 ? sn[0] : (g1 ? cs[2] : 0.0F);
	v[29] = g0 ? cs[0] : (g1 ? sn[3] : 0.0F);
	v[30] = g0 ? sn[1] : (g1 ? cs[3] : 0.0F);
	v[31] = g0 ? cs[1] : 0.0F;
#pragma unroll
	for (int u = 0; u < 16; ++u) out[u] = bf16_pack(v[2 * u], v[2 * u + 1]); }

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ inline bf16x8_t as_bf16x8(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
	return __builtin_bit_cast(bf16x8_t, (u32x4_t){ a, b, c, d }); }

// one wave: 64 samples through the network.  Lane (g, r) = (lane / 16,
// lane % 16) holds samples 16 j + r, j = 0..3: enc[j] is sample j's
// layer-1 K slots of group g, and lanes of group 0 get y[j], its three
// outputs.  Every product starts from its layer's biases; activations are
// ReLU'd to bf16 in registers and read by the next product where they lie;
// W (the LDS image) is the only operand read from memory.
__device__ inline void mlp_wave(const uint16_t* W, const float* bias, const uint32_t enc[4][16], float y[4][kMlpOut]) {
	const int lane = threadIdx.x & 63, r16 = lane & 15, g = lane >> 4;
	f32x4_t acc[4][4];
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		const f32x4_t bv = *reinterpret_cast<const f32x4_t*>(bias + 16 * i + 4 * g);
#pragma unroll
		for (int j = 0; j < 4; ++j) acc[i][j] = bv; }
	// layer 1, K = 128 in four chunks of 32 slots
#pragma unroll
	for (int c = 0; c < kMlpK1 / 32; ++c) {
		bf16x8_t a[4];
#pragma unroll
		for (int i = 0; i < 4; ++i) a[i] = lds_bf16x8(W + kImgW1 + (16 * i + r16) * kLd1 + 32 * c + g * 8);
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const bf16x8_t bj = as_bf16x8(enc[j][4 * c], enc[j][4 * c + 1], enc[j][4 * c + 2], enc[j][4 * c + 3]);
#pragma unroll
			for (int i = 0; i < 4; ++i) acc[i][j] = mfma_bf16(a[i], bj, acc[i][j]); }}
	// layers 2 and 3, then the output layer; h[c][j]: K chunk c of tile j
	bf16x8_t h[2][4];
#pragma unroll
	for (int layer = 0; layer < 3; ++layer) {
#pragma unroll
		for (int c = 0; c < 2; ++c)
#pragma unroll
			for (int j = 0; j < 4; ++j) {
				const f32x4_t lo = acc[2 * c][j], hi = acc[2 * c + 1][j];
				h[c][j] = as_bf16x8(relu_bf16x2(bf16_pack(lo[0], lo[1])), relu_bf16x2(bf16_pack(lo[2], lo[3])),
				                    relu_bf16x2(bf16_pack(hi[0], hi[1])), relu_bf16x2(bf16_pack(hi[2], hi[3]))); }
		if (layer == 2) break;
		const uint16_t* Wl = W + (layer == 0 ? kImgW2 : kImgW3);
		const float* bl = bias + (layer + 1) * kMlpHid;
#pragma unroll
		for (int i = 0; i < 4; ++i) {
			const f32x4_t bv = *reinterpret_cast<const f32x4_t*>(bl + 16 * i + 4 * g);
#pragma unroll
			for (int j = 0; j < 4; ++j) acc[i][j] = bv; }
#pragma unroll
		for (int c = 0; c < 2; ++c) {
			bf16x8_t a[4];
#pragma unroll
			for (int i = 0; i < 4; ++i) a[i] = lds_bf16x8(Wl + (16 * i + r16) * kLdH + 32 * c + g * 8);
#pragma unroll
			for (int j = 0; j < 4; ++j)
#pragma unroll
				for (int i = 0; i < 4; ++i) acc[i][j] = mfma_bf16(a[i], h[c][j], acc[i][j]); }}
	// output layer: 3 outputs in rows 0..2 of one 16-row tile (lane group 0)
	const f32x4_t b4 = g == 0 ? *reinterpret_cast<const f32x4_t*>(bias + 3 * kMlpHid) : f32x4_t{ 0.0F, 0.0F, 0.0F, 0.0F };
	f32x4_t o[4] = { b4, b4, b4, b4 };
#pragma unroll
	for (int c = 0; c < 2; ++c) {
		const bf16x8_t a = lds_bf16x8(W + kImgW4 + r16 * kLdH + 32 * c + g * 8);
#pragma unroll
		for (int j = 0; j < 4; ++j) o[j] = mfma_bf16(a, h[c][j], o[j]); }
#pragma unroll
	for (int j = 0; j < 4; ++j)
#pragma unroll
		for (int ch = 0; ch < kMlpOut; ++ch) y[j][ch] = o[j][ch]; }

// waves per inference workgroup (64 samples each); the workgroup's LDS
// holds only the weight image, and 8 waves leave each up to 256 VGPRs
constexpr int kMlpWaves = 8;

struct MlpShared {
	uint16_t w[kImgEnd];
	float bias[3 * kMlpHid + 4];
	float norm[4];
};

__device__ inline void mlp_load_image(MlpShared& sh, const uint32_t* __restrict__ img) {
	const uint4* src = reinterpret_cast<const uint4*>(img);
	uint4* dst = reinterpret_cast<uint4*>(sh.w);
	for (int i = threadIdx.x; i < kImgEnd / 8; i += blockDim.x) dst[i] = src[i];
	const float* fb = reinterpret_cast<const float*>(img) + kImgBiasF;
	for (int i = threadIdx.x; i < 3 * kMlpHid + 4; i += blockDim.x) sh.bias[i] = fb[i];
	if (threadIdx.x < 4) sh.norm[threadIdx.x] = fb[3 * kMlpHid + 4 + threadIdx.x]; }

__device__ inline uint32_t srgb_unit_to8(float v) {   // sRGB-space [0,1] -> byte
	v = v > 0.0F ? v : 0.0F;
	v = v < 1.0F ? v : 1.0F;
	return static_cast<uint32_t>(v * 255.0F + 0.5F); }

// dense inference: n samples of kMlpSample floats (13 inputs used) -> n x 3;
// wave w of block k takes samples 64 (k kMlpWaves + w) .. + 63
__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_forward(const float* __restrict__ in, int n,
		const uint32_t* __restrict__ img, float* __restrict__ out) {
	__shared__ __attribute__((aligned(16))) MlpShared sh;
	mlp_load_image(sh, img);
	__syncthreads();
	const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, g = lane >> 4;
	const long long base = (static_cast<long long>(blockIdx.x) * kMlpWaves + wv) * 64;
	uint32_t enc[4][16];
#pragma unroll
	for (int j = 0; j < 4; ++j) {
		// lane group g reads inputs 3 g .. 3 g + 2 and 12 of its sample;
		// group 2's are the position, normalised as mlp_inputs does
		const long long i = base + 16 * j + r16;
		const float* r = in + (i < n ? i : 0) * kMlpSample;
		float fa = r[3 * g], fb = r[3 * g + 1], fc = r[3 * g + 2];
		if (g == 2) { fa = (fa - sh.norm[0]) * sh.norm[3]; fb = (fb - sh.norm[1]) * sh.norm[3]; fc = (fc - sh.norm[2]) * sh.norm[3]; }
		enc_group(g, fa, fb, fc, r[12], enc[j]); }
	float y[4][kMlpOut];
	mlp_wave(sh.w, sh.bias, enc, y);
	if (g == 0)
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const long long i = base + 16 * j + r16;
			if (i < n) { out[i * 3 + 0] = y[j][0]; out[i * 3 + 1] = y[j][1]; out[i * 3 + 2] = y[j][2]; }}}

// ---------------------------------------------------------------------------
// learned resolve: like k_resolve, but the winner's shading inputs go
// through the network (a gather pass, then the network pass over the
// queued fragments).  Depth and primitive id are the same stores as the
// plain resolve.
// ---------------------------------------------------------------------------
template <class P>
__device__ inline bool resolve_inputs(const BatchDesc& b, const FrameDesc& f, const RasterArgs& a, const DrawLite* dl, int x, int y,
                                      const uint32_t* rcpTab, ShadeInput& s, float& dep, uint32_t& prim) {
	const unsigned long long k = a.keys[f.keyBase + static_cast<long long>(y) * f.W + x];
	dep = f.clearDepth;
	prim = 0xffffffffu;
	if (!(k < clear_key(f))) return false;
	const uint32_t slot = key_slot(k, f.depthFunc == 1);
	const int di = find_draw_slot(dl, b.nDraws, slot);
	const DrawLite& dd = dl[di];
	const uint32_t rel = slot - dd.slotBase;
	const uint32_t nTri = static_cast<uint32_t>(dd.tris) * static_cast<uint32_t>(dd.instances);
	const int tx = x / f.tileW, ty = y / f.tileH;
	Rect R;
	R.x0 = tx * f.tileW; R.y0 = ty * f.tileH;
	R.x1 = imin(R.x0 + f.tileW, f.W); R.y1 = imin(R.y0 + f.tileH, f.H);
	TriVerts tr;
	TriSetup t;
	uint32_t srcTri;
	const float* vary;
	if (rel < nTri) {
		srcTri = rel;
		inline_verts(a.tvert[dd.triBase + rel], a.vpos, tr, a.varyStride);
		setup_vector(R, tr.X, tr.Y, t);
		vary = a.vvary; }
	else {
		const uint32_t fslot = rel - nTri;
		srcTri = fslot / kFanMax;
		const uint32_t j = fslot - srcTri * kFanMax;
		const uint32_t cq = a.clipqOf[dd.triBase + srcTri];
		clipped_verts(cq * kFanMax + j, a.cvpos, tr, a.varyStride);
		setup_scalar(R, tr.X, tr.Y, t);
		vary = a.cvvary; }
	const int kx = ((x & ~1) - t.minx) >> 1, ky = ((y & ~1) - t.miny) >> 1;
	const int lane = (x & 1) | ((y & 1) << 1);
	const DrawDesc& d = a.draws[di];
	// the quad lane exactly as shade_pixel interpolates it
	const int32_t q1 = wadd(wadd(t.c1, wmul(kx, wmul(t.dy12, 2))), wmul(ky, wmul(t.dx12, 2)));
	const int32_t q2 = wadd(wadd(t.c2, wmul(kx, wmul(t.dy23, 2))), wmul(ky, wmul(t.dx23, 2)));
	const int32_t e1 = (lane == 0) ? q1 : (lane == 1) ? wadd(q1, t.dy12) : (lane == 2) ? wadd(q1, t.dx12) : wadd(wadd(q1, t.dx12), t.dy12);
	const int32_t e2 = (lane == 0) ? q2 : (lane == 1) ? wadd(q2, t.dy23) : (lane == 2) ? wadd(q2, t.dx23) : wadd(wadd(q2, t.dx23), t.dy23);
	const float* v0 = vary + tr.vary[0];
	const float* v1 = vary + tr.vary[1];
	const float* v2 = vary + tr.vary[2];
	const LaneBary own = lane_bary(t, tr, e1, e2, rcpTab);
	dep = own.depth;
	float vv[P::kVary];
#pragma unroll
	for (int k2 = 0; k2 < P::kVary; ++k2) vv[k2] = lane_interp(own, v0, v1, v2, k2);
	const LaneBary l0 = lane_bary(t, tr, q1, q2, rcpTab);
	const LaneBary l1 = lane_bary(t, tr, wadd(q1, t.dy12), wadd(q2, t.dy23), rcpTab);
	const int lod = texture_lod(d.tu0, lane_interp(l0, v0, v1, v2, P::kUV), lane_interp(l1, v0, v1, v2, P::kUV),
	                            lane_interp(l0, v0, v1, v2, P::kUV + 1), lane_interp(l1, v0, v1, v2, P::kUV + 1));
	P::inputs(d, vv, lod, s);
	prim = dd.primBase + srcTri;
	return true; }

// Per-tile error estimate: the plain shader at kMlpProbes jittered pixels of
// each tile (one in each cell of an 8 x 8 grid), compared with the
// network's colour there.  The probes' plain colours do not depend on the
// network: the gather pass computes them in leading workgroups of its own
// (one wave per tile) while its other workgroups queue the fragments; the
// comparison is made by the fallback pass, after the network has stored
// its colours.
constexpr int kMlpProbes = 64;
constexpr int kProbeTilesPerBlock = 256 / kMlpProbes;   // tiles per gather workgroup of probes
constexpr uint32_t kNoProbe = 0xffffffffu;   // probe pixel not covered (or no colour output)

__device__ inline uint32_t hash32(uint32_t v) {
	v ^= v >> 16; v *= 0x7feb352dU; v ^= v >> 15; v *= 0x846ca68bU; v ^= v >> 16; return v; }

// probe j of batch tile `tile`, whose pixel rect is R; false when the cell
// falls outside the rect
__device__ inline bool probe_pixel(const Rect& R, int tile, int j, int& x, int& y) {
	const int cw = (R.x1 - R.x0 + 7) / 8, ch = (R.y1 - R.y0 + 7) / 8;   // 8 x 8 cells
	const uint32_t id = static_cast<uint32_t>(tile) * kMlpProbes + static_cast<uint32_t>(j);
	const uint32_t h = hash32(id * 0x9e3779b9u + 17u);
	x = R.x0 + (j & 7) * cw + static_cast<int>(h % static_cast<uint32_t>(cw));
	y = R.y0 + (j >> 3) * ch + static_cast<int>((h >> 16) % static_cast<uint32_t>(ch));
	return x < R.x1 && y < R.y1; }

// one wave: the plain colour of each probe pixel of tile `tile` (lane j:
// probe j), or kNoProbe
template <class P>
__device__ inline void tile_probes(const BatchDesc& b, const RasterArgs& a, const DrawLite* dl, const int* ftb, int tile,
                                   const uint32_t* rcpTab, uint32_t* __restrict__ want) {
	const int j = static_cast<int>(threadIdx.x & 63);
	const FrameDesc& f = b.frames[find_base(ftb, b.nFrames, tile)];
	uint32_t w = kNoProbe;
	int x, y;
	if (f.outColor != nullptr && probe_pixel(tile_rect(f, tile - f.tileBase), tile, j, x, y)) {
		float col[3], dep;
		uint32_t prim;
		if (resolve_pixel<ShadeOne<P>>(b, f, a, dl, x, y, rcpTab, col, dep, prim)) w = post_pixel(f, col, x, y, kSrgb8Tab); }
	want[static_cast<long long>(tile) * kMlpProbes + j] = w; }

// A queued fragment, 32 bytes: its 13 network inputs (normalised as the
// network reads them) in 16-bit fixed point over [-2, 2] (steps of 6.1e-5,
// well inside bf16's rounding of the encoded values; the inputs of a
// fragment inside the normalisation box lie in [-1, 1]), its frame, and
// its pixel index.
constexpr int kQueueWords = 8;
constexpr float kQueueScale = 16383.5F;   // 32767 / 2

__device__ inline uint32_t queue_q16(float v) {
	v = v > 2.0F ? 2.0F : (v < -2.0F ? -2.0F : v);
	return static_cast<uint32_t>(static_cast<int>(__builtin_rintf(v * kQueueScale))) & 0xffffu; }

__device__ inline float queue_f16(uint32_t h) {
	return static_cast<float>(static_cast<int>(static_cast<int16_t>(static_cast<uint16_t>(h)))) * (1.0F / kQueueScale); }

// learned resolve, pass 1: like k_resolve for depth, ids and the clear
// colour; each covered pixel's network inputs are queued (the 32-byte
// record above) for the dense network pass.  The queue is
// cut into kQueueSegs segments with a counter each, picked by block index,
// so that consecutive workgroups (dealt round-robin to the XCDs) append
// through different counters.  The first probeBlocks workgroups compute
// the tiles' probes instead (kProbeTilesPerBlock tiles each), beside the
// pixel workgroups.
constexpr int kQueueSegs = 8;
constexpr int kQueueStride = 32;   // counters 128 bytes apart

template <class P>
__global__ __launch_bounds__(256) void k_mlp_gather(BatchDesc b, RasterArgs a, float* __restrict__ queue, uint32_t* __restrict__ count,
                                                    uint32_t segCap, uint32_t* __restrict__ want, int probeBlocks,
                                                    const uint32_t* __restrict__ img) {
	extern __shared__ uint4 dyn_lds[];   // the batch's draws
	DrawLite* dl = reinterpret_cast<DrawLite*>(dyn_lds);
	__shared__ int fbb[kMaxFrames + 1];
	const int blk = static_cast<int>(blockIdx.x);
	load_lite(dl, a.draws, b.nDraws);
	if (blk < probeBlocks) load_tile_bases(fbb, b);
	else load_block_bases(fbb, b);
	__syncthreads();
	// the rcpps and sRGB tables are read through the caches, as the plain
	// resolve reads them
	if (blk < probeBlocks) {   // uniform over the block
		const int tile = blk * kProbeTilesPerBlock + static_cast<int>(threadIdx.x >> 6);
		if (tile < b.nTiles) tile_probes<P>(b, a, dl, fbb, tile, kRcpMantissa, want);
		return; }
	const BlockPixel bp = block_pixel(b, fbb, blk - probeBlocks);
	const FrameDesc& f = b.frames[bp.frame];
	const int x = bp.x, y = bp.y;
	const bool on = x < f.W && y < f.H;
	ShadeInput s;
	float dep = f.clearDepth;
	uint32_t prim = 0xffffffffu;
	const bool hit = on && resolve_inputs<P>(b, f, a, dl, x, y, kRcpMantissa, s, dep, prim);
	const long long o = static_cast<long long>(y) * f.W + x;
	// one queue reservation per workgroup
	__shared__ uint32_t wcnt[5];
	const unsigned long long m = __ballot(hit);
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	if (lane == 0) wcnt[wv] = static_cast<uint32_t>(__popcll(m));
	__syncthreads();
	const int seg = static_cast<int>((blk - probeBlocks) % kQueueSegs);
	if (threadIdx.x == 0) {
		const uint32_t tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
		wcnt[4] = tot ? atomicAdd(&count[seg * kQueueStride], tot) : 0u; }
	__syncthreads();
	uint32_t base = wcnt[4];
	for (int w = 0; w < wv; ++w) base += wcnt[w];
	if (hit) {
		const uint32_t slot = base + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
		if (slot < segCap) {
			// the position's normalisation sits after the biases of the image
			float xin[kMlpFeat];
			mlp_inputs(s, reinterpret_cast<const float*>(img) + kImgBiasF + 3 * kMlpHid + 4, xin);
			uint32_t w[6];
#pragma unroll
			for (int c = 0; c < 6; ++c) w[c] = queue_q16(xin[2 * c]) | (queue_q16(xin[2 * c + 1]) << 16);
			uint4* r = reinterpret_cast<uint4*>(queue + (static_cast<long long>(seg) * segCap + slot) * kQueueWords);
			r[0] = make_uint4(w[0], w[1], w[2], w[3]);
			r[1] = make_uint4(w[4], w[5], queue_q16(xin[12]) | (static_cast<uint32_t>(bp.frame) << 16), static_cast<uint32_t>(o)); }}
	if (!on) return;
	if (f.outColor && !hit) f.outColor[o] = post_pixel(f, f.clearColor, x, y, kSrgb8Tab);
	if (f.outDepth) f.outDepth[o] = dep;
	if (f.outPrim) f.outPrim[o] = prim; }

// learned resolve, pass 2: the queued fragments through the network, 64
// per wave batch; persistent workgroups keep the weights in LDS, and each
// wave takes batches on its own (no workgroup barrier after the weights)
__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_shade(BatchDesc b, const float* __restrict__ queue,
		const uint32_t* __restrict__ count, uint32_t segCap, const uint32_t* __restrict__ img) {
	__shared__ __attribute__((aligned(16))) MlpShared sh;
	mlp_load_image(sh, img);
	__syncthreads();
	uint32_t cnt[kQueueSegs], first[kQueueSegs + 1];
	first[0] = 0;
#pragma unroll
	for (int q = 0; q < kQueueSegs; ++q) {
		cnt[q] = min(count[q * kQueueStride], segCap);
		first[q + 1] = first[q] + (cnt[q] + 63u) / 64u; }
	// the wave's index is uniform: its batch, segment and counts sit in
	// scalar registers
	const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)), lane = threadIdx.x & 63, g = lane >> 4;
	const uint32_t stride = gridDim.x * kMlpWaves;
	for (uint32_t item = blockIdx.x * kMlpWaves + wv; item < first[kQueueSegs]; item += stride) {   // uniform over the wave
		// the weights' LDS reads stay in the loop: hoisted, the 34 KB image
		// would not fit the registers (this emits no instruction)
		__asm__ volatile("" ::: "memory");
		int seg = 0;
#pragma unroll
		for (int q = 1; q < kQueueSegs; ++q) if (item >= first[q]) seg = q;
		const uint32_t base = (item - first[seg]) * 64u;
		// lane (g, r) encodes its slots of samples base + 16 j + r
		uint32_t enc[4][16], po[4], pf[4];
		bool pok[4];
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const uint32_t i = base + 16u * j + (lane & 15);
			pok[j] = i < cnt[seg];
			const uint4* r = reinterpret_cast<const uint4*>(queue + (static_cast<long long>(seg) * segCap + (pok[j] ? i : base)) * kQueueWords);
			const uint4 r0 = r[0], r1 = r[1];
			// input f is 16-bit slot f of the record: lane group g's inputs
			// 3 g, 3 g + 1, 3 g + 2 are picked from the words that hold them
			const uint32_t wa = g < 2 ? (g == 0 ? r0.x : r0.y) : (g == 2 ? r0.w : r1.x);
			const uint32_t wb = g < 2 ? (g == 0 ? r0.x : r0.z) : (g == 2 ? r0.w : r1.y);
			const uint32_t wc = g < 2 ? (g == 0 ? r0.y : r0.z) : (g == 2 ? r1.x : r1.y);
			const int odd = (g & 1) * 16;   // 3 g is odd for g = 1, 3
			enc_group(g, queue_f16(wa >> odd), queue_f16(wb >> (16 - odd)), queue_f16(wc >> odd), queue_f16(r1.z), enc[j]);
			po[j] = r1.w; pf[j] = r1.z >> 16; }
		float y[4][kMlpOut];
		mlp_wave(sh.w, sh.bias, enc, y);
		if (lane < 16)
#pragma unroll
			for (int j = 0; j < 4; ++j) {
				uint32_t* out = b.frames[pf[j]].outColor;
				if (pok[j] && out) out[po[j]] = (srgb_unit_to8(y[j][0]) << 16) | (srgb_unit_to8(y[j][1]) << 8) | srgb_unit_to8(y[j][2]); }}}

// the fallback decision of one tile, by one wave: the mean squared 8-bit
// error per channel at the tile's covered probes (the network's stored
// colour against the plain one) over the limit.  The result has a bit per
// tile quadrant, q = (right half) + 2 (bottom half), the unit the
// re-shading is dealt out in: all four or none.  (A decision per quadrant,
// from its 16 probes, missed quadrants whose probes all fell on good
// pixels: a 4K view then measured 39 dB however low the limit.)
__device__ inline uint32_t tile_over_limit(const FrameDesc& f, int tile, const uint32_t* __restrict__ want, float maxMse) {
	const int j = threadIdx.x & 63;
	int x, y;
	float se = 0.0F, cnt = 0.0F;
	const uint32_t w = want[static_cast<long long>(tile) * kMlpProbes + j];
	if (w != kNoProbe && probe_pixel(tile_rect(f, tile - f.tileBase), tile, j, x, y)) {
		const uint32_t got = f.outColor[static_cast<long long>(y) * f.W + x];
		for (int c = 0; c < 3; ++c) {
			const float dv = static_cast<float>(static_cast<int>((w >> (8 * c)) & 255u) - static_cast<int>((got >> (8 * c)) & 255u));
			se += dv * dv; }
		cnt = 3.0F; }
	for (int o = 32; o > 0; o >>= 1) { se += __shfl_xor(se, o, 64); cnt += __shfl_xor(cnt, o, 64); }
	return cnt > 0.0F && se / cnt > maxMse ? 0xfu : 0u; }

// the fallback decision: one wave per tile appends the quadrants of a tile
// over the limit (tile * 4 + quadrant) to `list`, counting them in `count`
// (zeroed at the frame's start)
__global__ __launch_bounds__(256) void k_mlp_decide(BatchDesc b, const uint32_t* __restrict__ want, float maxMse,
                                                    uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
	__shared__ int ftb[kMaxFrames + 1];
	load_tile_bases(ftb, b);
	__syncthreads();
	const int tile = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
	if (tile >= b.nTiles) return;   // uniform over the wave
	const FrameDesc& f = b.frames[find_base(ftb, b.nFrames, tile)];
	if (f.outColor == nullptr) return;
	const uint32_t over = tile_over_limit(f, tile, want, maxMse);
	if ((threadIdx.x & 63) == 0 && over) {
		uint32_t at = atomicAdd(count, static_cast<uint32_t>(__popc(over)));
		for (int q = 0; q < 4; ++q)
			if ((over >> q) & 1u) list[at++] = static_cast<uint32_t>(tile) * 4u + static_cast<uint32_t>(q); }}

// the listed tile quadrants re-shaded with the plain shader (colour only;
// depth and ids are already stored): subX x subY workgroups of 16 x 16
// pixels per list entry, one thread per pixel; workgroups past the list's
// end exit at once
template <class P>
__global__ __launch_bounds__(256) void k_mlp_fallback(BatchDesc b, RasterArgs a, const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count, int subX, int subY) {
	extern __shared__ uint4 dyn_lds[];   // the batch's draws
	DrawLite* dl = reinterpret_cast<DrawLite*>(dyn_lds);
	__shared__ int ftb[kMaxFrames + 1];
	const int per = subX * subY;
	const uint32_t item = blockIdx.x / static_cast<uint32_t>(per);
	if (item >= *count) return;   // uniform over the block
	load_lite(dl, a.draws, b.nDraws);
	load_tile_bases(ftb, b);
	__syncthreads();
	const uint32_t id = list[item];
	const int tile = static_cast<int>(id >> 2), q = static_cast<int>(id & 3u);
	const FrameDesc& f = b.frames[find_base(ftb, b.nFrames, tile)];
	// the quadrant: the tile split where its probe cells split
	const Rect R = tile_rect(f, tile - f.tileBase);
	const int cw = (R.x1 - R.x0 + 7) / 8, ch = (R.y1 - R.y0 + 7) / 8;
	const int qx0 = (q & 1) ? R.x0 + 4 * cw : R.x0, qx1 = (q & 1) ? R.x1 : min(R.x0 + 4 * cw, R.x1);
	const int qy0 = (q >> 1) ? R.y0 + 4 * ch : R.y0, qy1 = (q >> 1) ? R.y1 : min(R.y0 + 4 * ch, R.y1);
	const int sub = static_cast<int>(blockIdx.x % static_cast<uint32_t>(per));
	const int x = qx0 + (sub % subX) * 16 + static_cast<int>(threadIdx.x & 15);
	const int y = qy0 + (sub / subX) * 16 + static_cast<int>(threadIdx.x >> 4);
	if (x >= qx1 || y >= qy1) return;
	float col[3], dep;
	uint32_t prim;
	if (resolve_pixel<ShadeOne<P>>(b, f, a, dl, x, y, kRcpMantissa, col, dep, prim))
		f.outColor[static_cast<long long>(y) * f.W + x] = post_pixel(f, col, x, y, kSrgb8Tab); }

// training samples from a frame: covered pixels kept with probability
// `keep` (hash of pixel and seed), appended as 13 inputs + the plain
// shader's sRGB colour / 255
template <class P>
__global__ __launch_bounds__(256) void k_mlp_samples(BatchDesc b, RasterArgs a, float* __restrict__ out, uint32_t cap,
                                                     uint32_t* __restrict__ count, float keep, uint32_t seed) {
	extern __shared__ uint4 dyn_lds[];   // the batch's draws
	DrawLite* dl = reinterpret_cast<DrawLite*>(dyn_lds);
	__shared__ int fbb[kMaxFrames + 1];
	load_lite(dl, a.draws, b.nDraws);
	load_block_bases(fbb, b);
	__syncthreads();
	const BlockPixel bp = block_pixel(b, fbb);
	const FrameDesc& f = b.frames[bp.frame];
	const int x = bp.x, y = bp.y;
	if (x >= f.W || y >= f.H) return;
	const long long id = f.keyBase + static_cast<long long>(y) * f.W + x;
	const uint32_t h = hash32(static_cast<uint32_t>(id) ^ hash32(seed));
	if (static_cast<float>(h >> 8) * (1.0F / 16777216.0F) >= keep) return;
	ShadeInput s;
	float dep;
	uint32_t prim;
	if (!resolve_inputs<P>(b, f, a, dl, x, y, kRcpMantissa, s, dep, prim)) return;
	float col[3];
	resolve_pixel<ShadeOne<P>>(b, f, a, dl, x, y, kRcpMantissa, col, dep, prim);
	const uint32_t px = post_pixel(f, col, x, y, kSrgb8Tab);
	const uint32_t slot = atomicAdd(count, 1u);
	if (slot >= cap) return;
	float* r = out + static_cast<long long>(slot) * kMlpSample;
	r[0] = s.n[0]; r[1] = s.n[1]; r[2] = s.n[2]; r[3] = s.v[0]; r[4] = s.v[1]; r[5] = s.v[2];
	r[6] = s.p[0]; r[7] = s.p[1]; r[8] = s.p[2]; r[9] = s.albedo[0]; r[10] = s.albedo[1]; r[11] = s.albedo[2];
	r[12] = s.rough;
	r[13] = static_cast<float>((px >> 16) & 255u) * (1.0F / 255.0F);
	r[14] = static_cast<float>((px >> 8) & 255u) * (1.0F / 255.0F);
	r[15] = static_cast<float>(px & 255u) * (1.0F / 255.0F); }

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
// batch: encoded inputs X [B][128] (fp32) and targets T [B][3] for samples
// chosen by hash (or in order when seed == 0xffffffff)
__global__ void k_mlp_batch(const float* __restrict__ samples, long long n, int B, uint32_t seed, uint32_t step,
                            const float* __restrict__ prm, float* __restrict__ X, float* __restrict__ T) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= B) return;
	long long src;
	if (seed == 0xffffffffu) src = i;
	else src = static_cast<long long>(hash32(hash32(seed ^ (step * 0x9e3779b9u)) + static_cast<uint32_t>(i)) % static_cast<uint64_t>(n));
	const float* r = samples + src * kMlpSample;
	ShadeInput s;
	s.n[0] = r[0]; s.n[1] = r[1]; s.n[2] = r[2]; s.v[0] = r[3]; s.v[1] = r[4]; s.v[2] = r[5];
	s.p[0] = r[6]; s.p[1] = r[7]; s.p[2] = r[8]; s.albedo[0] = r[9]; s.albedo[1] = r[10]; s.albedo[2] = r[11];
	s.rough = r[12];
	float x[kMlpFeat];
	mlp_inputs(s, prm + kNorm, x);
	float* xo = X + static_cast<long long>(i) * kMlpK1;
	for (int e = 0; e < kMlpK1; ++e) xo[e] = mlp_enc_value(x, e);
	T[i * 3 + 0] = r[13]; T[i * 3 + 1] = r[14]; T[i * 3 + 2] = r[15]; }

// C = epilogue(op(A) op(B)) with op(A)(m,k) = transA ? A[k*lda+m] : A[m*lda+k]
// and op(B)(k,n) = transB ? B[n*ldb+k] : B[k*ldb+n]; operands are rounded
// to bf16 as they are staged in LDS, products accumulate in fp32.
// blockIdx.z splits K into chunks of kChunk; chunk z writes C + z*zStride.
struct GemmArgs {
	const float* A; const float* B; float* C;
	int M, N, K, lda, ldb, ldc, transA, transB;
	const float* bias;     // + bias[n]
	int relu;
	const float* mask; int ldm;   // * (mask[m*ldm+n] > 0)
	int kChunk; long long zStride;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
	__shared__ __attribute__((aligned(16))) uint16_t As[64 * 40];
	__shared__ __attribute__((aligned(16))) uint16_t Bs[64 * 40];
	const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
	const int kb = blockIdx.z * g.kChunk, ke = min(g.K, kb + g.kChunk);
	const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, r16 = lane & 15, q = lane >> 4;
	const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
	f32x4_t acc[2][2];
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{ 0.0F, 0.0F, 0.0F, 0.0F };
	for (int k0 = kb; k0 < ke; k0 += 32) {
		// stage 64 x 32 of op(A) and op(B) as [row][k] bf16
#pragma unroll
		for (int e = 0; e < 8; ++e) {
			const int idx = e * 256 + tid;
			int r, k;
			if (g.transA) { k = idx >> 6; r = idx & 63; } else { r = idx >> 5; k = idx & 31; }
			const int gm = m0 + r, gk = k0 + k;
			float v = 0.0F;
			if (gm < g.M && gk < ke) v = g.transA ? g.A[static_cast<long long>(gk) * g.lda + gm] : g.A[static_cast<long long>(gm) * g.lda + gk];
			As[r * 40 + k] = bf16_bits(v); }
#pragma unroll
		for (int e = 0; e < 8; ++e) {
			const int idx = e * 256 + tid;
			int r, k;
			if (g.transB) { r = idx >> 5; k = idx & 31; } else { k = idx >> 6; r = idx & 63; }
			const int gn = n0 + r, gk = k0 + k;
			float v = 0.0F;
			if (gn < g.N && gk < ke) v = g.transB ? g.B[static_cast<long long>(gn) * g.ldb + gk] : g.B[static_cast<long long>(gk) * g.ldb + gn];
			Bs[r * 40 + k] = bf16_bits(v); }
		__syncthreads();
		bf16x8_t a[2], b[2];
#pragma unroll
		for (int i = 0; i < 2; ++i) a[i] = lds_bf16x8(As + (wm + 16 * i + r16) * 40 + q * 8);
#pragma unroll
		for (int j = 0; j < 2; ++j) b[j] = lds_bf16x8(Bs + (wn + 16 * j + r16) * 40 + q * 8);
#pragma unroll
		for (int i = 0; i < 2; ++i)
#pragma unroll
			for (int j = 0; j < 2; ++j) acc[i][j] = mfma_bf16(a[i], b[j], acc[i][j]);
		__syncthreads(); }
	float* C = g.C + blockIdx.z * g.zStride;
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 2; ++j)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int m = m0 + wm + 16 * i + 4 * q + r, n = n0 + wn + 16 * j + r16;
				if (m >= g.M || n >= g.N) continue;
				float v = acc[i][j][r];
				if (g.bias) v += g.bias[n];
				if (g.relu) v = v > 0.0F ? v : 0.0F;
				if (g.mask && !(g.mask[static_cast<long long>(m) * g.ldm + n] > 0.0F)) v = 0.0F;
				C[static_cast<long long>(m) * g.ldc + n] = v; }}

// dY = 2 (Y - T) / (3 B); loss partial sums of squared error
__global__ void k_mlp_loss(const float* __restrict__ Y, const float* __restrict__ T, int B, float* __restrict__ dY,
                           float* __restrict__ lossSum) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	float se = 0.0F;
	if (i < 3 * B) {
		const float d = Y[i] - T[i];
		dY[i] = d * (2.0F / (3.0F * static_cast<float>(B)));
		se = d * d; }
	for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
	if ((threadIdx.x & 63) == 0 && se != 0.0F) atomicAdd(lossSum, se); }

// column sums of dZ [rows][cols] over row chunks -> partial z of a bias gradient
__global__ void k_mlp_colsum(const float* __restrict__ dZ, int rows, int cols, int rowChunk, float* __restrict__ out,
                             long long zStride) {
	const int c = threadIdx.x;
	if (c >= cols) return;
	const int r0 = blockIdx.x * rowChunk, r1 = min(rows, r0 + rowChunk);
	float s = 0.0F;
	for (int r = r0; r < r1; ++r) s += dZ[static_cast<long long>(r) * cols + c];
	out[blockIdx.x * zStride + c] = s; }

// sums the split partials; Adam update when lr > 0, else writes the gradient
__global__ void k_mlp_adam(const float* __restrict__ part, int nz, float* __restrict__ prm, float* __restrict__ mom,
                           float* __restrict__ vel, float lr, float b1, float b2, float eps, float bc1, float bc2,
                           float* __restrict__ gradOut) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= kMlpTrainable) return;
	float gsum = 0.0F;
	for (int z = 0; z < nz; ++z) gsum += part[static_cast<long long>(z) * kMlpTrainable + p];
	if (gradOut) { gradOut[p] = gsum; return; }
	const float m = b1 * mom[p] + (1.0F - b1) * gsum;
	const float v = b2 * vel[p] + (1.0F - b2) * gsum * gsum;
	mom[p] = m; vel[p] = v;
	prm[p] -= lr * (m / bc1) / (__builtin_sqrtf(v / bc2) + eps); }

}  // namespace rgl
