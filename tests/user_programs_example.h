// A program added at build time without editing the library's sources:
//   python -m rsr_amd.build --programs tests/user_programs_example.h
// It is written to the program concept of rgl_programs.h: a flat colour
// from uniforms [0..2], scaled per vertex by slot 6 (default 1).
#pragma once

namespace rgl {

struct ProgFlatScaled {
	static constexpr int kId = 200, kVary = 1, kUV = -1;
	static constexpr bool kEarlyZ = true, kDiscard = false, kInstanced = false;
	__device__ static void vertex(const DrawDesc& d, int, int idx, float pos[4], float* vo) {
		float p[4]; load_pos(d, idx, p);
		xform4(d.vpm, p, pos);
		vo[0] = buf_at(d.buffers[6], idx, 1.0F); }
	__device__ static bool fragment(const DrawDesc& d, const float* v, int, const FragCoord&, float color[4]) {
		color[0] = d.uniforms[0] * v[0]; color[1] = d.uniforms[1] * v[0]; color[2] = d.uniforms[2] * v[0];
		color[3] = 1.0F;
		return true; }};

using UserPrograms = ProgramList<ProgFlatScaled>;

}  // namespace rgl
