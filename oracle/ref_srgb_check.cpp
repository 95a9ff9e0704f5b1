// Exhaustive check of the oracle's sRGB encoder against the reference's own
// vendored encoder (3rdparty/ryg-srgb/ryg-srgb.h:183-219,
// float_to_srgb8_var2_SSE2), compiled here from the header where it lies.
// Test infrastructure: built by `make ref` into oracle/_ref/ and run by the
// CPU tests.  `ref_srgb_check [step]` visits every step-th group of four
// bit patterns (step 1: all 2^32).  Exit status 0 when all encode alike.
// `ref_srgb_check decode` prints the reference's sRGB8 -> float table
// (ryg-srgb.h:33-56, used by rglr_texture_load.cxx:55-64) as 256 hex words.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "ryg-srgb.h"

extern "C" uint32_t rglo_to_srgb8(float x);

int main(int argc, char** argv) {
	if (argc > 1 && std::strcmp(argv[1], "decode") == 0) {
		for (int i = 0; i < 256; ++i) {
			const float f = ryg::srgb8_to_float(static_cast<ryg::uint8>(i));
			uint32_t b; std::memcpy(&b, &f, 4); std::printf("%08x\n", b); }
		return 0; }
	const uint64_t step = 4ull * static_cast<uint64_t>(argc > 1 ? std::atoi(argv[1]) : 1);
	uint64_t bad = 0, n = 0;
	alignas(16) float in[4];
	alignas(16) uint32_t out[4];
	for (uint64_t u = 0; u < (1ull << 32); u += step) {
		for (int l = 0; l < 4; ++l) { const uint32_t b = static_cast<uint32_t>(u + l); std::memcpy(&in[l], &b, 4); }
		_mm_store_si128(reinterpret_cast<__m128i*>(out), ryg::float_to_srgb8_var2_SSE2(_mm_load_ps(in)));
		for (int l = 0; l < 4; ++l) {
			++n;
			if (out[l] != rglo_to_srgb8(in[l])) {
				if (bad < 8) std::printf("mismatch at 0x%08x: reference %u oracle %u\n",
				                         static_cast<uint32_t>(u + l), out[l], rglo_to_srgb8(in[l]));
				++bad; }}}
	std::printf("%llu inputs, %llu mismatches\n", static_cast<unsigned long long>(n), static_cast<unsigned long long>(bad));
	return bad ? 1 : 0; }
