/*
 * Captures the x86 RSQRTPS approximation that rmlv::normalize() uses
 * (reference src/rml/rmlv/rmlv_soa.hxx:243-248, rmlv_mvec4.hxx:499-500)
 * and checks that a 2 x 2048-entry mantissa table reproduces it for every
 * 32-bit input.
 *
 *   gen_rsqrtps_table table <out.bin>   write 4096 little-endian u32: the
 *                                       results for 1 + i/2048 (odd biased
 *                                       exponent), then 2 + i/1024 (even)
 *   gen_rsqrtps_table check <in.bin> [stride]
 *                                       compare the model with the live
 *                                       instruction over all inputs (stride
 *                                       1) or every stride-th input
 *
 * The model: for a positive normal input with biased exponent e and
 * mantissa m, the result takes the mantissa of table[e even][m >> 12] and
 * that entry's exponent minus (e - 127) / 2 for odd e, (e - 128) / 2 for
 * even e.  Zero and denormal inputs give infinity of the input's sign,
 * other negative inputs and negative infinity the default NaN, +infinity
 * gives +0, NaN the quieted NaN.
 *
 * The instruction is host-CPU dependent: the committed table was captured
 * on an Intel Xeon host.  Test infrastructure only.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static float hw_rsqrt(float x) { return _mm_cvtss_f32(_mm_rsqrt_ps(_mm_set1_ps(x))); }

static uint32_t model(const uint32_t* tab, uint32_t in) {
	const uint32_t e = (in >> 23) & 0xffu, s = in & 0x80000000u, m = in & 0x7fffffu;
	if (e == 0xffu) {
		if (m) return in | 0x400000u;
		return s ? 0xffc00000u : 0u; }
	if (e == 0u) return s | 0x7f800000u;
	if (s) return 0xffc00000u;
	const int even = (e & 1u) == 0u;
	const uint32_t t = tab[(even ? 2048 : 0) + (m >> 12)];
	const int re = (int)((t >> 23) & 0xffu) - (even ? ((int)e - 128) / 2 : ((int)e - 127) / 2);
	return ((uint32_t)re << 23) | (t & 0x7fffffu); }

int main(int argc, char** argv) {
	if (argc < 3) { fprintf(stderr, "usage: %s table|check <file> [stride]\n", argv[0]); return 2; }
	uint32_t tab[4096];
	if (strcmp(argv[1], "table") == 0) {
		for (uint32_t i = 0; i < 2048; ++i) {
			tab[i] = f2u(hw_rsqrt(u2f(0x3f800000u | (i << 12))));           /* [1, 2), e = 127 */
			tab[2048 + i] = f2u(hw_rsqrt(u2f(0x40000000u | (i << 12)))); }  /* [2, 4), e = 128 */
		FILE* f = fopen(argv[2], "wb");
		if (!f || fwrite(tab, 4, 4096, f) != 4096) { perror("write"); return 1; }
		fclose(f);
		return 0; }
	if (strcmp(argv[1], "check") == 0) {
		FILE* f = fopen(argv[2], "rb");
		if (!f || fread(tab, 4, 4096, f) != 4096) { perror("read"); return 1; }
		fclose(f);
		uint64_t stride = argc > 3 ? strtoull(argv[3], NULL, 10) : 1;
		if (stride == 0) stride = 1;
		uint64_t bad = 0, n = 0;
		for (uint64_t u = 0; u < (1ull << 32); u += stride, ++n) {
			const uint32_t in = (uint32_t)u;
			const uint32_t want = f2u(hw_rsqrt(u2f(in)));
			if (model(tab, in) != want) {
				if (bad < 8) printf("mismatch %08x: model %08x hw %08x\n", in, model(tab, in), want);
				++bad; }}
		printf("checked %llu inputs, %llu mismatches\n", (unsigned long long)n, (unsigned long long)bad);
		return bad ? 1 : 0; }
	return 2; }
