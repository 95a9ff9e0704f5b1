/*
 * Captures the x86 RCPPS approximation that rmlv::oneover() builds on
 * (reference src/rml/rmlv/rmlv_mvec4.hxx:630-650) and checks that a
 * 2048-entry mantissa table reproduces it for every 32-bit input.
 *
 *   gen_rcpps_table table <out.bin>   write 2048 little-endian u32 results
 *                                     of rcpps(1.0 + i/2048), i = 0..2047
 *   gen_rcpps_table check <in.bin> [stride]
 *                                     compare the table model with the live
 *                                     instruction over all inputs (stride 1)
 *                                     or every stride-th input
 *
 * The model: for a normal input with biased exponent e (1..254) and
 * mantissa m, the result has the sign of the input, exponent
 * 126 + 127 - e (flushed to signed zero when that is <= 0) and the
 * mantissa of table[m >> 12].  Zero and denormal inputs give signed
 * infinity, infinity gives signed zero, NaN gives the quieted NaN.
 *
 * The instruction is host-CPU dependent: the table committed with the
 * tests was captured on an Intel Xeon host.  Test infrastructure only.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static float hw_rcp(float x) { return _mm_cvtss_f32(_mm_rcp_ps(_mm_set1_ps(x))); }

static uint32_t model(const uint32_t* tab, uint32_t in) {
	uint32_t e = (in >> 23) & 0xffu, s = in & 0x80000000u, m = in & 0x7fffffu;
	if (e == 0xffu) return m ? (in | 0x400000u) : s;
	if (e == 0u) return s | 0x7f800000u;
	uint32_t t = tab[m >> 12];
	int re = (int)((t >> 23) & 0xffu) + 127 - (int)e;
	if (re <= 0) return s;
	if (re >= 255) return s | 0x7f800000u;
	return s | ((uint32_t)re << 23) | (t & 0x7fffffu); }

int main(int argc, char** argv) {
	if (argc < 3) { fprintf(stderr, "usage: %s table|check <file> [stride]\n", argv[0]); return 2; }
	uint32_t tab[2048];
	if (strcmp(argv[1], "table") == 0) {
		for (uint32_t i = 0; i < 2048; ++i) tab[i] = f2u(hw_rcp(u2f(0x3f800000u | (i << 12))));
		FILE* f = fopen(argv[2], "wb");
		if (!f || fwrite(tab, 4, 2048, f) != 2048) { perror("write"); return 1; }
		fclose(f);
		return 0; }
	if (strcmp(argv[1], "check") == 0) {
		FILE* f = fopen(argv[2], "rb");
		if (!f || fread(tab, 4, 2048, f) != 2048) { perror("read"); return 1; }
		fclose(f);
		uint64_t stride = argc > 3 ? strtoull(argv[3], NULL, 10) : 1;
		if (stride == 0) stride = 1;
		uint64_t bad = 0, n = 0;
		for (uint64_t u = 0; u < (1ull << 32); u += stride, ++n) {
			uint32_t in = (uint32_t)u;
			if (model(tab, in) != f2u(hw_rcp(u2f(in)))) ++bad; }
		printf("checked %llu inputs, %llu mismatches\n", (unsigned long long)n, (unsigned long long)bad);
		return bad ? 1 : 0; }
	return 2; }
