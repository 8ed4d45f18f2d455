// gf64_host.cpp -- host-side GF(2^64) element arithmetic used for the small
// dense work that stays on the CPU (k x k pivoting in SolveAndReconstruct,
// gf64_solve, coefficient gathering).  PCLMULQDQ when the host has it,
// otherwise a portable shift-and-xor product; both reduce modulo
// x^64 + x^4 + x^3 + x + 1 (gf64/gf64_global.h:8).
#include "gf64_host.h"

#include <cstring>
#include <vector>

#if defined(__x86_64__)
#include <smmintrin.h>
#include <wmmintrin.h>
#endif

namespace parhip {

static inline uint64_t reduce128(uint64_t lo, uint64_t hi) {
	uint64_t t_lo = hi ^ (hi << 1) ^ (hi << 3) ^ (hi << 4);
	uint64_t t_hi = (hi >> 63) ^ (hi >> 61) ^ (hi >> 60);
	uint64_t t2 = t_hi ^ (t_hi << 1) ^ (t_hi << 3) ^ (t_hi << 4);
	return lo ^ t_lo ^ t2;
}

static uint64_t mul_portable(uint64_t a, uint64_t b) {
	uint64_t lo = 0, hi = 0;
	for (int i = 0; i < 64; i++) {
		uint64_t m = (uint64_t)0 - ((b >> i) & 1u);
		lo ^= (a << i) & m;
		if (i) hi ^= (a >> (64 - i)) & m;
	}
	return reduce128(lo, hi);
}

#if defined(__x86_64__)
__attribute__((target("pclmul,sse4.1"))) static uint64_t mul_pclmul(uint64_t a, uint64_t b) {
	__m128i p = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a),
	                                 _mm_cvtsi64_si128((long long)b), 0x00);
	uint64_t lo = (uint64_t)_mm_cvtsi128_si64(p);
	uint64_t hi = (uint64_t)_mm_extract_epi64(p, 1);
	return reduce128(lo, hi);
}
static const bool g_has_pclmul = __builtin_cpu_supports("pclmul");
#endif

uint64_t host_gf64_mul(uint64_t a, uint64_t b) {
#if defined(__x86_64__)
	if (g_has_pclmul) return mul_pclmul(a, b);
#endif
	return mul_portable(a, b);
}

static int deg64(uint64_t v) { return v ? 63 - __builtin_clzll(v) : -1; }

uint64_t host_gf64_inv(uint64_t a) {
	if (a <= 1) return a;
	uint64_t r0 = 0x1Bu, r1 = a, s0 = 0, s1 = 1;
	int d0 = 64, d1 = deg64(r1);
	while (r1 > 1) {
		if (d0 < d1) {
			uint64_t t = r0; r0 = r1; r1 = t;
			t = s0; s0 = s1; s1 = t;
			int d = d0; d0 = d1; d1 = d;
		}
		int sh = d0 - d1;
		r0 ^= r1 << sh;
		s0 ^= s1 << sh;
		d0 = deg64(r0);
	}
	return r1 == 0 ? 0 : s1;
}

// Gauss-Jordan on A (n x n) with the reference's pivot rule
// (src/par3_engine.cc:989-1045): for each column the first row at or below the
// diagonal with a nonzero entry, row swap, scale to 1, eliminate every other
// row.  The same row operations are applied to E (initially the identity), so
// E ends as the transform that the reference applies to its RHS blocks:
// E = A^-1 on success, or the partial transform when it stops on a singular
// column.  Returns 0 or -1.
int host_gauss_jordan(uint64_t* A, uint64_t* E, size_t n) {
	for (size_t i = 0; i < n; i++)
		for (size_t j = 0; j < n; j++) E[i * n + j] = (i == j) ? 1u : 0u;
	for (size_t col = 0; col < n; col++) {
		size_t p = col;
		while (p < n && A[p * n + col] == 0) p++;
		if (p == n) return -1;
		if (p != col) {
			for (size_t j = 0; j < n; j++) {
				std::swap(A[col * n + j], A[p * n + j]);
				std::swap(E[col * n + j], E[p * n + j]);
			}
		}
		uint64_t pv = A[col * n + col];
		uint64_t pinv = host_gf64_inv(pv);
		if (pv != 1) {
			for (size_t j = 0; j < n; j++) {
				A[col * n + j] = host_gf64_mul(A[col * n + j], pinv);
				E[col * n + j] = host_gf64_mul(E[col * n + j], pinv);
			}
		}
		for (size_t row = 0; row < n; row++) {
			if (row == col) continue;
			uint64_t f = A[row * n + col];
			if (f == 0) continue;
			for (size_t j = 0; j < n; j++) {
				A[row * n + j] ^= host_gf64_mul(f, A[col * n + j]);
				E[row * n + j] ^= host_gf64_mul(f, E[col * n + j]);
			}
		}
	}
	return 0;
}

}  // namespace parhip
