// gf64_dev.h -- GF(2^64) element arithmetic for gfx950 device code.
//
// Field GF(2)[x]/(x^64 + x^4 + x^3 + x + 1), the PAR3 field
// (gf64/gf64_global.h:8).  CDNA has no carry-less multiply instruction, so the
// element multiply here is a 4-bit-window software product; the hot region
// kernels never call it per word -- they use per-coefficient split tables or a
// GF(2) bit-matrix on MFMA instead (gf64_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace parhip {

// hi*x^64 + lo reduced: x^64 == x^4+x^3+x+1; the <= 4 bits that hi*0x1B
// pushes past bit 63 are folded once more (same math as gf64/gf64_single.c:35-52).
__host__ __device__ __forceinline__ uint64_t gf64_reduce(uint64_t lo, uint64_t hi) {
	uint64_t t_lo = hi ^ (hi << 1) ^ (hi << 3) ^ (hi << 4);
	uint64_t t_hi = (hi >> 63) ^ (hi >> 61) ^ (hi >> 60);
	uint64_t t2 = t_hi ^ (t_hi << 1) ^ (t_hi << 3) ^ (t_hi << 4);
	return lo ^ t_lo ^ t2;
}

// c * x^j mod p for 0 <= j < 64.
__host__ __device__ __forceinline__ uint64_t gf64_mulx_pow(uint64_t c, unsigned j) {
	if (j == 0) return c;
	return gf64_reduce(c << j, c >> (64u - j));
}

// c * x^j mod p for 0 <= j < 64 without a branch on j (the high part is
// shifted in two steps so that j = 0 gives 0 instead of an undefined shift).
__host__ __device__ __forceinline__ uint64_t gf64_mulx_pow_any(uint64_t c, unsigned j) {
	return gf64_reduce(c << j, (c >> 1) >> (63u - j));
}

// c * x^4: the 4 bits shifted out fold back as h * (x^4 + x^3 + x + 1), at
// most 8 bits, so one fold suffices.
__host__ __device__ __forceinline__ uint64_t gf64_mulx4(uint64_t c) {
	const uint64_t h = c >> 60;
	return (c << 4) ^ h ^ (h << 1) ^ (h << 3) ^ (h << 4);
}

// c * n(x) for a 4-bit polynomial n: at most 3 bits overflow, folded once.
__host__ __device__ __forceinline__ uint64_t gf64_mul_nibble(uint64_t c, unsigned n) {
	const uint64_t m0 = 0 - (uint64_t)(n & 1u), m1 = 0 - (uint64_t)((n >> 1) & 1u);
	const uint64_t m2 = 0 - (uint64_t)((n >> 2) & 1u), m3 = 0 - (uint64_t)((n >> 3) & 1u);
	const uint64_t lo = (c & m0) ^ ((c << 1) & m1) ^ ((c << 2) & m2) ^ ((c << 3) & m3);
	const uint64_t h = ((c >> 63) & m1) ^ ((c >> 62) & m2) ^ ((c >> 61) & m3);
	return lo ^ h ^ (h << 1) ^ (h << 3) ^ (h << 4);
}

// Full element product a*b: 4-bit windows over b, 16 precomputed multiples of a
// kept as (lo, hi) 128-bit pairs, then one reduction.
__host__ __device__ inline uint64_t gf64_mul(uint64_t a, uint64_t b) {
	uint64_t lo = 0, hi = 0;
#pragma unroll 1
	for (int k = 60; k >= 0; k -= 4) {
		// (hi:lo) <<= 4
		hi = (hi << 4) | (lo >> 60);
		lo <<= 4;
		unsigned n = (unsigned)(b >> k) & 15u;
		// n * a carry-less: XOR of a<<t for the set bits t of n
		uint64_t plo = 0, phi = 0;
#pragma unroll
		for (int t = 0; t < 4; t++) {
			uint64_t m = (uint64_t)0 - (uint64_t)((n >> t) & 1u);
			plo ^= (a << t) & m;
			if (t) phi ^= (a >> (64 - t)) & m;
		}
		lo ^= plo;
		hi ^= phi;
	}
	return gf64_reduce(lo, hi);
}

__host__ __device__ __forceinline__ int gf64_deg(uint64_t v) {
#ifdef __HIP_DEVICE_COMPILE__
	return v ? 63 - __clzll(v) : -1;
#else
	return v ? 63 - __builtin_clzll(v) : -1;
#endif
}

// Multiplicative inverse by the degree-tracking extended Euclidean algorithm
// with the implicit x^64 term of the modulus; inv(0)=0, inv(1)=1
// (same contract as gf64/gf64_invert.c:29-67; the inverse is unique).
__host__ __device__ inline uint64_t gf64_inv(uint64_t a) {
	if (a <= 1) return a;
	uint64_t r0 = 0x1Bu, r1 = a, s0 = 0, s1 = 1;
	int d0 = 64, d1 = gf64_deg(r1);
	while (r1 > 1) {
		if (d0 < d1) {
			uint64_t t = r0; r0 = r1; r1 = t;
			t = s0; s0 = s1; s1 = t;
			int d = d0; d0 = d1; d1 = d;
		}
		int sh = d0 - d1;
		r0 ^= r1 << sh;
		s0 ^= s1 << sh;
		d0 = gf64_deg(r0);
	}
	return r1 == 0 ? 0 : s1;
}

}  // namespace parhip
