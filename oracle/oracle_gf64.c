/*
 * oracle_gf64.c -- CPU restatement of the reference's GF(2^64) recovery path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  The product never links this file.
 *
 * Field: GF(2)[x] / (x^64 + x^4 + x^3 + x + 1)   (gf64/gf64_global.h:8)
 * Words: slices are arrays of native little-endian uint64 words.
 *
 * Pins (tests/test_oracle.py):
 *   or_gf64_mul      - KATs from test/gf64-js.test.js:286-296; equals the
 *                      independent Russian-peasant multiply on random pairs;
 *                      field axioms (distributivity, associativity).
 *   or_gf64_inv      - inv(0)=0, inv(1)=1 (gf64/test/test.c:112-120); equals
 *                      the independent Fermat inverse a^(2^64-2); a*inv(a)=1.
 *   or_gf64_build_cauchy / or_gf64_compute_recovery
 *                    - test/par3-golden.bin CRC32 6A941C15 (test/par3-golden.js:35)
 *                      and the KAT 0x7ffffffffffffffa (test/par3-native-fidelity.js:240-244);
 *                      byte-equal to the reference engine built in oracle/_ref.
 *   region ops       - linearity and the reference's scalar semantics, byte-equal
 *                      to oracle/_ref's gf64_region_*_scalar functions
 *                      (test_region_ops_equal_reference_scalar).
 *   solve            - repaired blocks equal the dropped originals (round trip).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

/* Carry-less 64x64 -> 128 product, plain shift-and-xor.
 * Restates the schoolbook product of gf64/gf64_single.c:6-33 (there via PCLMUL). */
void or_gf64_clmul(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) {
	uint64_t l = 0, h = 0;
	for (int i = 0; i < 64; i++) {
		uint64_t m = (uint64_t)0 - ((b >> i) & 1u);
		l ^= (a << i) & m;
		if (i) h ^= (a >> (64 - i)) & m;
	}
	*lo = l;
	*hi = h;
}

/* Reduction of hi*x^64 + lo modulo x^64+x^4+x^3+x+1.
 * Restates gf64/gf64_single.c:35-52: x^64 == 0x1B, the 4 bits that overflow
 * hi*0x1B past bit 63 are folded a second time. */
uint64_t or_gf64_reduce(uint64_t lo, uint64_t hi) {
	uint64_t t_lo = hi ^ (hi << 1) ^ (hi << 3) ^ (hi << 4);
	uint64_t t_hi = (hi >> 63) ^ (hi >> 61) ^ (hi >> 60);
	uint64_t t2 = t_hi ^ (t_hi << 1) ^ (t_hi << 3) ^ (t_hi << 4);
	return lo ^ t_lo ^ t2;
}

/* gf64_mul_reference (gf64/gf64_single.c:6-53). */
uint64_t or_gf64_mul(uint64_t a, uint64_t b) {
	uint64_t lo, hi;
	or_gf64_clmul(a, b, &lo, &hi);
	return or_gf64_reduce(lo, hi);
}

/* Independent method (not a restatement): Russian-peasant multiply with the
 * reduction interleaved (multiply-by-x then conditional XOR of 0x1B).  Used
 * only to pin or_gf64_mul. */
uint64_t or_gf64_mul_peasant(uint64_t a, uint64_t b) {
	uint64_t r = 0;
	while (b) {
		if (b & 1) r ^= a;
		b >>= 1;
		uint64_t carry = a >> 63;
		a <<= 1;
		if (carry) a ^= 0x1Bu;
	}
	return r;
}

static int deg64(uint64_t v) { return v ? 63 - __builtin_clzll(v) : -1; }

/* gf64_inverse (gf64/gf64_invert.c:29-67): degree-tracking extended Euclid
 * with the implicit x^64 term of the modulus; inv(0)=0, inv(1)=1. */
uint64_t or_gf64_inv(uint64_t a) {
	if (a == 0) return 0;
	if (a == 1) return 1;
	uint64_t r0 = 0x1Bu, r1 = a, s0 = 0, s1 = 1;
	int d0 = 64, d1 = deg64(r1);
	while (r1 != 1 && r1 != 0) {
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

/* Independent method: Fermat, a^(2^64-2) by square-and-multiply. */
uint64_t or_gf64_inv_fermat(uint64_t a) {
	if (a == 0) return 0;
	uint64_t result = 1, base = a;
	uint64_t e = 0xFFFFFFFFFFFFFFFEull;
	while (e) {
		if (e & 1) result = or_gf64_mul(result, base);
		base = or_gf64_mul(base, base);
		e >>= 1;
	}
	return result;
}

/* GF64Controller::BuildCauchyMatrix (src/par3_engine.cc:358-429):
 * M[r][c] = inv((firstInput + c) XOR (firstRecovery + r)), a zero
 * denominator replaced by 1 before inversion (:392-394). Row-major R x N. */
void or_gf64_build_cauchy(uint64_t *out, size_t numInputs, size_t numRecovery,
                          uint64_t firstInput, uint64_t firstRecovery) {
	for (size_t r = 0; r < numRecovery; r++) {
		uint64_t y = firstRecovery + r;
		for (size_t c = 0; c < numInputs; c++) {
			uint64_t d = (firstInput + c) ^ y;
			out[r * numInputs + c] = or_gf64_inv(d ? d : 1);
		}
	}
}

/* gf64_region_mul_scalar (gf64/gf64_region_scalar.c:8-12). */
void or_gf64_region_mul(uint64_t *out, const uint64_t *in, size_t len, uint64_t c) {
	for (size_t i = 0; i < len; i++) out[i] = or_gf64_mul(in[i], c);
}

/* gf64_region_mul_scalar_arr (gf64/gf64_region_scalar.c:14-18): coefficient
 * index cycles, coeff[i % n_coeff]. */
void or_gf64_region_mul_arr(uint64_t *out, const uint64_t *in, const uint64_t *coeff,
                            size_t len, size_t n_coeff) {
	for (size_t i = 0; i < len; i++) out[i] = or_gf64_mul(in[i], coeff[i % n_coeff]);
}

/* gf64_region_muladd_scalar_arr (gf64/gf64_region_scalar.c:20-28): every
 * output word receives the XOR over ALL n_coeff products of the same input word. */
void or_gf64_region_muladd_arr(uint64_t *out, const uint64_t *in, const uint64_t *coeff,
                               size_t len, size_t n_coeff) {
	for (size_t i = 0; i < len; i++) {
		uint64_t s = 0;
		for (size_t c = 0; c < n_coeff; c++) s ^= or_gf64_mul(in[i], coeff[c]);
		out[i] ^= s;
	}
}

/* gf64_region_coupled_muladd_scalar_arr (gf64/gf64_region_scalar.c:45-55). */
void or_gf64_region_coupled_muladd_arr(uint64_t *out, const uint64_t *const *in_blocks,
                                       const uint64_t *coeff_blocks, size_t len, size_t G) {
	for (size_t g = 0; g < G; g++)
		for (size_t i = 0; i < len; i++) out[i] ^= or_gf64_mul(in_blocks[g][i], coeff_blocks[g]);
}

/* gf64_region_fused_output_muladd_scalar_arr (gf64/gf64_region_scalar.c:58-72):
 * outs[k][w] ^= in[w] * (*coeff_ptrs[k]). */
void or_gf64_region_fused_output_muladd_arr(uint64_t *const *outs, const uint64_t *in,
                                            const uint64_t *const *coeff_ptrs, size_t len, size_t K) {
	for (size_t i = 0; i < len; i++)
		for (size_t k = 0; k < K; k++) outs[k][i] ^= or_gf64_mul(in[i], *coeff_ptrs[k]);
}

/* gf64_region_2d_muladd_scalar_arr (gf64/gf64_region_scalar.c:75-91):
 * outs[k][w] ^= XOR_g in_blocks[g][w] * coeff2d[k*K_stride + g]. */
void or_gf64_region_2d_muladd_arr(uint64_t *const *outs, size_t K, const uint64_t *const *in_blocks,
                                  size_t G, const uint64_t *coeff2d, size_t K_stride, size_t len) {
	for (size_t g = 0; g < G; g++)
		for (size_t i = 0; i < len; i++)
			for (size_t k = 0; k < K; k++)
				outs[k][i] ^= or_gf64_mul(in_blocks[g][i], coeff2d[k * K_stride + g]);
}

/* ComputeRecoveryBlocksWithCoeff (src/par3_engine.cc:1165-1382) value semantics:
 * accumulate=false zero-fills the outputs first, accumulate=true XORs into them;
 * recovery[r][w] ^= XOR_i coeff[r*N+i] * inputs[i][w].  The reference's thread
 * decompositions only reorder XORs, so the result is this sum. */
void or_gf64_compute_recovery_with_coeff(const uint64_t *inputs, size_t numInputs,
                                         uint64_t *recovery, size_t numRecovery,
                                         size_t blockSize64, const uint64_t *coeff,
                                         int accumulate) {
	if (numInputs == 0 || numRecovery == 0) return;  /* par3_engine.cc:1173-1176 */
	if (!accumulate) memset(recovery, 0, numRecovery * blockSize64 * sizeof(uint64_t));
	for (size_t r = 0; r < numRecovery; r++) {
		uint64_t *out = recovery + r * blockSize64;
		for (size_t i = 0; i < numInputs; i++) {
			uint64_t c = coeff[r * numInputs + i];
			const uint64_t *in = inputs + i * blockSize64;
			for (size_t w = 0; w < blockSize64; w++) out[w] ^= or_gf64_mul(in[w], c);
		}
	}
}

/* ComputeRecoveryBlocks (src/par3_engine.cc:1061-1111): Cauchy build then the
 * coefficient contraction. */
void or_gf64_compute_recovery(const uint64_t *inputs, size_t numInputs,
                              uint64_t *recovery, size_t numRecovery, size_t blockSize64,
                              uint64_t firstInput, uint64_t firstRecovery, int accumulate) {
	if (numInputs == 0 || numRecovery == 0) return;
	uint64_t *m = (uint64_t *)malloc(numInputs * numRecovery * sizeof(uint64_t));
	if (!m) return;
	or_gf64_build_cauchy(m, numInputs, numRecovery, firstInput, firstRecovery);
	or_gf64_compute_recovery_with_coeff(inputs, numInputs, recovery, numRecovery, blockSize64, m,
	                                    accumulate);
	free(m);
}

/* The contraction evaluated only at the word columns cols[]; each column is
 * independent of the others, so this is exactly the full result's columns. */
void or_gf64_recovery_columns(const uint64_t *inputs, size_t numInputs, size_t blockSize64,
                              const uint64_t *coeff, size_t numRecovery,
                              const uint64_t *cols, size_t ncols, uint64_t *out) {
	for (size_t r = 0; r < numRecovery; r++)
		for (size_t j = 0; j < ncols; j++) {
			uint64_t s = 0;
			for (size_t i = 0; i < numInputs; i++)
				s ^= or_gf64_mul(inputs[i * blockSize64 + cols[j]], coeff[r * numInputs + i]);
			out[r * ncols + j] = s;
		}
}

/* GF64Controller::SolveAndReconstruct (src/par3_engine.cc:980-1048): in-place
 * Gauss-Jordan on A (n x n) with the first-nonzero pivot rule and row swaps,
 * the same row operations applied to the n RHS blocks.  Returns 0, or -1 when
 * a column has no nonzero pivot (singular). */
int or_gf64_solve_and_reconstruct(uint64_t *A, uint64_t *rhs, size_t n, size_t bs) {
	if (n == 0) return 0;
	for (size_t col = 0; col < n; col++) {
		size_t p = col;
		while (p < n && A[p * n + col] == 0) p++;
		if (p == n) return -1;
		if (p != col) {
			for (size_t j = 0; j < n; j++) {
				uint64_t t = A[col * n + j]; A[col * n + j] = A[p * n + j]; A[p * n + j] = t;
			}
			for (size_t w = 0; w < bs; w++) {
				uint64_t t = rhs[col * bs + w]; rhs[col * bs + w] = rhs[p * bs + w]; rhs[p * bs + w] = t;
			}
		}
		uint64_t pv = A[col * n + col];
		uint64_t pinv = or_gf64_inv(pv);
		if (pv != 1) {
			for (size_t j = 0; j < n; j++) A[col * n + j] = or_gf64_mul(A[col * n + j], pinv);
			for (size_t w = 0; w < bs; w++) rhs[col * bs + w] = or_gf64_mul(rhs[col * bs + w], pinv);
		}
		for (size_t row = 0; row < n; row++) {
			if (row == col) continue;
			uint64_t f = A[row * n + col];
			if (f == 0) continue;
			for (size_t j = 0; j < n; j++) A[row * n + j] ^= or_gf64_mul(f, A[col * n + j]);
			for (size_t w = 0; w < bs; w++) rhs[row * bs + w] ^= or_gf64_mul(rhs[col * bs + w], f);
		}
	}
	return 0;
}

/* gf64_solve (gf64/gf64_solve.c:39-96): Gauss-Jordan on [A|b]; A and b are
 * not modified. Returns 0 or -1 (singular / allocation failure). */
int or_gf64_solve(const uint64_t *A, const uint64_t *b, uint64_t *x, size_t n) {
	if (n == 0) return 0;
	uint64_t *Ac = (uint64_t *)malloc(n * n * sizeof(uint64_t));
	uint64_t *bc = (uint64_t *)malloc(n * sizeof(uint64_t));
	if (!Ac || !bc) { free(Ac); free(bc); return -1; }
	memcpy(Ac, A, n * n * sizeof(uint64_t));
	memcpy(bc, b, n * sizeof(uint64_t));
	int rc = or_gf64_solve_and_reconstruct(Ac, bc, n, 1);
	if (rc == 0) memcpy(x, bc, n * sizeof(uint64_t));
	free(Ac);
	free(bc);
	return rc;
}

/* CRC-32 (IEEE, reflected 0xEDB88320), as test/par3-golden.js:43-58 computes it. */
uint32_t or_crc32(const uint8_t *buf, size_t len) {
	uint32_t crc = 0xFFFFFFFFu;
	for (size_t i = 0; i < len; i++) {
		crc ^= buf[i];
		for (int k = 0; k < 8; k++) crc = (crc & 1) ? (0xEDB88320u ^ (crc >> 1)) : (crc >> 1);
	}
	return crc ^ 0xFFFFFFFFu;
}
