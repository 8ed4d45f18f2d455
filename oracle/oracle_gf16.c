/*
 * oracle_gf16.c -- CPU restatement of the reference's PAR2 GF(2^16) muladd.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  The product never links this file.
 *
 * Field: GF(2)[x] / 0x1100B, generator 2 (gf16/gf16_global.h:9).
 * Words: 16-bit little-endian; an odd trailing byte is a word whose high byte
 * is zero (test/gf16/test-ctrl.cpp:195-200).
 *
 * Pins (tests/test_oracle.py):
 *   or_gf16_mul       - equals the independent Russian-peasant multiply on all
 *                       pairs of a random sample and on every a*1, a*2; the
 *                       exp table has order exactly 65535 (2 is a generator).
 *   or_gf16_input_log - the PAR2 specification's constant sequence: the first
 *                       values 1,2,4,7,8,11,13,14,16,19,22,23 and every value
 *                       coprime to 3,5,17,257 in increasing order.
 *   or_gf16_coeff     - exponent 0 gives 1 for every input (plain XOR parity,
 *                       gf16/controller_cpu.cpp:495-513); coeff(i,e) equals
 *                       exp(log_i)^e by repeated multiplication; equals the
 *                       reference's gfmat_coeff compiled in oracle/_ref.
 *   or_par2_create    - restates the controller test's reference computation
 *                       (test/gf16/test-ctrl.cpp:185-202) over the pinned
 *                       multiply and coefficients; its odd-tail rule is
 *                       checked in test_par2_odd_tail.
 *   or_par2_recovery_columns
 *                     - equals or_par2_create on every column of small cases
 *                       (test_par2_columns_equal_create).
 */
#include "oracle.h"
#include <string.h>

static uint16_t g_exp[65536];
static uint16_t g_log[65536];
static uint16_t g_inlog[32768];
static int g_ready = 0;

static void gf16_init(void) {
	if (g_ready) return;
	/* exp/log tables of generator 2 modulo 0x1100B */
	uint32_t v = 1;
	for (uint32_t e = 0; e < 65535; e++) {
		g_exp[e] = (uint16_t)v;
		g_log[v] = (uint16_t)e;
		v <<= 1;
		if (v & 0x10000u) v ^= 0x1100Bu;
	}
	g_exp[65535] = g_exp[0];
	/* PAR2 input constants (gf16/gfmat_coeff.c:20-32): the i-th exponent,
	 * counting from 1, that is divisible by none of 3, 5, 17, 257. */
	uint32_t e = 0;
	for (uint32_t i = 0; i < 32768; i++) {
		do { e++; } while (e % 3 == 0 || e % 5 == 0 || e % 17 == 0 || e % 257 == 0);
		g_inlog[i] = (uint16_t)e;
	}
	g_ready = 1;
}

uint16_t or_gf16_exp(uint32_t e) { gf16_init(); return g_exp[e % 65535u]; }
uint16_t or_gf16_log(uint16_t v) { gf16_init(); return g_log[v]; }

/* Log/antilog multiply, as the reference's own test reference does
 * (test/gf16/test.h:59-76). */
uint16_t or_gf16_mul(uint16_t a, uint16_t b) {
	gf16_init();
	if (a == 0 || b == 0) return 0;
	return g_exp[((uint32_t)g_log[a] + g_log[b]) % 65535u];
}

/* Independent method: shift-and-add with interleaved reduction (pins or_gf16_mul). */
uint16_t or_gf16_mul_peasant(uint16_t a, uint16_t b) {
	uint32_t r = 0, x = a;
	while (b) {
		if (b & 1) r ^= x;
		b >>= 1;
		x <<= 1;
		if (x & 0x10000u) x ^= 0x1100Bu;
	}
	return (uint16_t)r;
}

/* gfmat_input_log (gf16/gfmat_coeff.c:80-82). */
uint16_t or_gf16_input_log(uint32_t inputNum) { gf16_init(); return g_inlog[inputNum & 32767u]; }

/* gfmat_coeff (gf16/gfmat_coeff.c:84-101): exp((log_i * e) mod 65535). */
uint16_t or_gf16_coeff(uint32_t inputNum, uint32_t exponent) {
	gf16_init();
	uint64_t l = (uint64_t)g_inlog[inputNum & 32767u] * exponent;
	return g_exp[l % 65535u];
}

/* out ^= coeff * in over 16-bit LE words; odd trailing byte zero-extended.
 * (test/gf16/test-ctrl.cpp:188-202) */
void or_gf16_muladd_region(uint8_t *out, const uint8_t *in, size_t bytes, uint16_t coeff) {
	size_t words = bytes / 2;
	for (size_t w = 0; w < words; w++) {
		uint16_t d = (uint16_t)(in[2 * w] | (in[2 * w + 1] << 8));
		uint16_t p = or_gf16_mul(d, coeff);
		out[2 * w] ^= (uint8_t)p;
		out[2 * w + 1] ^= (uint8_t)(p >> 8);
	}
	if (bytes & 1) {
		uint16_t p = or_gf16_mul(in[bytes - 1], coeff);
		out[bytes - 1] ^= (uint8_t)p;
		out[bytes] ^= (uint8_t)(p >> 8);
	}
}

/* PAR2 create value semantics (test/gf16/test-ctrl.cpp:188-202; PAR2Proc
 * first-batch-writes rule controller_ocl.cpp:626-634): every output slice
 * (sliceSize bytes, zero-initialised) is XOR_i coeff(inputNum_i, exp_e) * in_i,
 * where input i has sizes[i] <= sliceSize bytes. */
void or_par2_create(const uint8_t *const *inputs, const size_t *sizes, const uint16_t *inputNums,
                    size_t numInputs, uint8_t *const *outputs, const uint16_t *exponents,
                    size_t numOutputs, size_t sliceSize) {
	for (size_t o = 0; o < numOutputs; o++) {
		memset(outputs[o], 0, sliceSize);
		for (size_t i = 0; i < numInputs; i++)
			or_gf16_muladd_region(outputs[o], inputs[i], sizes[i],
			                      or_gf16_coeff(inputNums[i], exponents[o]));
	}
}

/* The PAR2 contraction at the 16-bit word columns cols[] only (columns are
 * independent); inputs contiguous numInputs x sliceSize bytes, out numOutputs x ncols. */
void or_par2_recovery_columns(const uint8_t *inputs, size_t numInputs, size_t sliceSize,
                              const uint16_t *inputNums, const uint16_t *exponents,
                              size_t numOutputs, const uint64_t *cols, size_t ncols,
                              uint16_t *out) {
	for (size_t o = 0; o < numOutputs; o++)
		for (size_t j = 0; j < ncols; j++) {
			uint16_t s = 0;
			for (size_t i = 0; i < numInputs; i++) {
				const uint8_t *p = inputs + i * sliceSize + 2 * cols[j];
				uint16_t d = (uint16_t)(p[0] | (p[1] << 8));
				s ^= or_gf16_mul(d, or_gf16_coeff(inputNums[i], exponents[o]));
			}
			out[o * ncols + j] = s;
		}
}

/* ---------------------------------------------------------------------------
 * PAR2 repair matrix: Galois16RecMatrix::Compute (gf16/gfmat_inv.cpp:562-676)
 * with Construct (:462-560) and the in-place Gauss-Jordan of scaleRows /
 * applyRows (:130-300), one row at a time.
 *
 * Columns: the valid inputs in input order, then the missing inputs in input
 * order (Construct's validCol / missingCol).  Row r holds recovery exponent
 * recovery[r] (sorted ascending first).  Pivot r sits at column validCount + r
 * and is never swapped; its entry is replaced by 1 before the row is scaled by
 * the pivot's reciprocal, and every other row's entry in that column is
 * replaced by 0 before the row gets factor * pivot row (REPLACE_WORD), which
 * leaves the inverse of the missing block in the missing columns and
 * inverse * valid block in the valid columns.  A zero pivot (PAR2's
 * non-invertible combinations) drops that recovery exponent and restarts; too
 * few exponents left -> 0.  On success the exponents are the numRec used and
 * out[r * numInputs + col] is GetFactor(col, r).
 * ------------------------------------------------------------------------- */
int or_gf16_rec_matrix_compute(const uint8_t *inputValid, size_t numInputs, size_t validCount,
                               uint16_t *recovery, size_t *recoveryCount, uint16_t *out) {
	const size_t numRec = numInputs - validCount;
	size_t cnt = *recoveryCount;
	/* std::sort(recovery) (gfmat_inv.cpp:594) */
	for (size_t i = 1; i < cnt; i++) {
		uint16_t v = recovery[i];
		size_t j = i;
		while (j > 0 && recovery[j - 1] > v) {
			recovery[j] = recovery[j - 1];
			j--;
		}
		recovery[j] = v;
	}
	for (;;) {
		if (numRec > cnt) {
			*recoveryCount = cnt;
			return 0;
		}
		/* Construct */
		size_t validCol = 0, missingCol = validCount;
		for (size_t i = 0; i < numInputs; i++) {
			const size_t col = inputValid[i] ? validCol++ : missingCol++;
			for (size_t r = 0; r < numRec; r++)
				out[r * numInputs + col] = or_gf16_coeff((uint32_t)i, recovery[r]);
		}
		/* eliminate */
		long bad = -1;
		for (size_t k = 0; k < numRec && bad < 0; k++) {
			uint16_t *row = out + k * numInputs;
			const size_t pc = validCount + k;
			const uint16_t p = row[pc];
			if (p == 0) {
				bad = (long)k;
				break;
			}
			row[pc] = 1;
			const uint16_t ip = or_gf16_exp(65535u - or_gf16_log(p));
			for (size_t c = 0; c < numInputs; c++) row[c] = or_gf16_mul(row[c], ip);
			for (size_t r = 0; r < numRec; r++) {
				if (r == k) continue;
				uint16_t *dst = out + r * numInputs;
				const uint16_t t = dst[pc];
				dst[pc] = 0;
				if (t == 0) continue;
				for (size_t c = 0; c < numInputs; c++) dst[c] ^= or_gf16_mul(t, row[c]);
			}
		}
		if (bad < 0) break;
		/* recovery.erase(recovery.begin() + badRow); goto invert_loop (:638-646) */
		for (size_t i = (size_t)bad; i + 1 < cnt; i++) recovery[i] = recovery[i + 1];
		cnt--;
	}
	*recoveryCount = numRec; /* recovery.resize(numRec) */
	return 1;
}
