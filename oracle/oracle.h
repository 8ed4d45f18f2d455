/*
 * oracle.h -- CPU restatement of the reference's recovery-slice arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (parparpar_amd/, include/)
 * includes, links or calls this code.  Only tests/, the smoke entry point and
 * bench.py's cpu_baseline leg load it, and only to check results.
 *
 * Every function cites the reference file:line it restates.  The pins that
 * check this restatement (golden vectors, KATs, independent methods, brute
 * force) live in tests/test_oracle.py.
 */
#ifndef PARHIP_ORACLE_H
#define PARHIP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- GF(2^64), poly x^64 + x^4 + x^3 + x + 1 ---------------- */
uint64_t or_gf64_mul(uint64_t a, uint64_t b);
uint64_t or_gf64_mul_peasant(uint64_t a, uint64_t b);
uint64_t or_gf64_inv(uint64_t a);
uint64_t or_gf64_inv_fermat(uint64_t a);
void or_gf64_clmul(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi);
uint64_t or_gf64_reduce(uint64_t lo, uint64_t hi);

void or_gf64_build_cauchy(uint64_t *out, size_t numInputs, size_t numRecovery,
                          uint64_t firstInput, uint64_t firstRecovery);

void or_gf64_region_mul(uint64_t *out, const uint64_t *in, size_t len, uint64_t c);
void or_gf64_region_mul_arr(uint64_t *out, const uint64_t *in, const uint64_t *coeff,
                            size_t len, size_t n_coeff);
void or_gf64_region_muladd_arr(uint64_t *out, const uint64_t *in, const uint64_t *coeff,
                               size_t len, size_t n_coeff);
void or_gf64_region_coupled_muladd_arr(uint64_t *out, const uint64_t *const *in_blocks,
                                       const uint64_t *coeff_blocks, size_t len, size_t G);
void or_gf64_region_fused_output_muladd_arr(uint64_t *const *outs, const uint64_t *in,
                                            const uint64_t *const *coeff_ptrs, size_t len, size_t K);
void or_gf64_region_2d_muladd_arr(uint64_t *const *outs, size_t K, const uint64_t *const *in_blocks,
                                  size_t G, const uint64_t *coeff2d, size_t K_stride, size_t len);

void or_gf64_compute_recovery_with_coeff(const uint64_t *inputs, size_t numInputs,
                                         uint64_t *recovery, size_t numRecovery,
                                         size_t blockSize64, const uint64_t *coeff,
                                         int accumulate);
void or_gf64_compute_recovery(const uint64_t *inputs, size_t numInputs,
                              uint64_t *recovery, size_t numRecovery, size_t blockSize64,
                              uint64_t firstInput, uint64_t firstRecovery, int accumulate);
/* Same contraction restricted to word columns cols[0..ncols): out is R x ncols. */
void or_gf64_recovery_columns(const uint64_t *inputs, size_t numInputs, size_t blockSize64,
                              const uint64_t *coeff, size_t numRecovery,
                              const uint64_t *cols, size_t ncols, uint64_t *out);
int or_gf64_solve_and_reconstruct(uint64_t *A, uint64_t *rhs, size_t n, size_t blockSize64);
int or_gf64_solve(const uint64_t *A, const uint64_t *b, uint64_t *x, size_t n);

/* ---------------- GF(2^16), poly 0x1100B (PAR2) ---------------- */
uint16_t or_gf16_mul(uint16_t a, uint16_t b);
uint16_t or_gf16_mul_peasant(uint16_t a, uint16_t b);
uint16_t or_gf16_exp(uint32_t e);
uint16_t or_gf16_log(uint16_t v);
uint16_t or_gf16_input_log(uint32_t inputNum);
uint16_t or_gf16_coeff(uint32_t inputNum, uint32_t exponent);
void or_gf16_muladd_region(uint8_t *out, const uint8_t *in, size_t bytes, uint16_t coeff);
void or_par2_create(const uint8_t *const *inputs, const size_t *sizes, const uint16_t *inputNums,
                    size_t numInputs, uint8_t *const *outputs, const uint16_t *exponents,
                    size_t numOutputs, size_t sliceSize);
int or_gf16_rec_matrix_compute(const uint8_t *inputValid, size_t numInputs, size_t validCount,
                               uint16_t *recovery, size_t *recoveryCount, uint16_t *out);
void or_par2_recovery_columns(const uint8_t *inputs, size_t numInputs, size_t sliceSize,
                              const uint16_t *inputNums, const uint16_t *exponents,
                              size_t numOutputs, const uint64_t *cols, size_t ncols,
                              uint16_t *out);

/* CPU baseline of the PAR2 create (baseline_gf16.c): out = r x words;
 * returns the threads used.  Checked against or_par2_create in the tests. */
int bl_par2_create(const uint16_t *in, size_t n, size_t words, const uint16_t *input_nums,
                   const uint16_t *exps, size_t r, uint16_t *out, int threads);

/* ---------------- PAR3 recovery packets + BLAKE3 seal ---------------- */
void or_blake3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[16]);
void or_blake3(const uint8_t *msg, size_t len, uint8_t out[32]);
void or_assemble_recovery_packets(const uint8_t *recovery, const uint8_t inputSetId[8],
                                  uint64_t firstRecoveryIndex, size_t numRecovery,
                                  size_t blockSize, uint8_t *out);

/* ---------------- helpers used by tests ---------------- */
uint32_t or_crc32(const uint8_t *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif
