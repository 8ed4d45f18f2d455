/*
 * parhip_par2.h -- C ABI of the PAR2 GF(2^16) recovery path on MI355X.
 *
 * Two levels:
 *   - device-resident contraction (the kernel-level replacement of
 *     Galois16Mul::mul_add_multi_packed, gf16/gf16mul.h:360-390, and of the
 *     OpenCL LOOKUP kernels, gf16/controller_ocl_init.cpp:115-200);
 *   - a backend object with the IPAR2ProcBackend contract
 *     (gf16/controller.h:77-217): the HIP sibling of PAR2ProcCPU / PAR2ProcOCL
 *     (gf16/controller_cpu.h, gf16/controller_ocl.h) that src/gf.cc's GfProc
 *     (:256-395, :700-731) would instantiate for a GPU device.
 * Words are 16-bit little-endian; odd byte counts are zero-padded to a word.
 */
#ifndef PARHIP_PAR2_H
#define PARHIP_PAR2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* gfmat_coeff (gf16/gfmat_coeff.c:99-101): coefficient of input `inputNum`
 * (0..32767) for recovery exponent `exponent`. Host computation. */
uint16_t parhip_gf16_coeff(uint16_t inputNum, uint16_t exponent);
/* GF(2^16) product modulo 0x1100B (gf16/gf16_global.h:9). Host computation. */
uint16_t parhip_gf16_mul(uint16_t a, uint16_t b);

/* d_coeff[e*numInputs + i] = gfmat_coeff(inputNums[i], exponents[e]) computed
 * on the device (host index arrays, device output). */
int parhip_d_gf16_coeffs(uint16_t *d_coeff, const uint16_t *inputNums, size_t numInputs,
                         const uint16_t *exponents, size_t numOutputs, void *stream);

/* out_e[w] (^)= XOR_i d_coeff[e*numInputs + i] * in_i[w] for w < words, on
 * device-resident slices (input i at d_inputs + i*inputStrideWords, output e
 * at d_outputs + e*outputStrideWords).  accumulate=0 overwrites. */
int parhip_d_gf16_muladd(const uint16_t *d_inputs, size_t numInputs, size_t inputStrideWords,
                         uint16_t *d_outputs, size_t numOutputs, size_t outputStrideWords,
                         size_t words, const uint16_t *d_coeff, int accumulate, void *stream);

/* ---- PAR2 repair matrix ------------------------------------------------- */
/* Galois16RecMatrix::Compute (gf16/gfmat_inv.cpp:562-676, Construct :462-560):
 * inputValid[i] != 0 for the inputs present (validCount of numInputs <= 32768);
 * recovery[0..*recoveryCount) the available exponents.  The exponents are
 * sorted; a recovery row whose pivot turns out zero (PAR2's non-invertible
 * combinations) is dropped and the inversion retried.  Returns 1 when
 * numRec = numInputs - validCount rows were inverted: recovery[0..numRec) are
 * the exponents used, *recoveryCount = numRec, and factors (numRec x
 * numInputs, row-major) hold GetFactor(col, row): columns 0..validCount-1 the
 * valid inputs in input order, then the missing inputs in input order.
 * Returns 0 (Compute's false) when too few exponents remain (*recoveryCount =
 * those left), a negative PARHIP_ERR_* on invalid arguments or device errors.
 * The inversion runs on the device. */
int parhip_gf16_rec_matrix_compute(const uint8_t *inputValid, uint32_t numInputs,
                                   uint32_t validCount, uint16_t *recovery,
                                   uint32_t *recoveryCount, uint16_t *factors);
/* Same, factors left in device memory (numRec x numInputs uint16) for a
 * device repair contraction; synchronises `stream` once per attempt. */
int parhip_d_gf16_rec_matrix_compute(const uint8_t *inputValid, uint32_t numInputs,
                                     uint32_t validCount, uint16_t *recovery,
                                     uint32_t *recoveryCount, uint16_t *d_factors, void *stream);

/* ---- PAR2ProcHIP backend (IPAR2ProcBackend contract) -------------------- */
typedef struct parhip_par2 parhip_par2;

/* setSliceSize + init: sliceSize bytes per input slice, inputBatchSize inputs
 * per kernel launch (0 = default). NULL on failure (parhip_last_error). */
parhip_par2 *parhip_par2_create(int device, size_t sliceSize, unsigned inputBatchSize);
void parhip_par2_destroy(parhip_par2 *p);
/* setRecoverySlices: exponents NULL = coefficients supplied per input. */
int parhip_par2_set_recovery_slices(parhip_par2 *p, unsigned numSlices, const uint16_t *exponents);
/* setCurrentSliceSize (between passes). */
int parhip_par2_set_current_slice_size(parhip_par2 *p, size_t size);
/* canAdd: 0 = PROC_ADD_OK, 1 = PROC_ADD_OK_BUSY, 2 = PROC_ADD_FULL. */
int parhip_par2_can_add(parhip_par2 *p);
/* addInput(buffer, size <= slice, inputNum, flush); returns once `buffer` may be reused. */
int parhip_par2_add_input(parhip_par2 *p, const void *buffer, size_t size, uint16_t inputNum,
                          int flush);
/* addInput with explicit per-output coefficients coeffs[numSlices]. */
int parhip_par2_add_input_coeffs(parhip_par2 *p, const void *buffer, size_t size,
                                 const uint16_t *coeffs, int flush);
/* Benchmarking, IPAR2ProcBackend::fillInput / dummyInput (gf16/controller.h:172,
 * :303-305): fill_input copies `buffer` (current slice size bytes) into the next
 * staging slot and returns 1 once every slot of every staging area holds it
 * (else 0); dummy_input then adds input `inputNum` over whatever its staging
 * slot holds, without new data. */
int parhip_par2_fill_input(parhip_par2 *p, const void *buffer);
int parhip_par2_dummy_input(parhip_par2 *p, uint16_t inputNum, int flush);
int parhip_par2_flush(parhip_par2 *p);
/* endInput: flush and wait until every added input has been processed. */
int parhip_par2_end_input(parhip_par2 *p);
/* getOutput: copies recovery slice `index` (current slice size bytes); returns
 * 1 when its checksum verified (cksumValid), 0 when not, < 0 on error. */
int parhip_par2_get_output(parhip_par2 *p, unsigned index, void *output);
/* discardOutput: the next batch overwrites instead of accumulating. */
int parhip_par2_discard_output(parhip_par2 *p);
int parhip_par2_input_batch_size(parhip_par2 *p);

#ifdef __cplusplus
}
#endif

#endif /* PARHIP_PAR2_H */
