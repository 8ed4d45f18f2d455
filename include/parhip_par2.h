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

/* ---- PAR2 backend engine (IPAR2ProcBackend contract) --------------------- */
/* The engine behind a PAR2 backend of the reference's controller
 * (IPAR2ProcBackend, gf16/controller.h:77-217).  The backend class a
 * maintainer adds to the reference tree (parparpar_amd/reference_backend/
 * controller_hip.{h,cpp}: class PAR2ProcHIP : public IPAR2ProcBackend, both
 * the futures and the USE_LIBUV flavour) forwards to these calls.
 *
 * Staging model (gf16/controller_ocl.cpp:382-456): inputs fill the slots of
 * the current staging area; an area goes to the device as one batch when it
 * holds inputBatchSize inputs, on flush, or early once it holds the minimum
 * batch size while nothing is on the device; the first batch of a pass
 * overwrites the recovery slices, later batches XOR into them; discard_output
 * starts a new pass.  Copies and checksums of the added slices run on the
 * engine's copy workers; a batch's host-to-device copy overlaps the previous
 * batch's contraction. */
typedef struct parhip_par2 parhip_par2;
/* Completion callbacks (they run on engine threads and must not block long):
 * done: the buffer given to an add may be reused (a copy worker);
 * valid: a recovery slice was copied out, cksumValid (a copy worker);
 * batch: a batch of numInputs inputs in staging area `area` was started
 * (on the calling thread, inside the call that closed the area) or finished on
 * the device, its area free again (the engine's device thread). */
typedef void (*parhip_par2_done_fn)(void *user);
typedef void (*parhip_par2_valid_fn)(void *user, int cksumValid);
typedef void (*parhip_par2_batch_fn)(void *user, unsigned area, unsigned numInputs);

/* setSliceSize + init: sliceSize bytes per input slice (0: set later with
 * parhip_par2_set_slice_size), inputBatchSize inputs per batch (0: chosen by
 * the engine from the slice size, about 64 MiB of slices, 16 to 512 inputs, as
 * the reference's backends choose theirs, gf16/controller_ocl_init.cpp:1352),
 * two staging areas. NULL on failure (parhip_last_error). */
parhip_par2 *parhip_par2_create(int device, size_t sliceSize, unsigned inputBatchSize);
/* The same with stagingAreas staging areas (0 = 2, at most 64), the count the
 * reference's backends take in their constructors (gf16/controller_ocl.h:171). */
parhip_par2 *parhip_par2_create_areas(int device, size_t sliceSize, unsigned inputBatchSize,
                                      unsigned stagingAreas);
void parhip_par2_destroy(parhip_par2 *p);
/* setSliceSize: (re)allocates staging and outputs for sliceSize bytes. */
int parhip_par2_set_slice_size(parhip_par2 *p, size_t sliceSize);
/* setRecoverySlices: exponents NULL = coefficients supplied per input. */
int parhip_par2_set_recovery_slices(parhip_par2 *p, unsigned numSlices, const uint16_t *exponents);
/* setCurrentSliceSize (between passes). */
int parhip_par2_set_current_slice_size(parhip_par2 *p, size_t size);
/* canAdd: 0 = PROC_ADD_OK, 1 = PROC_ADD_OK_BUSY, 2 = PROC_ADD_FULL (the
 * current staging area is still on the device). */
int parhip_par2_can_add(parhip_par2 *p);
/* waitForAdd: block until the current staging area is free. */
int parhip_par2_wait_for_add(parhip_par2 *p);
/* addInput(buffer, size <= slice, inputNum, flush); returns once `buffer` may
 * be reused (its copy and checksum are done). */
int parhip_par2_add_input(parhip_par2 *p, const void *buffer, size_t size, uint16_t inputNum,
                          int flush);
/* addInput with explicit per-output coefficients coeffs[numSlices] (read
 * before the call returns). */
int parhip_par2_add_input_coeffs(parhip_par2 *p, const void *buffer, size_t size,
                                 const uint16_t *coeffs, int flush);
/* Asynchronous addInput: returns once the slot is claimed (blocking only
 * while the current staging area is still on the device); done(user) fires
 * when `buffer` may be reused (done may be NULL if the caller keeps the buffer
 * unchanged until end_input returns). */
int parhip_par2_add_input_async(parhip_par2 *p, const void *buffer, size_t size, uint16_t inputNum,
                                int flush, parhip_par2_done_fn done, void *user);
int parhip_par2_add_input_coeffs_async(parhip_par2 *p, const void *buffer, size_t size,
                                       const uint16_t *coeffs, int flush, parhip_par2_done_fn done,
                                       void *user);
/* Batch start / finish notifications (either may be NULL). */
int parhip_par2_set_batch_callbacks(parhip_par2 *p, parhip_par2_batch_fn started,
                                    parhip_par2_batch_fn finished, void *user);
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
/* getOutput: copies recovery slice `index` (current slice size bytes) after
 * processing has finished; returns 1 when its checksum verified (cksumValid),
 * 0 when not, < 0 on error. */
int parhip_par2_get_output(parhip_par2 *p, unsigned index, void *output);
/* Asynchronous getOutput: fn(user, cksumValid) from a copy worker. */
int parhip_par2_get_output_async(parhip_par2 *p, unsigned index, void *output,
                                 parhip_par2_valid_fn fn, void *user);
/* discardOutput: the next batch overwrites instead of accumulating. */
int parhip_par2_discard_output(parhip_par2 *p);
int parhip_par2_set_min_input_batch_size(parhip_par2 *p, unsigned size);
int parhip_par2_input_batch_size(parhip_par2 *p);
int parhip_par2_batches_started(parhip_par2 *p);
int parhip_par2_staging_areas(parhip_par2 *p);
/* freeProcessingMem: recovery slices and read-back buffers. */
int parhip_par2_free_processing_mem(parhip_par2 *p);
/* Device address and row stride (bytes) of recovery slice `index` in HBM: the
 * slice's 16-bit words followed by its checksum word.  Valid until the slice
 * size or recovery count changes. */
int parhip_par2_output_device_ptr(parhip_par2 *p, unsigned index, void **d_row,
                                  size_t *strideBytes);

#ifdef __cplusplus
}
#endif

#endif /* PARHIP_PAR2_H */
