/*
 * parhip.h -- C ABI of the MI355X (gfx950) recovery-slice engine.
 *
 * Drop-in boundary for ParParPar's GF(2^64) PAR3 engine.  Each entry point
 * replaces one member of `class GF64Controller` (src/par3_engine.h:11-244) or
 * one region operation of gf64/gf64_global.h:23-29, with the same argument
 * meaning, word layout and value semantics, so the reference's N-API addon
 * (src/gf64_addon.cc) can bind it by swapping the call at each site.
 *
 * Differences from the reference, all additive:
 *   - every entry returns an int status (PARHIP_OK = 0) where the reference
 *     returns void and silently skips (src/par3_engine.cc:1069, :1173-1176);
 *     the no-op cases (numInputs == 0, numRecovery == 0) still return
 *     PARHIP_OK and touch nothing, as the reference does;
 *   - `numThreads` is accepted and ignored (work is spread over the GPU);
 *   - calls are re-entrant: no unsynchronised global state (the reference's
 *     coefficient LRU and dispatch pointers are not, src/par3_engine.cc:107-112);
 *   - `parhip_d_*` variants take device pointers and a hipStream_t (as void*),
 *     for data already resident in HBM; they enqueue work and return without
 *     synchronising.
 *
 * Words are native little-endian uint64_t; a slice of blockSize64 words is
 * blockSize64*8 bytes; inputs are contiguous numInputs x blockSize64 words and
 * recovery blocks contiguous numRecovery x blockSize64 words (par3_engine.h:14-23).
 */
#ifndef PARHIP_H
#define PARHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PARHIP_OK 0
#define PARHIP_ERR_SINGULAR (-1) /* singular matrix: the reference's -1            */
#define PARHIP_ERR_ARG (-2)      /* invalid argument (null pointer, bad size)      */
#define PARHIP_ERR_NOMEM (-3)    /* device or host allocation failed               */
#define PARHIP_ERR_DEVICE (-4)   /* HIP runtime error                              */
#define PARHIP_ERR_RANGE (-5)    /* size beyond a documented limit                 */

/* ---- multiply strategies for the contraction ---------------------------- */
#define PARHIP_METHOD_AUTO 0     /* per-shape choice (autotuner table)             */
#define PARHIP_METHOD_TABLE 1    /* split nibble tables in LDS                     */
#define PARHIP_METHOD_MFMA 2     /* GF(2) bit-matrix on fp4 MFMA (exact 0/1 counts) */
#define PARHIP_METHOD_MFMA_I8 3  /* GF(2) bit-matrix on int8 MFMA (int32 parity)    */
#define PARHIP_METHOD_TABLE8 4   /* 8-bit split tables in LDS (256-entry rows)      */
#define PARHIP_METHOD_DYADIC 5   /* plain Cauchy calls: 8 x 8, 16 x 16 or 32 x 32   */
                                 /* XOR blocks as subset convolutions (27 / 81 /    */
                                 /* 243 products per block), nibble tables in LDS;  */
                                 /* others as AUTO                                  */
#define PARHIP_METHOD_DYADIC8 6  /* as DYADIC with 8 x 8 blocks only                */
#define PARHIP_METHOD_DYADIC16 7 /* as DYADIC with 16 x 16 blocks wherever the call */
                                 /* allows them (plain Cauchy create)               */
#define PARHIP_METHOD_DYADIC32 8 /* as DYADIC with 32 x 32 blocks wherever the call */
                                 /* allows them (plain Cauchy create)               */

/* Library version (major*10000 + minor*100 + patch). */
int parhip_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char *parhip_last_error(void);
/* Select the HIP device used by calls from this thread (default 0). */
int parhip_set_device(int device);
/* Force a contraction strategy for this thread (PARHIP_METHOD_*), for the
 * GF(2^64) contraction and (MFMA, MFMA_I8) the PAR2 GF(2^16) one.  The MFMA and
 * 8-bit table strategies apply to strided slices; pointer-array calls use the
 * nibble tables. */
int parhip_set_method(int method);
/* Force the contraction launch shape for this thread (tuning; no counterpart
 * in the reference). Tables: rt = outputs per work-group, wpt = words per
 * thread, ib = inputs per LDS batch; MFMA: rt / wpt = the kernel's two tile
 * dimensions; groups = split-K rows; 0 = automatic. */
int parhip_tune_override(int rt, int wpt, int ib, int64_t groups);
/* Coefficient cache of this thread's plain Cauchy calls: on (default) keeps
 * built coefficient sets on the device in an LRU of 16, as the reference keeps
 * its Cauchy matrices (src/par3_engine.cc:451-482); off builds them on every
 * call, into stream-ordered memory freed after the call. */
int parhip_set_coeff_cache(int enabled);
/* Release cached device memory (coefficient cache, workspaces, this thread's
 * staging ring). */
int parhip_release_cache(void);

/* ======================================================================= */
/* GF64Controller, host pointers (synchronous: return after results land)   */
/* ======================================================================= */

/* GF64Controller::ComputeRecoveryBlocks (src/par3_engine.h:24-32,
 * src/par3_engine.cc:1061-1111). accumulate=0 overwrites the recovery blocks,
 * accumulate!=0 XORs into them (chunked create, lib/par3gen.js:2306-2316). */
int parhip_gf64_compute_recovery_blocks(const uint64_t *inputs, size_t numInputs,
                                        uint64_t *recovery, size_t numRecovery,
                                        size_t blockSize64, uint64_t firstInput,
                                        uint64_t firstRecovery, int numThreads, int accumulate);

/* GF64Controller::ComputeRecoveryBlocksFull (src/par3_engine.h:49-57, :1389-1399). */
int parhip_gf64_compute_recovery_blocks_full(const uint64_t *inputs, size_t numInputs,
                                             uint64_t *recovery, size_t numRecovery,
                                             size_t blockSize64, uint64_t firstInput,
                                             uint64_t firstRecovery, int numThreads,
                                             int accumulate);

/* GF64Controller::ComputeRecoveryBlocksBarycentric (src/par3_engine.h:209-217,
 * src/par3_engine_barycentric.cc:104-280; N-API compute_recovery_barycentric,
 * src/gf64_addon.cc:1636-1796).  For R <= 32 the reference delegates to
 * ComputeRecoveryBlocks; for R > 32 it inverts (firstInput+c) XOR
 * (firstRecovery+r) (0 -> 1) per row and accumulates, which is the same
 * Cauchy product: the values equal parhip_gf64_compute_recovery_blocks. */
int parhip_gf64_compute_recovery_blocks_barycentric(const uint64_t *inputs, size_t numInputs,
                                                    uint64_t *recovery, size_t numRecovery,
                                                    size_t blockSize64, uint64_t firstInput,
                                                    uint64_t firstRecovery, int numThreads,
                                                    int accumulate);

/* GF64Controller::ComputeRecoveryBlocksFenger (src/par3_engine.h:237-244,
 * src/par3_engine_fenger.cc:109-296;
 * N-API compute_recovery_fenger, src/gf64_addon.cc:1980-2078).  The reference
 * documents it as bit-exact with ComputeRecoveryBlocks for any workload (it is
 * an asymptotically fast evaluation of the same Cauchy product, chosen there
 * for R >= 8192); here it is that contraction.  Recovery is written. */
int parhip_gf64_compute_recovery_blocks_fenger(const uint64_t *inputs, size_t numInputs,
                                               uint64_t *recovery, size_t numRecovery,
                                               size_t blockSize64, uint64_t firstInput,
                                               uint64_t firstRecovery, int numThreads);

/* N-API compute_recovery_streaming (src/gf64_addon.cc:2349-2650): the input
 * slices arrive as numBatches host buffers, batch b holding batchBlocks[b]
 * whole slices; concatenated in order they are the numInputs inputs (global
 * indices firstInput, firstInput+1, ...).  The addon copies them into one
 * buffer and calls the engine once (Fenger, Barycentric or Cauchy: the same
 * values); here they are streamed to the device from where they lie.
 * Recovery is written, not accumulated.  numInputs == 0 or numRecovery == 0
 * returns without touching recovery; batch counts that do not sum to
 * numInputs give PARHIP_ERR_RANGE. */
int parhip_gf64_compute_recovery_batches(const uint64_t *const *batches, const size_t *batchBlocks,
                                         size_t numBatches, uint64_t *recovery, size_t numInputs,
                                         size_t numRecovery, size_t blockSize64,
                                         uint64_t firstInput, uint64_t firstRecovery,
                                         int numThreads);

/* GF64Controller::ComputeRecoveryBlocksWithCoeff (src/par3_engine.h:65-72,
 * src/par3_engine.cc:1165-1382); coeff is numRecovery x numInputs row-major. */
int parhip_gf64_compute_recovery_blocks_with_coeff(const uint64_t *inputs, size_t numInputs,
                                                   uint64_t *recovery, size_t numRecovery,
                                                   size_t blockSize64, const uint64_t *coeff,
                                                   int numThreads, int accumulate);

/* GF64Controller::BuildCauchyMatrix (src/par3_engine.h:119-123, :358-429):
 * M[r][c] = inv((firstInput+c) XOR (firstRecovery+r)), zero denominators -> 1. */
int parhip_gf64_build_cauchy_matrix(uint64_t *coeffMatrix, size_t numInputs, size_t numRecovery,
                                    uint64_t firstInput, uint64_t firstRecovery);

/* GF64Controller::BuildCauchyMatrixAlloc (src/par3_engine.h:84-87, :432-441):
 * malloc'd result (free with free() or parhip_free), NULL on N==0, R==0 or OOM. */
uint64_t *parhip_gf64_build_cauchy_matrix_alloc(size_t numInputs, size_t numRecovery,
                                                uint64_t firstInput, uint64_t firstRecovery);
void parhip_free(void *p);

/* GF64Controller::MultiplyAccumulate (src/par3_engine.h:133-138):
 * out[i] ^= XOR_j in[j] * coeffMatrix[i*numIn + j]. */
int parhip_gf64_multiply_accumulate(uint64_t *out, size_t numOut, const uint64_t *in,
                                    size_t numIn, const uint64_t *coeffMatrix, size_t blockSize64);

/* GF64Controller::SolveAndReconstruct (src/par3_engine.h:162-168, :980-1048):
 * A (n x n, row-major) is eliminated in place exactly as the reference does it
 * (first nonzero pivot, row swaps); rhs (n x blockSize64) becomes A^-1 * rhs.
 * Returns PARHIP_OK (0), or PARHIP_ERR_SINGULAR (-1, the reference's value)
 * with A and rhs left in the reference's partially-eliminated state. */
int parhip_gf64_solve_and_reconstruct(uint64_t *A, uint64_t *rhsBlocks, size_t n,
                                      size_t blockSize64, int numThreads);

/* gf64_solve (gf64/gf64_solve.c:39-96): x = A^-1 b for one word per row. */
int parhip_gf64_solve(const uint64_t *A, const uint64_t *b, uint64_t *x, size_t n);

/* gf64_mul / gf64_inverse element operations (gf64/gf64_single.c:6-53,
 * gf64/gf64_invert.c:29-67). Host computation. */
uint64_t parhip_gf64_mul(uint64_t a, uint64_t b);
uint64_t parhip_gf64_inverse(uint64_t a);

/* Host inputs larger than one staging chunk are streamed to the device through
 * a ring of pinned slots (copy threads + DMA + contraction overlapped, the
 * chunks XOR-accumulated); results are identical to one contraction.
 * Chunk size for this thread in bytes (0 = default 64 MiB, minimum 4096). */
int parhip_set_stream_chunk(size_t bytes);

/* par3_create_streaming (src/gf64_create_streaming.cc:93-240, N-API
 * validation :327-470): read sourcePath, zero-pad the tail to whole blocks,
 * numInputs = ceil(size / blockSize), compute every recovery block with the
 * Cauchy points (firstInput, firstRecovery).  The file is read in chunks into
 * pinned memory and reduced on the device while the next chunk is read.
 * *recoveryOut: numRecovery * blockSize bytes from aligned_alloc(64) (free
 * with free() or parhip_free); durationMs covers read + compute;
 * throughputMBps = recovery MiB / s as the reference reports it.  useMmap and
 * numThreads are accepted for signature parity. Errors: an open failure is
 * reported with its Node-style code (ENOENT, EACCES, ...) in
 * parhip_last_error(); an empty file and a blockSize that is not a positive
 * multiple of 8 are rejected. */
int parhip_par3_create_streaming(const char *sourcePath, int32_t numRecovery, int64_t blockSize,
                                 uint64_t firstInput, uint64_t firstRecovery, int32_t numThreads,
                                 int useMmap, uint8_t **recoveryOut, size_t *recoveryBytes,
                                 double *durationMs, double *throughputMBps);

/* GF64Controller::ComputeRecoveryBlocksFromFile (src/par3_engine.h:104-110,
 * src/par3_engine.cc:1445-1537): numInputs = file size / (blockSize64 * 8),
 * the size a positive multiple of it; the numRecovery recovery blocks go into
 * the caller's buffer.  As in the reference, the call is gated by the
 * environment variable PAR3_GF64_USE_MMAP=1 and returns 0 on success, -1 on
 * any failure (gate off, null argument, open/fstat failure, size mismatch);
 * parhip_last_error() says which.  The file is read in chunks into pinned
 * memory and reduced on the device while the next chunk is read.  numThreads
 * is accepted for signature parity. */
int parhip_gf64_compute_recovery_blocks_from_file(const char *sourcePath, uint64_t *recovery,
                                                  size_t numRecovery, size_t blockSize64,
                                                  uint64_t firstInput, uint64_t firstRecovery,
                                                  int numThreads);

/* ======================================================================= */
/* PAR3 recovery packets                                                    */
/* ======================================================================= */

/* assemble_recovery_packets (AssembleRecoveryPackets_NAPI,
 * src/gf64_addon.cc:2790-2938): numRecovery packets of 64 + blockSize bytes
 * into out (caller-allocated, numRecovery * (64 + blockSize) bytes): magic
 * "PAR3\0PKT", 16-byte BLAKE3 of packet[24..end), LE64 packet length, the
 * 8-byte inputSetId, "PAR REC\0", LE64 block index firstRecoveryIndex + i,
 * LE64 count 1, block i of `recovery` (numRecovery * blockSize bytes).
 * Assembled and hashed on the device.  numRecovery and blockSize must be
 * positive and their product at most 1 TiB (PARHIP_ERR_RANGE otherwise). */
int parhip_assemble_recovery_packets(const uint8_t *recovery, const uint8_t *inputSetId,
                                     int64_t firstRecoveryIndex, int64_t numRecovery,
                                     int64_t blockSize, uint8_t *out);
/* Device-pointer form (inputSetId stays a host pointer to 8 bytes); enqueues
 * on `stream` without synchronising. */
int parhip_d_assemble_recovery_packets(const void *d_recovery, const uint8_t *inputSetId,
                                       int64_t firstRecoveryIndex, int64_t numRecovery,
                                       int64_t blockSize, void *d_out, void *stream);

/* ======================================================================= */
/* Region operations (gf64/gf64_global.h:23-29), host pointers              */
/* ======================================================================= */
/* out[w] = in[w] * c */
int parhip_gf64_region_mul(uint64_t *out, const uint64_t *in, size_t len, uint64_t constant);
/* out[w] = in[w] * coeff[w % n_coeff]   (cycling, gf64/gf64_region_scalar.c:14-18) */
int parhip_gf64_region_mul_arr(uint64_t *out, const uint64_t *in, const uint64_t *coeff,
                               size_t len, size_t n_coeff);
/* out[w] ^= XOR_c in[w] * coeff[c]     (sum, gf64/gf64_region_scalar.c:20-28) */
int parhip_gf64_region_muladd_arr(uint64_t *out, const uint64_t *in, const uint64_t *coeff,
                                  size_t len, size_t n_coeff);
/* out[w] ^= XOR_g in_blocks[g][w] * coeff_blocks[g]   (gf64_region_scalar.c:45-55) */
int parhip_gf64_region_coupled_muladd_arr(uint64_t *out, const uint64_t *const *in_blocks,
                                          const uint64_t *coeff_blocks, size_t len, size_t G);
/* outs[k][w] ^= in[w] * (*coeff_block_starts[k])      (gf64_region_scalar.c:58-72) */
int parhip_gf64_region_fused_output_muladd_arr(uint64_t *const *outs, const uint64_t *in,
                                               const uint64_t *const *coeff_block_starts,
                                               size_t len, size_t K);
/* outs[k][w] ^= XOR_g in_blocks[g][w] * coeff2d[k*K_stride + g]  (gf64_region_scalar.c:75-91) */
int parhip_gf64_region_2d_muladd_arr(uint64_t *const *outs, size_t K,
                                     const uint64_t *const *in_blocks, size_t G,
                                     const uint64_t *coeff_block_2d, size_t K_stride, size_t len);

/* ======================================================================= */
/* Device-resident variants (pointers into HBM, stream = hipStream_t or NULL)*/
/* ======================================================================= */

/* Cauchy matrix into device memory d_coeff (numRecovery x numInputs). */
int parhip_d_gf64_build_cauchy(uint64_t *d_coeff, size_t numInputs, size_t numRecovery,
                               uint64_t firstInput, uint64_t firstRecovery, void *stream);

/* Contraction with an explicit device coefficient matrix (ld = numInputs).
 * Input i starts at d_inputs + i*inputStride64, output r at
 * d_recovery + r*recoveryStride64 (strides in words, >= blockSize64). */
int parhip_d_gf64_recovery_with_coeff(const uint64_t *d_inputs, size_t numInputs,
                                      size_t inputStride64, uint64_t *d_recovery,
                                      size_t numRecovery, size_t recoveryStride64,
                                      size_t blockSize64, const uint64_t *d_coeff,
                                      int accumulate, void *stream);

/* Contraction with Cauchy coefficients for points (firstInput+i, firstRecovery+r)
 * (GF64Controller::ComputeRecoveryBlocks, src/par3_engine.cc:1061-1111, on
 * device buffers); the coefficients are built on the device and cached per
 * shape (parhip_set_coeff_cache). */
int parhip_d_gf64_recovery(const uint64_t *d_inputs, size_t numInputs, size_t inputStride64,
                           uint64_t *d_recovery, size_t numRecovery, size_t recoveryStride64,
                           size_t blockSize64, uint64_t firstInput, uint64_t firstRecovery,
                           int accumulate, void *stream);

/* parhip_d_gf64_recovery without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate -- work recorded under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and bytes; the coefficients and split-K partials live in one
 * fixed 64 MiB device array per GPU, held until the process ends (the partials
 * take what the coefficients leave: fewer split-K rows, never an error).
 * Shapes whose coefficients need more than the array -- (ceil((firstInput%8 +
 * numInputs)/8) x ceil((firstRecovery%8 + numRecovery)/8)) blocks of 64 B above
 * 64 MiB -- return PARHIP_ERR_RANGE.  The array keeps the last call's
 * contents; calls that overlap on one GPU share it and race, so order them on
 * one stream or with events. */
int parhip_d_gf64_recovery_static(const uint64_t *d_inputs, size_t numInputs, size_t inputStride64,
                                  uint64_t *d_recovery, size_t numRecovery, size_t recoveryStride64,
                                  size_t blockSize64, uint64_t firstInput, uint64_t firstRecovery,
                                  int accumulate, void *stream);

/* Contraction over scattered slices: d_in_ptrs / d_out_ptrs are device arrays
 * of device pointers (e.g. the known blocks of an archive during repair);
 * d_coeff is numOut x numIn with leading dimension coeff_ld. */
int parhip_d_gf64_recovery_ptrs(const uint64_t *const *d_in_ptrs, size_t numIn,
                                uint64_t *const *d_out_ptrs, size_t numOut, size_t blockSize64,
                                const uint64_t *d_coeff, size_t coeff_ld, int accumulate,
                                void *stream);

/* ======================================================================= */
/* Kernel timing (no counterpart in the reference, which only has wall-     */
/* clock hooks, src/gf64_addon.cc:1583-1615). Per calling thread.           */
/* ======================================================================= */
/* Bracket every kernel launched by this thread with HIP events (on its stream). */
int parhip_prof_enable(int on);
/* Bracket only the kernels of this name ("mac_dyadic", ...; NULL or "" = all):
 * two events per step instead of two per kernel. */
int parhip_prof_only(const char *kernel);
/* Drop accumulated timings. */
int parhip_prof_reset(void);
/* Total device time and launch count of a kernel: "mac_tab", "xor_reduce",
 * "cauchy". Synchronises with the recorded events. */
int parhip_prof_get(const char *kernel, double *total_ms, uint64_t *launches);

/* The device's LDS gather rate for the table kernels' access pattern, measured
 * by a kernel that does only that (random 16-byte, or 8-byte, reads of one entry
 * of a 16-entry 256-byte table row per lane, many waves per CU): nanoseconds per
 * wave-level read per CU. No counterpart in the reference; for measurement
 * (the bound of the XOR-block and table kernels). Runs on the null stream. */
int parhip_measure_lds_gather(double *ns_b128, double *ns_b64);

/* Device workspace the contraction of this shape needs (bytes). Informational. */
size_t parhip_gf64_workspace_bytes(size_t numInputs, size_t numRecovery, size_t blockSize64);

#ifdef __cplusplus
}
#endif

#endif /* PARHIP_H */
