/*
 * parhip_batch.h -- batched PAR3 create and the launch-shape autotuner.
 *
 * No single counterpart in the reference, which runs one create per call
 * (lib/par3gen.js:2441 -> ComputeRecoveryBlocksWithCoeff, src/par3_engine.cc:1165)
 * and only times block sizes for telemetry (PAR3_GF64_BLOCK_SIZE_AUTOTUNE,
 * src/par3_engine.cc:655-732).  Each job below has exactly the value
 * semantics of parhip_d_gf64_recovery (include/parhip.h) with accumulate = 0:
 * recovery[r] = XOR_i cauchy(firstInput+i, firstRecovery+r) * input[i].
 */
#ifndef PARHIP_BATCH_H
#define PARHIP_BATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	const uint64_t *d_inputs; /* device: numInputs slices, inputStride64 words apart */
	size_t numInputs;
	size_t inputStride64;
	uint64_t *d_recovery;     /* device: numRecovery slices, recoveryStride64 words apart */
	size_t numRecovery;
	size_t recoveryStride64;
	size_t blockSize64;       /* words per slice */
	uint64_t firstInput;
	uint64_t firstRecovery;
} parhip_gf64_job;

typedef struct parhip_batch parhip_batch;

#define PARHIP_BATCH_GRAPH 1u    /* capture the launch sequence in a hipGraph */
/* Time the candidates per job shape (cached): the table launch variants in the
 * persistent kernel, and the strategy itself where the GF(2) bit-matrix on MFMA
 * qualifies (its rows, 512 B per coefficient, no larger than the job's input:
 * 64 * numRecovery <= blockSize64): the job alone as full-device launches of
 * the table kernel and of the MFMA kernel.  A job that goes to MFMA runs as its
 * own launches after the persistent kernel. */
#define PARHIP_BATCH_AUTOTUNE 2u
/* Run every job that qualifies as above on the MFMA strategy, untimed. */
#define PARHIP_BATCH_MFMA 4u

/* Plan a batch: per-job launch variant (autotuned when asked), the device tile
 * queue, coefficient storage; optionally capture the graph. Buffers named by
 * the jobs must stay valid while the batch is used. NULL on failure. */
parhip_batch *parhip_batch_create(const parhip_gf64_job *jobs, size_t njobs, unsigned flags);
/* Enqueue one run of every job on `stream` (graph replay when captured). */
int parhip_batch_run(parhip_batch *b, void *stream);
int parhip_batch_info(const parhip_batch *b, uint64_t *ntiles, uint64_t *natomic_jobs,
                      uint64_t *coeff_bytes);
/* Launch variant chosen for job `job` (its index in the array given to
 * parhip_batch_create): table tiles: outputs per work-group, words per thread,
 * inputs per LDS batch; XOR-block tiles: points per block (8, 16 or 32), words
 * per thread, 1; MFMA: outputs per work-group, 32-word column blocks per wave,
 * inputs per stage.  PARHIP_ERR_ARG for an empty job (no inputs, outputs or
 * words). */
int parhip_batch_job_config(const parhip_batch *b, size_t job, int *rt, int *wpt, int *ib);
/* Strategy of job `job`: PARHIP_METHOD_DYADIC (XOR-block tiles), PARHIP_METHOD_TABLE
 * or PARHIP_METHOD_MFMA (include/parhip.h); PARHIP_ERR_ARG for an empty job. */
int parhip_batch_job_method(const parhip_batch *b, size_t job);
void parhip_batch_destroy(parhip_batch *b);

/* Autotuner cache: look up a tuned shape (launch variant, or its strategy),
 * persist, reload, clear. */
int parhip_autotune_lookup(size_t numInputs, size_t numRecovery, size_t blockSize64, int *rt,
                           int *wpt, int *ib, double *ns_per_mac);
int parhip_autotune_lookup_method(size_t numInputs, size_t numRecovery, size_t blockSize64);
int parhip_autotune_save(const char *path);
int parhip_autotune_load(const char *path);
int parhip_autotune_clear(void);

#ifdef __cplusplus
}
#endif

#endif /* PARHIP_BATCH_H */
