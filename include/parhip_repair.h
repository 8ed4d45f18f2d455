/*
 * parhip_repair.h -- PAR3 repair on the device, and a repair server.
 *
 * parhip_d_gf64_repair replaces the repair pipeline of the reference:
 * RHS assembly (lib/par3gen.js:3840-3995: rhs[eq] = REC[firstRecovery+eq] XOR
 * sum over known slices of c(eq, j) * D[j]) followed by
 * GF64Controller::SolveAndReconstruct (src/par3_engine.cc:980-1048) on the k x k
 * Cauchy sub-matrix.  The result -- the k lost slices -- is unique, so it is
 * byte-identical to the reference's; the inverse is computed in closed form.
 *
 * The server has no counterpart in the reference (repair there is one CLI
 * call, lib/par3gen.js:3499): it keeps archives resident in HBM, admits
 * requests in arrival order while the requests in flight and their device
 * memory stay within the limits, overlaps admitted requests on separate
 * streams, and reports host-observed latency percentiles.
 */
#ifndef PARHIP_REPAIR_H
#define PARHIP_REPAIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Rebuild the k input slices `missing` (host array of input indices) of an
 * archive: data slice i at d_data + i*dataStride64 (the lost slots' memory may
 * hold anything: it may be read but never contributes),
 * recovery slice r at d_recovery + r*recoveryStride64 with Cauchy points
 * (firstInput + i, firstRecovery + r).  recoveryRows (host, k indices) selects
 * the recovery slices used, NULL = the first k (lib/par3gen.js:3892-3901).
 * Slice m of the result goes to d_out + m*outStride64.  Returns after the
 * work on `stream` has finished. */
int parhip_d_gf64_repair(const uint64_t *d_data, size_t numInputs, size_t dataStride64,
                         const uint64_t *d_recovery, size_t numRecovery, size_t recoveryStride64,
                         size_t blockSize64, uint64_t firstInput, uint64_t firstRecovery,
                         const uint64_t *missing, size_t k, const uint64_t *recoveryRows,
                         uint64_t *d_out, size_t outStride64, void *stream);

/* Device bytes one repair of this shape needs besides its output. */
size_t parhip_gf64_repair_workspace_bytes(size_t numInputs, size_t blockSize64, size_t k);

typedef struct parhip_repair_server parhip_repair_server;

/* hbmBudgetBytes bounds the server's device memory: the workspace of requests
 * in flight, outputs not yet released and cached inverses.  A request waits
 * while it does not fit (after evicting unused cached inverses), whether or not
 * others are in flight: a caller that holds outputs unreleased while it waits
 * for a later request can stall it.  A request that can never fit is refused
 * at submit.  maxInflight bounds concurrent requests; numStreams (0 =
 * maxInflight) streams. */
parhip_repair_server *parhip_repair_server_create(size_t hbmBudgetBytes, unsigned maxInflight,
                                                  unsigned numStreams);
/* Register a device-resident archive; returns its id (>= 0) or an error. */
int parhip_repair_server_add_archive(parhip_repair_server *s, const uint64_t *d_data,
                                     size_t numInputs, size_t dataStride64,
                                     const uint64_t *d_recovery, size_t numRecovery,
                                     size_t recoveryStride64, size_t blockSize64,
                                     uint64_t firstInput, uint64_t firstRecovery);
/* Queue a repair of k lost input slices; returns a request id (>= 0) or an error. */
int64_t parhip_repair_submit(parhip_repair_server *s, int archive, const uint64_t *missing,
                             size_t k);
/* Block until the request completes; latency_ms = submit -> completion. */
int parhip_repair_wait(parhip_repair_server *s, int64_t id, double *latency_ms);
/* Device address of the k repaired slices (k x blockSize64 words), valid until release. */
int parhip_repair_result(parhip_repair_server *s, int64_t id, uint64_t **d_out, size_t *k);
/* Free the request's output and return its memory to the budget. */
int parhip_repair_release(parhip_repair_server *s, int64_t id);
int parhip_repair_server_stats(parhip_repair_server *s, uint64_t *completed, double *p50_ms,
                               double *p99_ms, unsigned *peak_inflight, uint64_t *peak_bytes);
/* The resident inverse cache: inverted k x k matrices kept in HBM, keyed by
 * (archive, missing slices in request order), charged to the HBM budget and
 * evicted least recently used first when a request needs the memory.  hits /
 * misses count admitted requests; entries / bytes are the current contents. */
int parhip_repair_server_cache_stats(parhip_repair_server *s, uint64_t *hits, uint64_t *misses,
                                     uint64_t *entries, uint64_t *bytes, uint64_t *evictions);
/* Drains queued work, then frees everything. */
void parhip_repair_server_destroy(parhip_repair_server *s);

#ifdef __cplusplus
}
#endif

#endif /* PARHIP_REPAIR_H */
