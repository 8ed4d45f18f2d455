// engine_host.cpp -- host-pointer entries of the C ABI (include/parhip.h).
//
// These mirror GF64Controller's static methods and the region operations
// exactly as the reference's N-API addon calls them: caller-owned host
// buffers in, results written back before return.  Each call stages its
// operands into HBM on the calling thread's stream, runs the device
// contraction (engine.cpp / gf64_kernels.hip) and copies the results back.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "gf64_host.h"
#include "ingest.h"
#include "kernels.h"

using namespace parhip;

#define RC(expr)                  \
	do {                          \
		int rc_ = (expr);         \
		if (rc_) return rc_;      \
	} while (0)
#define HIPC(expr, what)                                   \
	do {                                                   \
		hipError_t e_ = (expr);                            \
		if (e_ != hipSuccess) return hip_fail(e_, what);   \
	} while (0)

static bool mul_overflows(size_t a, size_t b, size_t* out) {
	if (a && b > SIZE_MAX / a) return true;
	*out = a * b;
	return false;
}

static int words_bytes(size_t count, size_t words, size_t* bytes) {
	size_t w;
	if (mul_overflows(count, words, &w) || mul_overflows(w, sizeof(uint64_t), bytes))
		return fail(PARHIP_ERR_RANGE, "size overflow (%zu x %zu words)", count, words);
	return PARHIP_OK;
}

static int h2d(void* d, const void* h, size_t bytes, hipStream_t s) {
	if (bytes == 0) return PARHIP_OK;
	HIPC(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
	return PARHIP_OK;
}
static int d2h(void* h, const void* d, size_t bytes, hipStream_t s) {
	if (bytes == 0) return PARHIP_OK;
	HIPC(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
	return PARHIP_OK;
}
static int sync(hipStream_t s) {
	HIPC(hipStreamSynchronize(s), "hipStreamSynchronize");
	return PARHIP_OK;
}

// Generic host-pointer contraction: G input blocks (pointers), K output blocks
// (pointers), coefficients K x G (ld = coeff_ld), len words each.
// Run `body` (which enqueues on s); on failure, drain s before returning so
// no copy still in flight reads host memory the caller's frame is releasing.
static int drained(hipStream_t s, const std::function<int()>& body) {
	const int rc = body();
	if (rc) {
		(void)hipStreamSynchronize(s);
		return rc;
	}
	return sync(s);
}

static int host_contract(uint64_t* const* outs, size_t K, const uint64_t* const* ins, size_t G,
                         const uint64_t* coeff, size_t coeff_ld, size_t len, int accumulate) {
	if (K == 0 || G == 0 || len == 0) return PARHIP_OK;
	size_t in_bytes, out_bytes, blk = len * sizeof(uint64_t), cb;
	RC(words_bytes(G, len, &in_bytes));
	RC(words_bytes(K, len, &out_bytes));
	RC(words_bytes(K, G, &cb));
	hipStream_t s;
	RC(host_stream(&s));
	std::vector<uint64_t> c(K * G);
	DevBuf din, dout, dco;
	return drained(s, [&]() -> int {
		RC(din.alloc(in_bytes, s));
		RC(dout.alloc(out_bytes, s));
		RC(dco.alloc(cb, s));
		for (size_t g = 0; g < G; g++) RC(h2d(din.as<uint64_t>() + g * len, ins[g], blk, s));
		if (accumulate)
			for (size_t k = 0; k < K; k++) RC(h2d(dout.as<uint64_t>() + k * len, outs[k], blk, s));
		for (size_t k = 0; k < K; k++)
			for (size_t g = 0; g < G; g++) c[k * G + g] = coeff[k * coeff_ld + g];
		RC(h2d(dco.p, c.data(), cb, s));
		MacArgs a{};
		a.in_base = din.as<uint64_t>();
		a.in_stride = len;
		a.out_base = dout.as<uint64_t>();
		a.out_stride = len;
		a.coeff = dco.as<uint64_t>();
		a.coeff_ld = G;
		a.N = G;
		a.R = K;
		a.W = len;
		a.accumulate = accumulate;
		RC(run_contraction(a, s));
		for (size_t k = 0; k < K; k++) RC(d2h(outs[k], dout.as<uint64_t>() + k * len, blk, s));
		return PARHIP_OK;
	});
}

// Contraction of N contiguous host input slices (W words each) against a
// device coefficient matrix (R x N, leading dimension ld) into device outputs,
// streamed through the pinned staging ring chunk by chunk.  The first chunk
// writes (or XORs when `accumulate`), later chunks XOR: linearity makes it the
// single contraction.
static int stream_contract(const FillFn& fill, uint64_t N, uint64_t W, uint64_t* dout, uint64_t R,
                           const uint64_t* dco, uint64_t ld, int accumulate, hipStream_t s) {
	const ConsumeFn consume = [&](const void* src, uint64_t first, uint64_t count,
	                              hipStream_t st) {
		MacArgs a{};
		a.in_base = (const uint64_t*)src;
		a.in_stride = W;
		a.out_base = dout;
		a.out_stride = W;
		a.coeff = dco + first;
		a.coeff_ld = ld;
		a.N = count;
		a.R = R;
		a.W = W;
		a.accumulate = (accumulate || first > 0) ? 1 : 0;
		return run_contraction(a, st);
	};
	return stream_slices(N, (size_t)W * sizeof(uint64_t), fill, consume, s);
}

// The plain Cauchy contraction (points fi + i, fr + r) of N host slices,
// streamed like stream_contract: chunk [first, first + count) runs as that part
// of the whole call, on one cached coefficient set.
static int stream_cauchy(const FillFn& fill, uint64_t N, uint64_t W, uint64_t* dout, uint64_t R,
                         uint64_t fi, uint64_t fr, int accumulate, hipStream_t s) {
	const ConsumeFn consume = [&](const void* src, uint64_t first, uint64_t count,
	                              hipStream_t st) {
		MacArgs a{};
		a.in_base = (const uint64_t*)src;
		a.in_stride = W;
		a.out_base = dout;
		a.out_stride = W;
		a.N = count;
		a.R = R;
		a.W = W;
		a.accumulate = (accumulate || first > 0) ? 1 : 0;
		return run_cauchy_contraction(a, fi, fr, st, first, N);
	};
	return stream_slices(N, (size_t)W * sizeof(uint64_t), fill, consume, s);
}

// memcpy split into 4 MiB parts over the copy threads.
static void parallel_copy(void* dst, const void* src, size_t bytes) {
	const size_t part = 4u << 20;
	parallel_parts((bytes + part - 1) / part, [&](size_t p) {
		const size_t o = p * part;
		memcpy((char*)dst + o, (const char*)src + o, std::min(part, bytes - o));
	});
}

// Parallel copy of `count` slices from a contiguous host buffer into a pinned slot.
static FillFn memcpy_fill(const uint64_t* inputs, size_t slice_bytes) {
	return [inputs, slice_bytes](void* dst, uint64_t first, uint64_t count) {
		parallel_copy(dst, (const char*)inputs + first * slice_bytes, (size_t)count * slice_bytes);
		return PARHIP_OK;
	};
}

// Slices [first, first + count) gathered from a list of host batches whose
// running slice offsets are `starts` (starts[b] = first slice of batch b,
// starts.back() = total); a chunk may span several batches.
static FillFn batches_fill(const uint64_t* const* batches, const std::vector<uint64_t>& starts,
                           size_t slice_bytes) {
	return [batches, &starts, slice_bytes](void* dst, uint64_t first, uint64_t count) {
		const uint64_t end = first + count;
		size_t b = std::upper_bound(starts.begin(), starts.end(), first) - starts.begin() - 1;
		for (uint64_t j = first; j < end; b++) {
			const uint64_t take = std::min(end, starts[b + 1]) - j;
			if (take == 0) continue;  // empty batch
			parallel_copy((char*)dst + (j - first) * slice_bytes,
			              (const char*)batches[b] + (j - starts[b]) * slice_bytes,
			              (size_t)take * slice_bytes);
			j += take;
		}
		return PARHIP_OK;
	};
}

// Inputs larger than one staging chunk are streamed (overlapped copies);
// smaller ones take one copy.
static bool use_streaming(size_t in_bytes) { return in_bytes > ingest_chunk_bytes(); }

extern "C" {

// ----------------------------------------------------------- GF64Controller
int parhip_gf64_compute_recovery_blocks_with_coeff(const uint64_t* inputs, size_t numInputs,
                                                   uint64_t* recovery, size_t numRecovery,
                                                   size_t blockSize64, const uint64_t* coeff,
                                                   int numThreads, int accumulate) {
	(void)numThreads;
	if (numInputs == 0 || numRecovery == 0) return PARHIP_OK;
	if (!inputs || !recovery || !coeff)
		return fail(PARHIP_ERR_ARG, "compute_recovery_blocks_with_coeff: null pointer");
	if (blockSize64 == 0) return PARHIP_OK;
	size_t in_bytes, out_bytes, cb;
	RC(words_bytes(numInputs, blockSize64, &in_bytes));
	RC(words_bytes(numRecovery, blockSize64, &out_bytes));
	RC(words_bytes(numRecovery, numInputs, &cb));
	hipStream_t s;
	RC(host_stream(&s));
	if (use_streaming(in_bytes)) {
		DevBuf dout, dco;
		RC(dout.alloc(out_bytes, s));
		RC(dco.alloc(cb, s));
		RC(h2d(dco.p, coeff, cb, s));
		if (accumulate) RC(h2d(dout.p, recovery, out_bytes, s));
		RC(stream_contract(memcpy_fill(inputs, blockSize64 * 8), numInputs, blockSize64,
		                   dout.as<uint64_t>(), numRecovery, dco.as<uint64_t>(), numInputs, accumulate,
		                   s));
		RC(d2h(recovery, dout.p, out_bytes, s));
		return sync(s);
	}
	DevBuf din, dout, dco;
	RC(din.alloc(in_bytes, s));
	RC(dout.alloc(out_bytes, s));
	RC(dco.alloc(cb, s));
	RC(h2d(din.p, inputs, in_bytes, s));
	RC(h2d(dco.p, coeff, cb, s));
	if (accumulate) RC(h2d(dout.p, recovery, out_bytes, s));
	RC(parhip_d_gf64_recovery_with_coeff(din.as<uint64_t>(), numInputs, blockSize64,
	                                     dout.as<uint64_t>(), numRecovery, blockSize64, blockSize64,
	                                     dco.as<uint64_t>(), accumulate, s));
	RC(d2h(recovery, dout.p, out_bytes, s));
	return sync(s);
}

int parhip_gf64_compute_recovery_blocks(const uint64_t* inputs, size_t numInputs,
                                        uint64_t* recovery, size_t numRecovery,
                                        size_t blockSize64, uint64_t firstInput,
                                        uint64_t firstRecovery, int numThreads, int accumulate) {
	(void)numThreads;
	if (numInputs == 0 || numRecovery == 0) return PARHIP_OK;
	if (!inputs || !recovery)
		return fail(PARHIP_ERR_ARG, "compute_recovery_blocks: null pointer");
	if (blockSize64 == 0) return PARHIP_OK;
	size_t in_bytes, out_bytes;
	RC(words_bytes(numInputs, blockSize64, &in_bytes));
	RC(words_bytes(numRecovery, blockSize64, &out_bytes));
	hipStream_t s;
	RC(host_stream(&s));
	if (use_streaming(in_bytes)) {
		DevBuf dout;
		RC(dout.alloc(out_bytes, s));
		if (accumulate) RC(h2d(dout.p, recovery, out_bytes, s));
		RC(stream_cauchy(memcpy_fill(inputs, blockSize64 * 8), numInputs, blockSize64,
		                 dout.as<uint64_t>(), numRecovery, firstInput, firstRecovery, accumulate, s));
		RC(d2h(recovery, dout.p, out_bytes, s));
		return sync(s);
	}
	DevBuf din, dout;
	RC(din.alloc(in_bytes, s));
	RC(dout.alloc(out_bytes, s));
	RC(h2d(din.p, inputs, in_bytes, s));
	if (accumulate) RC(h2d(dout.p, recovery, out_bytes, s));
	RC(parhip_d_gf64_recovery(din.as<uint64_t>(), numInputs, blockSize64, dout.as<uint64_t>(),
	                          numRecovery, blockSize64, blockSize64, firstInput, firstRecovery,
	                          accumulate, s));
	RC(d2h(recovery, dout.p, out_bytes, s));
	return sync(s);
}

int parhip_gf64_compute_recovery_blocks_barycentric(const uint64_t* inputs, size_t numInputs,
                                                    uint64_t* recovery, size_t numRecovery,
                                                    size_t blockSize64, uint64_t firstInput,
                                                    uint64_t firstRecovery, int numThreads,
                                                    int accumulate) {
	// Both of the reference's branches compute the Cauchy product
	// (src/par3_engine_barycentric.cc:121-128 delegate, :170-229 per-row
	// inverse + muladd_arr).
	return parhip_gf64_compute_recovery_blocks(inputs, numInputs, recovery, numRecovery,
	                                           blockSize64, firstInput, firstRecovery, numThreads,
	                                           accumulate);
}

int parhip_gf64_compute_recovery_blocks_fenger(const uint64_t* inputs, size_t numInputs,
                                               uint64_t* recovery, size_t numRecovery,
                                               size_t blockSize64, uint64_t firstInput,
                                               uint64_t firstRecovery, int numThreads) {
	return parhip_gf64_compute_recovery_blocks(inputs, numInputs, recovery, numRecovery,
	                                           blockSize64, firstInput, firstRecovery, numThreads, 0);
}

int parhip_gf64_compute_recovery_batches(const uint64_t* const* batches, const size_t* batchBlocks,
                                         size_t numBatches, uint64_t* recovery, size_t numInputs,
                                         size_t numRecovery, size_t blockSize64,
                                         uint64_t firstInput, uint64_t firstRecovery,
                                         int numThreads) {
	(void)numThreads;
	if (numInputs == 0 || numRecovery == 0) return PARHIP_OK;
	if (!recovery || (numBatches && (!batches || !batchBlocks)))
		return fail(PARHIP_ERR_ARG, "compute_recovery_batches: null pointer");
	std::vector<uint64_t> starts(numBatches + 1, 0);
	for (size_t b = 0; b < numBatches; b++) {
		if (batchBlocks[b] && !batches[b])
			return fail(PARHIP_ERR_ARG, "compute_recovery_batches: batch %zu is null", b);
		if (batchBlocks[b] > numInputs - starts[b])
			return fail(PARHIP_ERR_RANGE, "compute_recovery_batches: batches hold more than %zu slices",
			            numInputs);
		starts[b + 1] = starts[b] + batchBlocks[b];
	}
	if (starts[numBatches] != numInputs)
		return fail(PARHIP_ERR_RANGE, "compute_recovery_batches: batches hold %llu slices, numInputs is %zu",
		            (unsigned long long)starts[numBatches], numInputs);
	if (blockSize64 == 0) return PARHIP_OK;
	size_t in_bytes, out_bytes;
	RC(words_bytes(numInputs, blockSize64, &in_bytes));
	RC(words_bytes(numRecovery, blockSize64, &out_bytes));
	hipStream_t s;
	RC(host_stream(&s));
	const size_t slice_bytes = blockSize64 * 8;
	DevBuf dout;
	RC(dout.alloc(out_bytes, s));
	if (use_streaming(in_bytes)) {
		RC(stream_cauchy(batches_fill(batches, starts, slice_bytes), numInputs, blockSize64,
		                 dout.as<uint64_t>(), numRecovery, firstInput, firstRecovery, 0, s));
	} else {
		DevBuf din;
		RC(din.alloc(in_bytes, s));
		for (size_t b = 0; b < numBatches; b++)
			if (batchBlocks[b])
				RC(h2d(din.as<char>() + starts[b] * slice_bytes, batches[b], batchBlocks[b] * slice_bytes, s));
		RC(parhip_d_gf64_recovery(din.as<uint64_t>(), numInputs, blockSize64, dout.as<uint64_t>(),
		                          numRecovery, blockSize64, blockSize64, firstInput, firstRecovery, 0,
		                          s));
	}
	RC(d2h(recovery, dout.p, out_bytes, s));
	return sync(s);
}

int parhip_gf64_compute_recovery_blocks_full(const uint64_t* inputs, size_t numInputs,
                                             uint64_t* recovery, size_t numRecovery,
                                             size_t blockSize64, uint64_t firstInput,
                                             uint64_t firstRecovery, int numThreads,
                                             int accumulate) {
	// The reference's Full variant is the same value computation on its
	// single-pass path (src/par3_engine.cc:1389-1399).
	return parhip_gf64_compute_recovery_blocks(inputs, numInputs, recovery, numRecovery,
	                                           blockSize64, firstInput, firstRecovery, numThreads,
	                                           accumulate);
}

int parhip_gf64_build_cauchy_matrix(uint64_t* coeffMatrix, size_t numInputs, size_t numRecovery,
                                    uint64_t firstInput, uint64_t firstRecovery) {
	if (numInputs == 0 || numRecovery == 0) return PARHIP_OK;
	if (!coeffMatrix) return fail(PARHIP_ERR_ARG, "build_cauchy_matrix: null output");
	size_t cb;
	RC(words_bytes(numRecovery, numInputs, &cb));
	hipStream_t s;
	RC(host_stream(&s));
	DevBuf d;
	RC(d.alloc(cb, s));
	RC(parhip_d_gf64_build_cauchy(d.as<uint64_t>(), numInputs, numRecovery, firstInput,
	                              firstRecovery, s));
	RC(d2h(coeffMatrix, d.p, cb, s));
	return sync(s);
}

uint64_t* parhip_gf64_build_cauchy_matrix_alloc(size_t numInputs, size_t numRecovery,
                                                uint64_t firstInput, uint64_t firstRecovery) {
	if (numInputs == 0 || numRecovery == 0) return nullptr;
	size_t cb;
	if (words_bytes(numRecovery, numInputs, &cb)) return nullptr;
	uint64_t* m = (uint64_t*)malloc(cb);
	if (!m) return nullptr;
	if (parhip_gf64_build_cauchy_matrix(m, numInputs, numRecovery, firstInput, firstRecovery)) {
		free(m);
		return nullptr;
	}
	return m;
}

int parhip_gf64_multiply_accumulate(uint64_t* out, size_t numOut, const uint64_t* in,
                                    size_t numIn, const uint64_t* coeffMatrix,
                                    size_t blockSize64) {
	return parhip_gf64_compute_recovery_blocks_with_coeff(in, numIn, out, numOut, blockSize64,
	                                                      coeffMatrix, 0, 1);
}

int parhip_gf64_solve_and_reconstruct(uint64_t* A, uint64_t* rhsBlocks, size_t n,
                                      size_t blockSize64, int numThreads) {
	(void)numThreads;
	if (n == 0) return PARHIP_OK;
	if (!A || (!rhsBlocks && blockSize64))
		return fail(PARHIP_ERR_ARG, "solve_and_reconstruct: null pointer");
	size_t eb;
	RC(words_bytes(n, n, &eb));
	std::vector<uint64_t> E(n * n);
	const int singular = host_gauss_jordan(A, E.data(), n);
	// rhs <- E * rhs on the device (E = A^-1, or the partial transform the
	// reference has applied when it stops on a singular column).
	if (blockSize64) {
		size_t rb;
		RC(words_bytes(n, blockSize64, &rb));
		hipStream_t s;
		RC(host_stream(&s));
		DevBuf drhs, dout, dE;
		RC(drained(s, [&]() -> int {
			RC(drhs.alloc(rb, s));
			RC(dout.alloc(rb, s));
			RC(dE.alloc(eb, s));
			RC(h2d(drhs.p, rhsBlocks, rb, s));
			RC(h2d(dE.p, E.data(), eb, s));
			RC(parhip_d_gf64_recovery_with_coeff(drhs.as<uint64_t>(), n, blockSize64,
			                                     dout.as<uint64_t>(), n, blockSize64, blockSize64,
			                                     dE.as<uint64_t>(), 0, s));
			RC(d2h(rhsBlocks, dout.p, rb, s));
			return PARHIP_OK;
		}));
	}
	if (singular) return fail(PARHIP_ERR_SINGULAR, "solve_and_reconstruct: singular matrix");
	return PARHIP_OK;
}

int parhip_gf64_solve(const uint64_t* A, const uint64_t* b, uint64_t* x, size_t n) {
	if (n == 0) return PARHIP_OK;
	if (!A || !b || !x) return fail(PARHIP_ERR_ARG, "gf64_solve: null pointer");
	std::vector<uint64_t> Ac(A, A + n * n), E(n * n);
	if (host_gauss_jordan(Ac.data(), E.data(), n))
		return fail(PARHIP_ERR_SINGULAR, "gf64_solve: singular matrix");
	for (size_t i = 0; i < n; i++) {
		uint64_t s = 0;
		for (size_t j = 0; j < n; j++) s ^= host_gf64_mul(E[i * n + j], b[j]);
		x[i] = s;
	}
	return PARHIP_OK;
}

// ----------------------------------------------------------- region operations
int parhip_gf64_region_mul(uint64_t* out, const uint64_t* in, size_t len, uint64_t constant) {
	if (len == 0) return PARHIP_OK;
	if (!out || !in) return fail(PARHIP_ERR_ARG, "region_mul: null pointer");
	uint64_t* outs[1] = {out};
	const uint64_t* ins[1] = {in};
	return host_contract(outs, 1, ins, 1, &constant, 1, len, 0);
}

int parhip_gf64_region_mul_arr(uint64_t* out, const uint64_t* in, const uint64_t* coeff,
                               size_t len, size_t n_coeff) {
	if (len == 0) return PARHIP_OK;
	if (!out || !in || !coeff || n_coeff == 0)
		return fail(PARHIP_ERR_ARG, "region_mul_arr: null pointer or n_coeff == 0");
	if (n_coeff == 1) return parhip_gf64_region_mul(out, in, len, coeff[0]);
	size_t bytes, cb;
	RC(words_bytes(1, len, &bytes));
	RC(words_bytes(1, n_coeff, &cb));
	hipStream_t s;
	RC(host_stream(&s));
	DevBuf din, dout, dco;
	return drained(s, [&]() -> int {
		RC(din.alloc(bytes, s));
		RC(dout.alloc(bytes, s));
		RC(dco.alloc(cb, s));
		RC(h2d(din.p, in, bytes, s));
		RC(h2d(dco.p, coeff, cb, s));
		HIPC(launch_region_mul_arr(dout.as<uint64_t>(), din.as<uint64_t>(), dco.as<uint64_t>(), len,
		                           n_coeff, s),
		     "k_region_mul_arr");
		RC(d2h(out, dout.p, bytes, s));
		return PARHIP_OK;
	});
}

int parhip_gf64_region_muladd_arr(uint64_t* out, const uint64_t* in, const uint64_t* coeff,
                                  size_t len, size_t n_coeff) {
	if (len == 0 || n_coeff == 0) return PARHIP_OK;
	if (!out || !in || !coeff) return fail(PARHIP_ERR_ARG, "region_muladd_arr: null pointer");
	// XOR_c in*coeff[c] == in * (XOR_c coeff[c]) by distributivity.
	uint64_t c = 0;
	for (size_t i = 0; i < n_coeff; i++) c ^= coeff[i];
	uint64_t* outs[1] = {out};
	const uint64_t* ins[1] = {in};
	return host_contract(outs, 1, ins, 1, &c, 1, len, 1);
}

int parhip_gf64_region_coupled_muladd_arr(uint64_t* out, const uint64_t* const* in_blocks,
                                          const uint64_t* coeff_blocks, size_t len, size_t G) {
	if (len == 0 || G == 0) return PARHIP_OK;
	if (!out || !in_blocks || !coeff_blocks)
		return fail(PARHIP_ERR_ARG, "region_coupled_muladd_arr: null pointer");
	for (size_t g = 0; g < G; g++)
		if (!in_blocks[g]) return fail(PARHIP_ERR_ARG, "region_coupled_muladd_arr: null block");
	uint64_t* outs[1] = {out};
	return host_contract(outs, 1, in_blocks, G, coeff_blocks, G, len, 1);
}

int parhip_gf64_region_fused_output_muladd_arr(uint64_t* const* outs, const uint64_t* in,
                                               const uint64_t* const* coeff_block_starts,
                                               size_t len, size_t K) {
	if (len == 0 || K == 0) return PARHIP_OK;
	if (!outs || !in || !coeff_block_starts)
		return fail(PARHIP_ERR_ARG, "region_fused_output_muladd_arr: null pointer");
	std::vector<uint64_t> c(K);
	for (size_t k = 0; k < K; k++) {
		if (!outs[k] || !coeff_block_starts[k])
			return fail(PARHIP_ERR_ARG, "region_fused_output_muladd_arr: null entry");
		c[k] = *coeff_block_starts[k];
	}
	const uint64_t* ins[1] = {in};
	return host_contract(outs, K, ins, 1, c.data(), 1, len, 1);
}

int parhip_gf64_region_2d_muladd_arr(uint64_t* const* outs, size_t K,
                                     const uint64_t* const* in_blocks, size_t G,
                                     const uint64_t* coeff_block_2d, size_t K_stride,
                                     size_t len) {
	if (len == 0 || K == 0 || G == 0) return PARHIP_OK;
	if (!outs || !in_blocks || !coeff_block_2d)
		return fail(PARHIP_ERR_ARG, "region_2d_muladd_arr: null pointer");
	if (K_stride < G) return fail(PARHIP_ERR_ARG, "region_2d_muladd_arr: K_stride < G");
	for (size_t k = 0; k < K; k++)
		if (!outs[k]) return fail(PARHIP_ERR_ARG, "region_2d_muladd_arr: null output");
	for (size_t g = 0; g < G; g++)
		if (!in_blocks[g]) return fail(PARHIP_ERR_ARG, "region_2d_muladd_arr: null input");
	return host_contract(outs, K, in_blocks, G, coeff_block_2d, K_stride, len, 1);
}

// ---------------------------------------------------------- streaming create
// src/gf64_create_streaming.cc:93-240 (Par3csExecute): read the source file,
// zero-pad the tail to a whole block, compute every recovery block with
// ComputeRecoveryBlocksFull.  Here the file is read chunk by chunk straight
// into pinned staging slots (pread on the host copy threads) and each chunk
// is reduced on the device while the next one is read.
// Chunks of a file (blocks of S bytes, `total` bytes in all) read straight
// into the pinned staging slots by the host copy threads; bytes past the end
// of the file read as zeros.
static FillFn file_fill(int fd, size_t total, size_t S, std::atomic<int>* read_err) {
	return [fd, total, S, read_err](void* dst, uint64_t first, uint64_t count) {
		const size_t off = (size_t)first * S, bytes = (size_t)count * S, part = 4u << 20;
		parallel_parts((bytes + part - 1) / part, [&](size_t p) {
			const size_t o = p * part, len = std::min(part, bytes - o);
			size_t got = 0;
			while (got < len && off + o + got < total) {
				const size_t want = std::min(len - got, total - (off + o + got));
				const ssize_t r = ::pread(fd, (char*)dst + o + got, want, (off_t)(off + o + got));
				if (r < 0 && errno == EINTR) continue;
				if (r <= 0) {
					int expect = 0;
					read_err->compare_exchange_strong(expect, r < 0 ? errno : EIO);
					return;
				}
				got += (size_t)r;
			}
			if (got < len) memset((char*)dst + o + got, 0, len - got);  // zero tail
		});
		const int err = read_err->load();
		if (err) return fail(PARHIP_ERR_ARG, "read failed: errno=%d (%s)", err, strerror(err));
		return (int)PARHIP_OK;
	};
}

static const char* errno_code(int err) {
	switch (err) {
		case ENOENT: return "ENOENT";
		case EACCES: return "EACCES";
		case EISDIR: return "EISDIR";
		case EFBIG: return "EFBIG";
		case ENOMEM: return "ENOMEM";
		case ENAMETOOLONG: return "ENAMETOOLONG";
		default: return "EIO";
	}
}

int parhip_par3_create_streaming(const char* sourcePath, int32_t numRecovery, int64_t blockSize,
                                 uint64_t firstInput, uint64_t firstRecovery, int32_t numThreads,
                                 int useMmap, uint8_t** recoveryOut, size_t* recoveryBytes,
                                 double* durationMs, double* throughputMBps) {
	(void)numThreads;
	(void)useMmap;  // the data path is pread -> pinned -> HBM either way
	if (!recoveryOut) return fail(PARHIP_ERR_ARG, "par3_create_streaming: null result pointer");
	*recoveryOut = nullptr;
	if (recoveryBytes) *recoveryBytes = 0;
	if (!sourcePath || !sourcePath[0])
		return fail(PARHIP_ERR_ARG, "sourcePath must be a non-empty string");
	if (numRecovery <= 0) return fail(PARHIP_ERR_RANGE, "recoverySlices must be a positive integer");
	if (blockSize <= 0 || blockSize % 8 != 0)
		return fail(PARHIP_ERR_RANGE, "blockSize must be positive and a multiple of 8");
	const int fd = ::open(sourcePath, O_RDONLY);
	if (fd < 0) {
		const int err = errno;
		return fail(PARHIP_ERR_ARG, "%s: open failed for sourcePath: errno=%d (%s)", errno_code(err),
		            err, strerror(err));
	}
	struct stat st;
	if (::fstat(fd, &st) != 0) {
		const int err = errno;
		::close(fd);
		return fail(PARHIP_ERR_ARG, "fstat failed: errno=%d (%s)", err, strerror(err));
	}
	if (st.st_size <= 0) {
		::close(fd);
		return fail(PARHIP_ERR_ARG, "sourcePath is empty");
	}
	const size_t total = (size_t)st.st_size, S = (size_t)blockSize, W = S / 8;
	const uint64_t N = (total + S - 1) / S;
	const size_t R = (size_t)numRecovery;
	size_t out_bytes, cb;
	int rc = words_bytes(R, W, &out_bytes);
	if (!rc) rc = words_bytes(R, N, &cb);
	uint8_t* rec = nullptr;
	if (!rc) {
		rec = (uint8_t*)aligned_alloc(64, (out_bytes + 63) & ~(size_t)63);
		if (!rec) rc = fail(PARHIP_ERR_NOMEM, "aligned_alloc failed for recovery buffer");
	}
	const auto t0 = std::chrono::steady_clock::now();
	hipStream_t s = nullptr;
	if (!rc) rc = host_stream(&s);
	if (!rc) {
		DevBuf dout;
		rc = dout.alloc(out_bytes, s);
		std::atomic<int> read_err{0};
		const FillFn fill = file_fill(fd, total, S, &read_err);
		if (!rc) rc = stream_cauchy(fill, N, W, dout.as<uint64_t>(), R, firstInput, firstRecovery, 0, s);
		if (!rc) rc = d2h(rec, dout.p, out_bytes, s);
		if (!rc) rc = sync(s);
	}
	::close(fd);
	if (rc) {
		free(rec);
		return rc;
	}
	const double ms =
	    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	*recoveryOut = rec;
	if (recoveryBytes) *recoveryBytes = out_bytes;
	if (durationMs) *durationMs = ms;
	if (throughputMBps) *throughputMBps = ms > 0 ? ((double)out_bytes / 1048576.0) / (ms / 1000.0) : 0;
	return PARHIP_OK;
}

// GF64Controller::ComputeRecoveryBlocksFromFile (src/par3_engine.cc:1445-1537):
// the inputs are the file's blocks (its size a positive multiple of the block
// size), the recovery blocks go into the caller's buffer.  The reference maps
// the file behind the PAR3_GF64_USE_MMAP=1 gate and returns -1 without it; the
// same gate and the same -1 for every failure are kept here, while the data
// path is the chunked read into pinned memory of the streaming create (the
// device reduces each chunk while the next one is read).
int parhip_gf64_compute_recovery_blocks_from_file(const char* sourcePath, uint64_t* recovery,
                                                  size_t numRecovery, size_t blockSize64,
                                                  uint64_t firstInput, uint64_t firstRecovery,
                                                  int numThreads) {
	(void)numThreads;
	const char* env = std::getenv("PAR3_GF64_USE_MMAP");
	if (!env || !*env || std::atoi(env) != 1) {
		fail(PARHIP_ERR_ARG, "PAR3_GF64_USE_MMAP=0; use compute_recovery_blocks with the inputs in memory");
		return -1;
	}
	if (!sourcePath || !recovery) {
		fail(PARHIP_ERR_ARG, "compute_recovery_blocks_from_file: null sourcePath or recovery");
		return -1;
	}
	if (blockSize64 == 0) {
		fail(PARHIP_ERR_ARG, "compute_recovery_blocks_from_file: blockSize64 == 0");
		return -1;
	}
	const int fd = ::open(sourcePath, O_RDONLY);
	if (fd < 0) {
		const int err = errno;
		fail(PARHIP_ERR_ARG, "compute_recovery_blocks_from_file: open(%s) failed: %s", sourcePath,
		     strerror(err));
		return -1;
	}
	struct stat st;
	if (::fstat(fd, &st) != 0) {
		const int err = errno;
		::close(fd);
		fail(PARHIP_ERR_ARG, "compute_recovery_blocks_from_file: fstat(%s) failed: %s", sourcePath,
		     strerror(err));
		return -1;
	}
	const size_t total = (size_t)st.st_size, S = blockSize64 * 8;
	if (blockSize64 > SIZE_MAX / 8 || total < S || total % S != 0) {
		::close(fd);
		fail(PARHIP_ERR_ARG,
		     "compute_recovery_blocks_from_file: file size %zu is not a positive multiple of "
		     "blockSize64 * 8 = %zu", total, S);
		return -1;
	}
	const uint64_t N = total / S;
	int rc = PARHIP_OK;
	if (numRecovery) {
		size_t out_bytes;
		hipStream_t s = nullptr;
		rc = words_bytes(numRecovery, blockSize64, &out_bytes);
		if (!rc) rc = host_stream(&s);
		DevBuf dout;
		if (!rc) rc = dout.alloc(out_bytes, s);
		std::atomic<int> read_err{0};
		if (!rc)
			rc = stream_cauchy(file_fill(fd, total, S, &read_err), N, blockSize64, dout.as<uint64_t>(),
			                   numRecovery, firstInput, firstRecovery, 0, s);
		if (!rc) rc = d2h(recovery, dout.p, out_bytes, s);
		if (!rc) rc = sync(s);
	}
	::close(fd);
	return rc ? -1 : 0;
}

}  // extern "C"
