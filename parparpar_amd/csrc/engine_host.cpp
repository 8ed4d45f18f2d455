// engine_host.cpp -- host-pointer entries of the C ABI (include/parhip.h).
//
// These mirror GF64Controller's static methods and the region operations
// exactly as the reference's N-API addon calls them: caller-owned host
// buffers in, results written back before return.  Each call stages its
// operands into HBM on the calling thread's stream, runs the device
// contraction (engine.cpp / gf64_kernels.hip) and copies the results back.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "gf64_host.h"
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
static int host_contract(uint64_t* const* outs, size_t K, const uint64_t* const* ins, size_t G,
                         const uint64_t* coeff, size_t coeff_ld, size_t len, int accumulate) {
	if (K == 0 || G == 0 || len == 0) return PARHIP_OK;
	size_t in_bytes, out_bytes, blk = len * sizeof(uint64_t), cb;
	RC(words_bytes(G, len, &in_bytes));
	RC(words_bytes(K, len, &out_bytes));
	RC(words_bytes(K, G, &cb));
	hipStream_t s;
	RC(host_stream(&s));
	DevBuf din, dout, dco;
	RC(din.alloc(in_bytes, s));
	RC(dout.alloc(out_bytes, s));
	RC(dco.alloc(cb, s));
	for (size_t g = 0; g < G; g++) RC(h2d(din.as<uint64_t>() + g * len, ins[g], blk, s));
	if (accumulate)
		for (size_t k = 0; k < K; k++) RC(h2d(dout.as<uint64_t>() + k * len, outs[k], blk, s));
	std::vector<uint64_t> c(K * G);
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
	return sync(s);
}

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
		RC(drhs.alloc(rb, s));
		RC(dout.alloc(rb, s));
		RC(dE.alloc(eb, s));
		RC(h2d(drhs.p, rhsBlocks, rb, s));
		RC(h2d(dE.p, E.data(), eb, s));
		RC(parhip_d_gf64_recovery_with_coeff(drhs.as<uint64_t>(), n, blockSize64,
		                                     dout.as<uint64_t>(), n, blockSize64, blockSize64,
		                                     dE.as<uint64_t>(), 0, s));
		RC(d2h(rhsBlocks, dout.p, rb, s));
		RC(sync(s));
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
	RC(din.alloc(bytes, s));
	RC(dout.alloc(bytes, s));
	RC(dco.alloc(cb, s));
	RC(h2d(din.p, in, bytes, s));
	RC(h2d(dco.p, coeff, cb, s));
	HIPC(launch_region_mul_arr(dout.as<uint64_t>(), din.as<uint64_t>(), dco.as<uint64_t>(), len,
	                           n_coeff, s),
	     "k_region_mul_arr");
	RC(d2h(out, dout.p, bytes, s));
	return sync(s);
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

}  // extern "C"
