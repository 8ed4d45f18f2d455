// gf16_inv.hip -- PAR2 repair matrix on the device: Gauss-Jordan inversion of
// the missing-input block of the recovery matrix over GF(2^16).
//
// Galois16RecMatrix::Compute (gf16/gfmat_inv.cpp:562-676) eliminates, pivot by
// pivot and without row exchanges, the full recovery x input matrix in place:
// the pivot entry is replaced by 1 before its row is scaled by the pivot's
// reciprocal, every other row's entry in the pivot column is replaced by 0
// before that row receives factor x pivot row.  The missing-input block ends
// up holding its inverse A^-1 and the valid block A^-1 x C_valid.  Here only
// A (numRec x numRec) is eliminated, one pivot per launch, reading one buffer
// and writing the other (every entry of a step depends on the previous step's
// pivot row and pivot column); A^-1 x C_valid is one GF(2^16) contraction on
// the PAR2 kernels afterwards.  A zero pivot is reported through *bad (the
// first such pivot; later launches see it and do nothing), exactly the row
// the reference drops before it retries.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace parhip {

__global__ __launch_bounds__(256) void k_gf16_gj_step(const uint16_t* __restrict__ in,
                                                      uint16_t* __restrict__ out, uint32_t n,
                                                      uint32_t k,
                                                      const uint16_t* __restrict__ exp_tab,
                                                      const uint16_t* __restrict__ log_tab,
                                                      int* __restrict__ bad) {
	if (*bad >= 0) return;
	const uint16_t p = in[(uint64_t)k * n + k];
	if (p == 0) {
		if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *bad = (int)k;
		return;
	}
	const uint32_t lip = (65535u - log_tab[p]) % 65535u;  // log(1/p)
	const uint32_t r = blockIdx.y;
	for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
		// pivot row after "replace pivot by 1, scale by 1/p": s_c = (c == k ? 1 : A[k][c]) / p
		const uint16_t kc = (c == k) ? (uint16_t)1 : in[(uint64_t)k * n + c];
		const uint32_t ls = kc ? (log_tab[kc] + lip) % 65535u : 0xFFFFFFFFu;
		if (r == k) {
			out[(uint64_t)k * n + c] = kc ? exp_tab[ls] : (uint16_t)0;
		} else {
			const uint16_t t = in[(uint64_t)r * n + k];
			uint16_t v = (c == k) ? (uint16_t)0 : in[(uint64_t)r * n + c];
			if (t && kc) v ^= exp_tab[(log_tab[t] + ls) % 65535u];
			out[(uint64_t)r * n + c] = v;
		}
	}
}

__global__ void k_set_int(int* p, int v) { *p = v; }

// n pivots; returns (through *result) the buffer holding A^-1: a or b.
hipError_t launch_gf16_invert(uint16_t* a, uint16_t* b, uint32_t n, const uint16_t* exp_tab,
                              const uint16_t* log_tab, int* bad, uint16_t** result, hipStream_t s) {
	hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, bad, -1);
	uint16_t* cur = a;
	uint16_t* nxt = b;
	const uint32_t bx = (n + 255) / 256 > 64 ? 64 : (n + 255) / 256;
	for (uint32_t k = 0; k < n; k++) {
		hipLaunchKernelGGL(k_gf16_gj_step, dim3(bx, n), dim3(256), 0, s, cur, nxt, n, k, exp_tab,
		                   log_tab, bad);
		uint16_t* t = cur;
		cur = nxt;
		nxt = t;
	}
	*result = cur;
	return hipGetLastError();
}

}  // namespace parhip
