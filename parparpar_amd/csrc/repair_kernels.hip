// repair_kernels.hip -- device side of PAR3 repair (k lost slices).
//
// The reference repairs by Gauss-Jordan elimination of the k x k Cauchy
// sub-matrix A[eq][m] = 1/(y_eq + x_m) (recovery point y_eq = firstRecovery+eq,
// missing input x_m) applied to the RHS blocks (src/par3_engine.cc:980-1048,
// lib/par3gen.js:3840-4012): O(k^3) scalar work plus O(k^2) region MACs.  The
// inverse of a nonsingular matrix is unique, so any exact method gives the
// same bytes; here the closed form of a Cauchy matrix's inverse is used,
//
//   (A^-1)[m][eq] = A[eq][m] * alpha_eq * beta_m,
//   alpha_eq = prod_m' (y_eq + x_m') / prod_{eq' != eq} (y_eq + y_eq'),
//   beta_m   = prod_eq' (y_eq' + x_m) / prod_{m' != m} (x_m + x_m'),
//
// (characteristic 2: every difference is a XOR), O(k^2) products and k^2 + 2k
// inversions, all on the device.  The RHS build and the application of A^-1
// are both contractions and run on the table kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf64_dev.h"
#include "kernels.h"

namespace parhip {

// alpha (which == 0) or beta (which == 1) factors: one work-group per factor,
// products over the k terms reduced in LDS.
__global__ __launch_bounds__(256) void k_cauchy_inv_factors(const uint64_t* __restrict__ ys,
                                                            const uint64_t* __restrict__ xs,
                                                            uint64_t k, uint64_t* __restrict__ alpha,
                                                            uint64_t* __restrict__ beta) {
	__shared__ uint64_t num_s[256], den_s[256];
	const uint64_t idx = blockIdx.x;  // [0, k) -> alpha, [k, 2k) -> beta
	const bool is_alpha = idx < k;
	const uint64_t me = is_alpha ? ys[idx] : xs[idx - k];
	const uint64_t self = is_alpha ? idx : idx - k;
	const uint64_t* other = is_alpha ? xs : ys;  // numerator terms
	const uint64_t* same = is_alpha ? ys : xs;   // denominator terms
	uint64_t num = 1, den = 1;
	for (uint64_t t = threadIdx.x; t < k; t += blockDim.x) {
		num = gf64_mul(num, me ^ other[t]);
		if (t != self) den = gf64_mul(den, me ^ same[t]);
	}
	num_s[threadIdx.x] = num;
	den_s[threadIdx.x] = den;
	__syncthreads();
	for (unsigned s = blockDim.x / 2; s > 0; s >>= 1) {
		if (threadIdx.x < s) {
			num_s[threadIdx.x] = gf64_mul(num_s[threadIdx.x], num_s[threadIdx.x + s]);
			den_s[threadIdx.x] = gf64_mul(den_s[threadIdx.x], den_s[threadIdx.x + s]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		const uint64_t f = gf64_mul(num_s[0], gf64_inv(den_s[0]));
		if (is_alpha)
			alpha[self] = f;
		else
			beta[self] = f;
	}
}

// inv[m][eq] = alpha_eq * beta_m / (y_eq + x_m)   (k x k, row-major by m)
__global__ void k_cauchy_inverse(const uint64_t* __restrict__ ys, const uint64_t* __restrict__ xs,
                                 uint64_t k, const uint64_t* __restrict__ alpha,
                                 const uint64_t* __restrict__ beta, uint64_t* __restrict__ inv) {
	const uint64_t total = k * k;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t m = idx / k, eq = idx - m * k;
		const uint64_t d = ys[eq] ^ xs[m];
		inv[idx] = gf64_mul(gf64_mul(gf64_inv(d ? d : 1), alpha[eq]), beta[m]);
	}
}

hipError_t launch_cauchy_inverse(const uint64_t* ys, const uint64_t* xs, uint64_t k,
                                 uint64_t* alpha, uint64_t* beta, uint64_t* inv, hipStream_t s) {
	if (!k) return hipSuccess;
	hipLaunchKernelGGL(k_cauchy_inv_factors, dim3((unsigned)(2 * k)), dim3(256), 0, s, ys, xs, k,
	                   alpha, beta);
	uint64_t blocks = (k * k + 255) / 256;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_cauchy_inverse, dim3((unsigned)blocks), dim3(256), 0, s, ys, xs, k, alpha,
	                   beta, inv);
	return hipGetLastError();
}

// RHS coefficients of a repair as one k x ld matrix over the inputs
// [known slices..., used recovery slices...]:
//   out[eq][c] = 1/(xs_known[c] + ys[eq])       for c <  nk
//   out[eq][nk + eq'] = (eq == eq')             the recovery block itself
__global__ void k_repair_coeffs(uint64_t* __restrict__ out, uint64_t ld,
                                const uint64_t* __restrict__ xs_known, uint64_t nk,
                                const uint64_t* __restrict__ ys, uint64_t k) {
	const uint64_t total = k * (nk + k);
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t eq = idx / (nk + k), c = idx - eq * (nk + k);
		uint64_t v;
		if (c < nk) {
			const uint64_t d = xs_known[c] ^ ys[eq];
			v = gf64_inv(d ? d : 1);
		} else {
			v = (c - nk == eq) ? 1u : 0u;
		}
		out[eq * ld + c] = v;
	}
}

hipError_t launch_repair_coeffs(uint64_t* out, uint64_t ld, const uint64_t* xs_known, uint64_t nk,
                                const uint64_t* ys, uint64_t k, hipStream_t s) {
	const uint64_t total = k * (nk + k);
	if (!total) return hipSuccess;
	uint64_t blocks = (total + 255) / 256;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_repair_coeffs, dim3((unsigned)blocks), dim3(256), 0, s, out, ld, xs_known,
	                   nk, ys, k);
	return hipGetLastError();
}

// Gather/scatter helper: dst[r] row pointers for a list of slice indices.
__global__ void k_slice_ptrs(const uint64_t* __restrict__ base, uint64_t stride,
                             const uint64_t* __restrict__ idx, uint64_t n,
                             const uint64_t** __restrict__ out) {
	for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	     t += (uint64_t)gridDim.x * blockDim.x)
		out[t] = base + idx[t] * stride;
}

hipError_t launch_slice_ptrs(const uint64_t* base, uint64_t stride, const uint64_t* idx, uint64_t n,
                             const uint64_t** out, hipStream_t s) {
	if (!n) return hipSuccess;
	uint64_t blocks = (n + 255) / 256;
	if (blocks > 4096) blocks = 4096;
	hipLaunchKernelGGL(k_slice_ptrs, dim3((unsigned)blocks), dim3(256), 0, s, base, stride, idx, n,
	                   out);
	return hipGetLastError();
}

}  // namespace parhip
