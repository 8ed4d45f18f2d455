// gf64_dyadic.hip -- the GF(2^64) Cauchy contraction through the XOR-block
// structure of its coefficient matrix (gfx950): the plain kernel, the
// coefficient kernels and their launchers.  The method and the device code are
// described in gf64_dyadic_dev.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf64_dyadic16_dev.h"
#include "gf64_dyadic32_dev.h"
#include "gf64_dyadic_dev.h"

namespace parhip {

// Work-group: word tile blockIdx.x, split-K row blockIdx.y, output group
// blockIdx.z.  FAST: every tile is full and 16-byte aligned (checked by the
// launcher).  MASK: inputs flagged in a.lost count as zero.  GP: coefficients
// from a.gp (else computed here).
template <int WPT, bool FAST, bool MASK, bool GP, bool PIPE = false>
__global__ __launch_bounds__(256, WPT == 2 ? 4 : 2) void k_mac_dyadic(DyadicArgs a) {
	// tab[(tbl*16 + k)*16 + n] = (g'_SA * n x^{4k}, g'_SB * n x^{4k})
	__shared__ uint4 tab[kDyTables * 256];
	__shared__ uint64_t coef[kDyChunk * 8];  // coefficients computed here (a.gp null)
	const uint64_t g = blockIdx.y;
	const uint64_t b_begin = g * a.blocks_per_group;
	uint64_t b_end = b_begin + a.blocks_per_group;
	if (b_end > a.nblocks) b_end = a.nblocks;
	dy_tile<WPT, FAST, MASK, GP, PIPE>(a, tab, coef, blockIdx.z, (uint64_t)blockIdx.x * 256 * WPT, b_begin, b_end,
	                         g, a.partial ? DY_OUT_PARTIAL : (a.atomic_out ? DY_OUT_ATOMIC : DY_OUT_DIRECT));
}

// The 16 x 16 form (gf64_dyadic16_dev.h): same grid, blocks and groups of 16
// points; 58 KiB of LDS, two work-groups per CU.
template <int WPT, bool FAST, bool GP, bool MASK = false>
__global__ __launch_bounds__(256, 2) void k_mac_dyadic16(DyadicArgs a) {
	__shared__ uint4 tab[2 * kDyTables * 256];   // two sets of seven pair tables
	__shared__ uint64_t coef[kD16Chunk * 16];    // coefficients computed here (GP false)
	const uint64_t g = blockIdx.y;
	const uint64_t b_begin = g * a.blocks_per_group;
	uint64_t b_end = b_begin + a.blocks_per_group;
	if (b_end > a.nblocks) b_end = a.nblocks;
	d16_tile<WPT, FAST, GP, MASK>(a, tab, coef, blockIdx.z, (uint64_t)blockIdx.x * 256 * WPT, b_begin, b_end, g,
	                        a.partial ? DY_OUT_PARTIAL : (a.atomic_out ? DY_OUT_ATOMIC : DY_OUT_DIRECT));
}

// The 32 x 32 form (gf64_dyadic32_dev.h): tiles of 512 words, one work-group
// of 512 threads per CU (116 KiB of LDS).
template <bool FAST, bool GP, bool MASK>
__global__ __launch_bounds__(kD32Threads, 1) void k_mac_dyadic32(DyadicArgs a) {
	__shared__ uint4 tab[4 * kDyTables * 256];  // four sets of seven pair tables
	__shared__ uint64_t coef[kD32Chunk * 32];   // coefficients computed here (GP false)
	const uint64_t g = blockIdx.y;
	const uint64_t b_begin = g * a.blocks_per_group;
	uint64_t b_end = b_begin + a.blocks_per_group;
	if (b_end > a.nblocks) b_end = a.nblocks;
	d32_tile<FAST, GP, MASK>(a, tab, coef, blockIdx.z, (uint64_t)blockIdx.x * kD32Threads, b_begin, b_end, g,
	                         a.partial ? DY_OUT_PARTIAL : (a.atomic_out ? DY_OUT_ATOMIC : DY_OUT_DIRECT));
}

// Transformed coefficients of every (output group q, input block b) of the
// 32 x 32 form: g(u) = f(X_b XOR Y_q XOR u), X_b = xbase + 32b,
// Y_q = ybase + 32q, out[(q*nblocks + b)*32 + S] = XOR over supersets T of S
// of g(T); 8 (q, b) pairs per work-group.
__global__ __launch_bounds__(256) void k_dyadic32_coeffs(uint64_t* __restrict__ out, uint64_t nblocks,
                                                         uint64_t Q, uint64_t xbase, uint64_t ybase) {
	const uint64_t total = nblocks * Q;
	const int u = threadIdx.x & 31;
	for (uint64_t base_pair = (uint64_t)blockIdx.x * 8; base_pair < total; base_pair += (uint64_t)gridDim.x * 8) {
		const uint64_t idx = base_pair + (threadIdx.x >> 5);
		uint64_t c = 0;
		if (idx < total) {
			const uint64_t q = idx / nblocks, b = idx - q * nblocks;
			const uint64_t z = ((xbase + 32 * b) ^ (ybase + 32 * q)) ^ (uint64_t)u;
			c = z ? gf64_inv(z) : 1u;  // the reference's 0 -> 1 (src/par3_engine.cc:392-394)
		}
#pragma unroll
		for (int bit = 1; bit < 32; bit <<= 1) {
			const uint32_t lo = __shfl_xor((uint32_t)c, bit, 64);
			const uint32_t hi = __shfl_xor((uint32_t)(c >> 32), bit, 64);
			if (!(u & bit)) c ^= ((uint64_t)hi << 32) | lo;
		}
		if (idx < total) out[idx * 32 + u] = c;
	}
}

hipError_t launch_dyadic32_coeffs(uint64_t* out, uint64_t nblocks, uint64_t Q, uint64_t xbase,
                                  uint64_t ybase, hipStream_t s) {
	const uint64_t total = nblocks * Q;
	if (!total) return hipSuccess;
	uint64_t blocks = (total + 7) / 8;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_dyadic32_coeffs, dim3((unsigned)blocks), dim3(256), 0, s, out, nblocks, Q, xbase,
	                   ybase);
	return hipGetLastError();
}

hipError_t launch_mac_dyadic32(const DyadicArgs& a, uint64_t groups, hipStream_t s) {
	const uint64_t Q = (a.pre_out + a.R + 31) / 32;
	const uint64_t tiles = (a.W + kD32Threads - 1) / kD32Threads;
	if (!tiles || !groups || !Q) return hipSuccess;
	if (tiles > 0x7fffffffu || groups > 65535 || Q > 65535) return hipErrorInvalidValue;
	if (a.lost && !a.gp) return hipErrorInvalidValue;
	dim3 grid((unsigned)tiles, (unsigned)groups, (unsigned)Q);
	const bool fast = a.W % kD32Threads == 0;
	const int v = (a.lost ? 4 : 0) | (fast ? 2 : 0) | (a.gp ? 1 : 0);
	switch (v) {
#define D32_CASE(V, F, G, M) \
	case V: hipLaunchKernelGGL((k_mac_dyadic32<F, G, M>), grid, dim3(kD32Threads), 0, s, a); break;
		D32_CASE(7, true, true, true)
		D32_CASE(5, false, true, true)
		D32_CASE(3, true, true, false)
		D32_CASE(2, true, false, false)
		D32_CASE(1, false, true, false)
		default: hipLaunchKernelGGL((k_mac_dyadic32<false, false, false>), grid, dim3(kD32Threads), 0, s, a); break;
#undef D32_CASE
	}
	return hipGetLastError();
}

// Transformed coefficients of every (output group q, input block b) of the
// 16 x 16 form: g(u) = f(X_b XOR Y_q XOR u), X_b = xbase + 16b,
// Y_q = ybase + 16q, out[(q*nblocks + b)*16 + S] = XOR over supersets T of S
// of g(T); 16 (q, b) pairs per work-group.
__global__ __launch_bounds__(256) void k_dyadic16_coeffs(uint64_t* __restrict__ out, uint64_t nblocks,
                                                         uint64_t Q, uint64_t xbase, uint64_t ybase) {
	__shared__ uint64_t gs[256];
	const uint64_t total = nblocks * Q;
	const int u = threadIdx.x & 15;
	for (uint64_t base_pair = (uint64_t)blockIdx.x * 16; base_pair < total;
	     base_pair += (uint64_t)gridDim.x * 16) {
		const uint64_t idx = base_pair + (threadIdx.x >> 4);
		uint64_t c = 0;
		if (idx < total) {
			const uint64_t q = idx / nblocks, b = idx - q * nblocks;
			const uint64_t z = ((xbase + 16 * b) ^ (ybase + 16 * q)) ^ (uint64_t)u;
			c = z ? gf64_inv(z) : 1u;
		}
		gs[threadIdx.x] = c;
		__syncthreads();
		uint64_t acc = 0;
#pragma unroll
		for (int t = 0; t < 16; t++)
			if ((t & u) == u) acc ^= gs[(threadIdx.x & ~15) + t];
		if (idx < total) out[idx * 16 + u] = acc;
		__syncthreads();
	}
}

hipError_t launch_dyadic16_coeffs(uint64_t* out, uint64_t nblocks, uint64_t Q, uint64_t xbase,
                                  uint64_t ybase, hipStream_t s) {
	const uint64_t total = nblocks * Q;
	if (!total) return hipSuccess;
	uint64_t blocks = (total + 15) / 16;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_dyadic16_coeffs, dim3((unsigned)blocks), dim3(256), 0, s, out, nblocks, Q,
	                   xbase, ybase);
	return hipGetLastError();
}

hipError_t launch_mac_dyadic16(const DyadicArgs& a, uint64_t groups, int wpt, hipStream_t s) {
	const uint64_t Q = (a.pre_out + a.R + 15) / 16;
	const uint64_t tiles = (a.W + 256ull * wpt - 1) / (256ull * wpt);
	if (!tiles || !groups || !Q) return hipSuccess;
	if (tiles > 0x7fffffffu || groups > 65535 || Q > 65535) return hipErrorInvalidValue;
	if ((wpt != 1 && wpt != 2) || (a.lost && !a.gp)) return hipErrorInvalidValue;
	dim3 grid((unsigned)tiles, (unsigned)groups, (unsigned)Q);
	const bool fast = a.W % (256ull * wpt) == 0;
	if (a.lost) {
		// lost-input masks (the repair's right-hand side; its coefficients
		// always come from gp)
		if (wpt == 2 && fast) hipLaunchKernelGGL((k_mac_dyadic16<2, true, true, true>), grid, dim3(256), 0, s, a);
		else if (wpt == 2) hipLaunchKernelGGL((k_mac_dyadic16<2, false, true, true>), grid, dim3(256), 0, s, a);
		else if (fast) hipLaunchKernelGGL((k_mac_dyadic16<1, true, true, true>), grid, dim3(256), 0, s, a);
		else hipLaunchKernelGGL((k_mac_dyadic16<1, false, true, true>), grid, dim3(256), 0, s, a);
		return hipGetLastError();
	}
	const int v = (wpt == 2 ? 4 : 0) | (fast ? 2 : 0) | (a.gp ? 1 : 0);
	switch (v) {
#define D16_CASE(V, W, F, G) \
	case V: hipLaunchKernelGGL((k_mac_dyadic16<W, F, G>), grid, dim3(256), 0, s, a); break;
		D16_CASE(7, 2, true, true)
		D16_CASE(6, 2, true, false)
		D16_CASE(5, 2, false, true)
		D16_CASE(4, 2, false, false)
		D16_CASE(3, 1, true, true)
		D16_CASE(2, 1, true, false)
		D16_CASE(1, 1, false, true)
		default: hipLaunchKernelGGL((k_mac_dyadic16<1, false, false>), grid, dim3(256), 0, s, a); break;
#undef D16_CASE
	}
	return hipGetLastError();
}

// Transformed coefficients of every (output group q, input block b):
//   g(u) = f(X_b XOR Y_q XOR u), X_b = xbase + 8b, Y_q = ybase + 8q,
//   out[(q*nblocks + b)*8 + S] = XOR over supersets T of S of g(T).
// One inversion per thread (an inversion is a serial chain of ~100 dependent
// steps), 32 (q, b) pairs per work-group; the 8 values of a pair meet in LDS.
__global__ __launch_bounds__(256) void k_dyadic_coeffs(uint64_t* __restrict__ out, uint64_t nblocks,
                                                       uint64_t Q, uint64_t xbase, uint64_t ybase) {
	__shared__ uint64_t gs[256];
	const uint64_t total = nblocks * Q;
	const int u = threadIdx.x & 7;
	for (uint64_t base_pair = (uint64_t)blockIdx.x * 32; base_pair < total;
	     base_pair += (uint64_t)gridDim.x * 32) {
		const uint64_t idx = base_pair + (threadIdx.x >> 3);
		uint64_t c = 0;
		if (idx < total) {
			const uint64_t q = idx / nblocks, b = idx - q * nblocks;
			const uint64_t z = ((xbase + 8 * b) ^ (ybase + 8 * q)) ^ (uint64_t)u;
			c = z ? gf64_inv(z) : 1u;
		}
		gs[threadIdx.x] = c;
		__syncthreads();
		uint64_t acc = 0;
#pragma unroll
		for (int t = 0; t < 8; t++)
			if ((t & u) == u) acc ^= gs[(threadIdx.x & ~7) + t];
		if (idx < total) out[idx * 8 + u] = acc;
		__syncthreads();
	}
}

hipError_t launch_dyadic_coeffs(uint64_t* out, uint64_t nblocks, uint64_t Q, uint64_t xbase,
                                uint64_t ybase, hipStream_t s) {
	const uint64_t total = nblocks * Q;
	if (!total) return hipSuccess;
	uint64_t blocks = (total + 31) / 32;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_dyadic_coeffs, dim3((unsigned)blocks), dim3(256), 0, s, out, nblocks, Q,
	                   xbase, ybase);
	return hipGetLastError();
}

// Rows of `words` words at `stride` set to zero (a kernel: safe to record in
// a graph, where memset nodes are not).
__global__ void k_zero_rows(uint64_t* __restrict__ out, uint64_t rows, uint64_t words, uint64_t stride) {
	const uint64_t r = blockIdx.y;
	for (uint64_t w = 2 * ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x); w < words;
	     w += 2 * (uint64_t)gridDim.x * blockDim.x) {
		uint64_t* p = out + r * stride + w;
		if (w + 1 < words && ((reinterpret_cast<uintptr_t>(p) & 15u) == 0)) {
			*reinterpret_cast<ulonglong2*>(p) = make_ulonglong2(0, 0);
		} else {
			p[0] = 0;
			if (w + 1 < words) p[1] = 0;
		}
	}
}

hipError_t launch_zero_rows(uint64_t* out, uint64_t rows, uint64_t words, uint64_t stride,
                            hipStream_t s) {
	if (!rows || !words) return hipSuccess;
	if (rows > 65535) return hipErrorInvalidValue;
	uint64_t bx = (words / 2 + 255) / 256;
	if (bx > 1024) bx = 1024;
	if (!bx) bx = 1;
	hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)bx, (unsigned)rows), dim3(256), 0, s, out, rows, words,
	                   stride);
	return hipGetLastError();
}

bool mac_dyadic_supported(int wpt, int ib) { return (wpt == 4 || wpt == 2) && ib == 1; }

hipError_t launch_mac_dyadic(const DyadicArgs& a, uint64_t groups, int wpt, int ib, hipStream_t s) {
	const uint64_t Q = (a.pre_out + a.R + 7) / 8;
	const uint64_t tiles = (a.W + 256ull * wpt - 1) / (256ull * wpt);
	if (!tiles || !groups || !Q) return hipSuccess;
	if (tiles > 0x7fffffffu || groups > 65535 || Q > 65535) return hipErrorInvalidValue;
	dim3 grid((unsigned)tiles, (unsigned)groups, (unsigned)Q);
	const bool fast = ((reinterpret_cast<uintptr_t>(a.in_base) & 15u) == 0) && (a.in_stride & 1) == 0 &&
	                  a.W % (256ull * wpt) == 0;
	if (ib != 1 || (wpt != 4 && wpt != 2)) return hipErrorInvalidValue;
	const int v = (wpt == 4 ? 8 : 0) | (fast ? 4 : 0) | (a.lost ? 2 : 0) | (a.gp ? 1 : 0);
	if (v == 12 && Q == 1) {
		// one output group (R <= 8): tables read one step ahead
		hipLaunchKernelGGL((k_mac_dyadic<4, true, false, false, true>), grid, dim3(256), 0, s, a);
		return hipGetLastError();
	}
	if (v == 13 && Q == 1) {
		hipLaunchKernelGGL((k_mac_dyadic<4, true, false, true, true>), grid, dim3(256), 0, s, a);
		return hipGetLastError();
	}
	switch (v) {
#define DY_CASE(V, W, F, M, G) \
	case V: hipLaunchKernelGGL((k_mac_dyadic<W, F, M, G>), grid, dim3(256), 0, s, a); break;
		DY_CASE(15, 4, true, true, true)
		DY_CASE(14, 4, true, true, false)
		DY_CASE(13, 4, true, false, true)
		DY_CASE(12, 4, true, false, false)
		DY_CASE(11, 4, false, true, true)
		DY_CASE(10, 4, false, true, false)
		DY_CASE(9, 4, false, false, true)
		DY_CASE(8, 4, false, false, false)
		DY_CASE(7, 2, true, true, true)
		DY_CASE(6, 2, true, true, false)
		DY_CASE(5, 2, true, false, true)
		DY_CASE(4, 2, true, false, false)
		DY_CASE(3, 2, false, true, true)
		DY_CASE(2, 2, false, true, false)
		DY_CASE(1, 2, false, false, true)
		default: hipLaunchKernelGGL((k_mac_dyadic<2, false, false, false>), grid, dim3(256), 0, s, a); break;
#undef DY_CASE
	}
	return hipGetLastError();
}

}  // namespace parhip
