// gf16_kernels.hip -- PAR2 GF(2^16) muladd kernels for gfx950 (MI355X).
//
//     out[e][w] ^= XOR_i coeff(i, e) * in_i[w]      16-bit LE words, mod 0x1100B
//
// the contraction behind PAR2 create (gf16/controller_cpu.cpp:463-537,
// controller_ocl_init.cpp:115-200).  Coefficients follow the PAR2 rule
// coeff(i, e) = 2^(log_i * e mod 65535) with log_i the i-th exponent not
// divisible by 3, 5, 17 or 257 (gf16/gfmat_coeff.c:9-101); they are computed
// on the device from two constant tables.
//
// k_gf16_tab: split nibble tables in LDS.  For every input and group of eight
// outputs, four 16-entry tables T_k[n] = (c_e * n x^{4k}) for the 8 outputs
// packed in one 16-byte entry: one ds_read_b128 yields a nibble's products for
// 8 outputs, and a 16-entry table fills exactly one 256-byte LDS row (no bank
// conflicts for any nibble values).  Each thread owns 4 consecutive words (one
// 64-bit load per input) and 8*NG outputs of accumulators in VGPRs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf16_dev.h"
#include "kernels.h"

namespace parhip {

__device__ __forceinline__ uint32_t xor3_16(uint32_t a, uint32_t b, uint32_t c) {
	return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// coeff[e*N + i] = exp((inlog[inputNums[i]] * exps[e]) mod 65535)
__global__ void k_gf16_coeffs(uint16_t* __restrict__ coeff, const uint16_t* __restrict__ input_nums,
                              uint64_t N, const uint16_t* __restrict__ exps, uint64_t R,
                              const uint16_t* __restrict__ exp_tab,
                              const uint16_t* __restrict__ inlog_tab) {
	const uint64_t total = N * R;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t e = idx / N, i = idx - e * N;
		const uint32_t l = inlog_tab[input_nums[i] & 32767u];
		const uint32_t p = (uint32_t)(((uint64_t)l * exps[e]) % 65535u);
		coeff[idx] = exp_tab[p];
	}
}

// Loads of one batch of IB inputs for k_gf16_tab: this thread's 4 words of
// each input (zeros past the input range or the slice end) and, for thread
// t < IB*4*NG, the coefficients of output pair t % (4*NG) of input t / (4*NG)
// (build phase 1).  Coefficients first, so a wait for them leaves the data
// loads in flight.
template <int NG, int IB>
__device__ __forceinline__ void gf16_tab_load(const Gf16Args& a, uint64_t ib0, uint64_t i_end,
                                              uint64_t w0, uint64_t r0, uint2 (&d)[IB],
                                              uint32_t& c0, uint32_t& c1) {
	const int tid = threadIdx.x;
	c0 = 0;
	c1 = 0;
	if (tid < IB * 4 * NG) {
		const int pr = tid % (4 * NG);
		const int ib = tid / (4 * NG);
		const uint64_t i = ib0 + ib, e0 = r0 + 2 * pr;
		if (i < i_end) {
			if (e0 < a.R) c0 = a.coeff[e0 * a.coeff_ld + i];
			if (e0 + 1 < a.R) c1 = a.coeff[(e0 + 1) * a.coeff_ld + i];
		}
	}
#pragma unroll
	for (int ib = 0; ib < IB; ib++) {
		const uint64_t i = ib0 + ib;
		uint64_t v = 0;
		if (i < i_end) {
			const uint16_t* src = a.in_base + i * a.in_stride;
			if (w0 + 3 < a.W && ((reinterpret_cast<uintptr_t>(src) & 7u) == 0)) {
				v = *reinterpret_cast<const uint64_t*>(src + w0);
			} else {
#pragma unroll
				for (int u = 0; u < 4; u++)
					if (w0 + u < a.W) v |= (uint64_t)src[w0 + u] << (16 * u);
			}
		}
		d[ib] = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
	}
}

template <int NG, int IB>
__global__ __launch_bounds__(256, 2) void k_gf16_tab(Gf16Args a) {
	static_assert(IB * 4 * NG <= 256, "one basis pair per thread");
	constexpr int TILE = 256 * 4;  // words per work-group
	__shared__ uint4 tab[IB][NG][4][16];
	// [input][j][output pair]; rows padded by 16 B so the 4 position rows a wave
	// reads together (j = 4k + b, k = 0..3) fall on different banks
	__shared__ __align__(16) uint32_t bas[IB][16][4 * NG + 4];

	const int tid = threadIdx.x;
	const uint64_t w0 = (uint64_t)blockIdx.x * TILE + 4u * tid;
	const uint64_t r0 = (uint64_t)blockIdx.z * (8 * NG);
	const uint64_t g = blockIdx.y;
	const uint64_t i_begin = g * a.in_per_group;
	uint64_t i_end = i_begin + a.in_per_group;
	if (i_end > a.N) i_end = a.N;

	// acc[j][q] : word w0+j, dword d of output group q -> outputs 8q+2d (lo16), 8q+2d+1 (hi16)
	uint4 acc[4][NG];
#pragma unroll
	for (int j = 0; j < 4; j++)
#pragma unroll
		for (int q = 0; q < NG; q++) acc[j][q] = make_uint4(0, 0, 0, 0);

	// Software pipeline: batch b+1's loads are issued before batch b's lookups.
	uint2 d[IB];
	uint32_t c0, c1;
	gf16_tab_load<NG, IB>(a, i_begin, i_end, w0, r0, d, c0, c1);
	for (uint64_t ib0 = i_begin; ib0 < i_end; ib0 += IB) {
		__syncthreads();  // previous batch finished reading tab / bas
		// build, phase 1: basis c_e * x^j (j < 16) of every (input, output) of
		// the batch, an output pair per dword: bas[ib][j][pair].  One thread per
		// (input, pair) writes whole dwords; the lanes of a store cover
		// consecutive dwords of one row (no bank conflicts).
		if (tid < IB * 4 * NG) {
			const int pr = tid % (4 * NG);
			const int ib = tid / (4 * NG);
			uint32_t b0 = c0, b1 = c1;
#pragma unroll
			for (int j = 0; j < 16; j++) {
				bas[ib][j][pr] = b0 | (b1 << 16);
				b0 = gf16_mulx(b0);
				b1 = gf16_mulx(b1);
			}
		}
		__syncthreads();
		// phase 2: tab[ib][q][k][n] = XOR of the basis rows j = 4k + b for the set
		// bits b of n, output pairs 4q..4q+3 in the four dwords.  Two threads per
		// (ib, q, k) row, entries 8*hf .. 8*hf + 7 each, from the row's four basis
		// vectors held in registers: the first entry from all four, then one XOR
		// per entry in Gray-code order (g(s) = s ^ s >> 1 changes one bit per
		// step).  The low three bits are offset by the row (XOR with row & 7), so
		// the 16 lanes of a write pass (8 rows x 2 halves) write 16 different
		// entries of their 256-byte rows: no bank conflicts.
		for (int t = tid; t < IB * NG * 4 * 2; t += 256) {
			const int row = t >> 1;
			const unsigned hf = t & 1;
			const int k = row & 3;
			const int q = (row >> 2) % NG;
			const int ib = (row >> 2) / NG;
			uint4 bv[4];
#pragma unroll
			for (int b = 0; b < 4; b++) bv[b] = *reinterpret_cast<const uint4*>(&bas[ib][4 * k + b][4 * q]);
			const unsigned off = (unsigned)row & 7u;
			uint4 ent = make_uint4(0, 0, 0, 0);
#pragma unroll
			for (int b = 0; b < 4; b++) {
				const uint32_t m = 0u - (((off | (hf << 3)) >> b) & 1u);
				ent.x ^= bv[b].x & m;
				ent.y ^= bv[b].y & m;
				ent.z ^= bv[b].z & m;
				ent.w ^= bv[b].w & m;
			}
			uint4* dst = &tab[ib][q][k][0];
#pragma unroll
			for (int st = 0; st < 8; st++) {
				if (st) {
					// bit that g(st) flips against g(st - 1): the lowest set bit of st
					const int fb = st & 1 ? 0 : (st & 2 ? 1 : 2);
					ent.x ^= bv[fb].x;
					ent.y ^= bv[fb].y;
					ent.z ^= bv[fb].z;
					ent.w ^= bv[fb].w;
				}
				dst[(hf << 3) | ((unsigned)(st ^ (st >> 1)) ^ off)] = ent;
			}
		}
		__syncthreads();
		// next batch's loads, in flight during this batch's lookups
		uint2 dn[IB];
		uint32_t n0c, n1c;
		gf16_tab_load<NG, IB>(a, ib0 + IB, i_end, w0, r0, dn, n0c, n1c);
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const uint32_t word = (j < 2 ? d[ib].x : d[ib].y) >> (16 * (j & 1));
				const unsigned n0 = word & 15u, n1 = (word >> 4) & 15u, n2 = (word >> 8) & 15u,
				               n3 = (word >> 12) & 15u;
#pragma unroll
				for (int q = 0; q < NG; q++) {
					const uint4 v0 = tab[ib][q][0][n0];
					const uint4 v1 = tab[ib][q][1][n1];
					const uint4 v2 = tab[ib][q][2][n2];
					const uint4 v3 = tab[ib][q][3][n3];
					uint4 c = acc[j][q];
					c.x = xor3_16(xor3_16(c.x, v0.x, v1.x), v2.x, v3.x);
					c.y = xor3_16(xor3_16(c.y, v0.y, v1.y), v2.y, v3.y);
					c.z = xor3_16(xor3_16(c.z, v0.z, v1.z), v2.z, v3.z);
					c.w = xor3_16(xor3_16(c.w, v0.w, v1.w), v2.w, v3.w);
					acc[j][q] = c;
				}
			}
		}
#pragma unroll
		for (int ib = 0; ib < IB; ib++) d[ib] = dn[ib];
		c0 = n0c;
		c1 = n1c;
	}

	// write: output e = r0 + 8q + m, words w0..w0+3
#pragma unroll
	for (int q = 0; q < NG; q++) {
#pragma unroll
		for (int m = 0; m < 8; m++) {
			const uint64_t e = r0 + 8 * q + m;
			if (e >= a.R) continue;
			uint16_t wv[4];
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const uint4 c = acc[j][q];
				const uint32_t dw = (m >> 1) == 0 ? c.x : (m >> 1) == 1 ? c.y : (m >> 1) == 2 ? c.z : c.w;
				wv[j] = (uint16_t)((m & 1) ? (dw >> 16) : dw);
			}
			uint16_t* dst;
			bool xr;
			if (a.partial) {
				dst = a.partial + (g * a.R + e) * a.W;
				xr = false;
			} else {
				dst = a.out_base + e * a.out_stride;
				xr = a.accumulate != 0;
			}
			if (w0 + 3 < a.W && ((reinterpret_cast<uintptr_t>(dst) & 7u) == 0)) {
				uint64_t v = (uint64_t)wv[0] | ((uint64_t)wv[1] << 16) | ((uint64_t)wv[2] << 32) |
				             ((uint64_t)wv[3] << 48);
				uint64_t* p = reinterpret_cast<uint64_t*>(dst + w0);
				if (xr) v ^= *p;
				*p = v;
			} else {
#pragma unroll
				for (int j = 0; j < 4; j++)
					if (w0 + j < a.W) dst[w0 + j] = xr ? (uint16_t)(dst[w0 + j] ^ wv[j]) : wv[j];
			}
		}
	}
}

// Fold split-K partials: dst[h][e][w] = XOR of F partials (one thread per 4 words).
__global__ void k_gf16_fold(const uint16_t* __restrict__ src, uint64_t G, uint64_t RW, uint64_t F,
                            uint16_t* __restrict__ dst) {
	const uint64_t H = (G + F - 1) / F;
	const uint64_t total = H * RW;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t h = idx / RW, e = idx - h * RW;
		const uint64_t g0 = h * F, g1 = (g0 + F < G) ? g0 + F : G;
		uint16_t s = 0;
		for (uint64_t gg = g0; gg < g1; gg++) s ^= src[gg * RW + e];
		dst[idx] = s;
	}
}

__global__ void k_gf16_reduce(const uint16_t* __restrict__ partial, uint64_t G, uint64_t R,
                              uint64_t W, uint16_t* out_base, uint64_t out_stride, int accumulate) {
	const uint64_t total = R * W;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t e = idx / W, w = idx - e * W;
		uint16_t s = 0;
		for (uint64_t gg = 0; gg < G; gg++) s ^= partial[(gg * R + e) * W + w];
		uint16_t* dst = out_base + e * out_stride;
		dst[w] = accumulate ? (uint16_t)(dst[w] ^ s) : s;
	}
}

// ---------------------------------------------------------------- launchers
hipError_t launch_gf16_coeffs(uint16_t* coeff, const uint16_t* input_nums, uint64_t N,
                              const uint16_t* exps, uint64_t R, const uint16_t* exp_tab,
                              const uint16_t* inlog_tab, hipStream_t s) {
	const uint64_t total = N * R;
	if (!total) return hipSuccess;
	uint64_t blocks = (total + 255) / 256;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_gf16_coeffs, dim3((unsigned)blocks), dim3(256), 0, s, coeff, input_nums,
	                   N, exps, R, exp_tab, inlog_tab);
	return hipGetLastError();
}

template <int NG, int IB>
static hipError_t launch_gf16_tab_t(const Gf16Args& a, uint64_t groups, hipStream_t s) {
	const uint64_t tiles = (a.W + 1023) / 1024;
	const uint64_t rtiles = (a.R + 8 * NG - 1) / (8 * NG);
	dim3 grid((unsigned)tiles, (unsigned)groups, (unsigned)rtiles);
	hipLaunchKernelGGL((k_gf16_tab<NG, IB>), grid, dim3(256), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_gf16_tab(const Gf16Args& a, uint64_t groups, int ng, hipStream_t s) {
	switch (ng) {
		case 1: return launch_gf16_tab_t<1, 4>(a, groups, s);
		case 2: return launch_gf16_tab_t<2, 4>(a, groups, s);
		case 4: return launch_gf16_tab_t<4, 4>(a, groups, s);
		case 8: return launch_gf16_tab_t<8, 4>(a, groups, s);
		default: return hipErrorInvalidValue;
	}
}

uint64_t gf16_reduce_scratch_words(uint64_t G, uint64_t R, uint64_t W) {
	const uint64_t F = 16;
	return (G + (G > F ? (G + F - 1) / F : 0)) * R * W;
}

hipError_t launch_gf16_reduce(uint16_t* partial, uint64_t G, uint64_t R, uint64_t W,
                              uint16_t* out_base, uint64_t out_stride, int accumulate,
                              hipStream_t s) {
	const uint64_t RW = R * W;
	if (!RW) return hipSuccess;
	const uint64_t F = 16;
	uint16_t* front = partial;
	uint16_t* tail = partial + G * RW;
	uint16_t* src = front;
	while (G > F) {
		uint16_t* dst = (src == front) ? tail : front;
		const uint64_t H = (G + F - 1) / F;
		uint64_t blocks = (H * RW + 255) / 256;
		if (blocks > 65536) blocks = 65536;
		hipLaunchKernelGGL(k_gf16_fold, dim3((unsigned)blocks), dim3(256), 0, s, src, G, RW, F, dst);
		hipError_t e = hipGetLastError();
		if (e != hipSuccess) return e;
		src = dst;
		G = H;
	}
	uint64_t blocks = (RW + 255) / 256;
	if (blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(k_gf16_reduce, dim3((unsigned)blocks), dim3(256), 0, s, src, G, R, W, out_base,
	                   out_stride, accumulate);
	return hipGetLastError();
}

}  // namespace parhip
