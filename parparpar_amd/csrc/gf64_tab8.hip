// gf64_tab8.hip -- GF(2^64) contraction with 8-bit split tables in LDS.
//
// The byte-table form of the table strategy: each coefficient's product is
// split into 8 tables of 256 entries, T_k[n] = c * n(x) * x^{8k}, so a word
// costs 8 lookups instead of the 16 of the nibble tables (gf64_kernels.hip).
// Two outputs share a 16-byte entry, as there.  What it trades: a 256-entry
// row is 4 KiB, so a wave's random byte values collide on LDS banks (entries
// n and n + 16 share banks), and a table is 16x larger to build per
// coefficient.
//
// Build: row (input, pair, k) is written by 16 threads, thread n_lo writing
// entries 16 n_hi + n_lo for n_hi = 0..15 as low(n_lo) ^ high(n_hi), where
// low = c * n_lo * x^{8k} and high(n_hi) = c * n_hi * x^{8k+4} is stepped
// through the 16 nibble values by XOR of a 4-element basis; in each store the
// 16 lanes of a group write 16 consecutive entries (no write conflicts).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf64_dev.h"
#include "kernels.h"

namespace parhip {

namespace {
__device__ __forceinline__ uint32_t t8xor3(uint32_t a, uint32_t b, uint32_t c) {
	return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
}  // namespace

template <int RT, int WPT, int IB>
__global__ __launch_bounds__(256) void k_mac_tab8(MacArgs a) {
	constexpr int NP = RT / 2;
	constexpr int ROWS = IB * NP * 8;        // 256-entry rows per batch
	constexpr int RPT = (ROWS * 16 + 255) / 256;  // rows per thread group of 16
	__shared__ uint4 tab[ROWS * 256];
#define TAB8(ib, p, k, n) tab[((((ib) * NP + (p)) * 8 + (k)) << 8) + (n)]
	const int tid = threadIdx.x;
	const uint64_t w0 = (uint64_t)blockIdx.x * 256 * WPT;
	const uint64_t r0 = (uint64_t)blockIdx.z * RT;
	const uint64_t g = blockIdx.y;
	const uint64_t i_begin = g * a.in_per_group;
	uint64_t i_end = i_begin + a.in_per_group;
	if (i_end > a.N) i_end = a.N;

	uint4 acc[WPT][NP];
#pragma unroll
	for (int j = 0; j < WPT; j++)
#pragma unroll
		for (int p = 0; p < NP; p++) acc[j][p] = make_uint4(0, 0, 0, 0);

	for (uint64_t ib0 = i_begin; ib0 < i_end; ib0 += IB) {
		uint2 d[IB][WPT];
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
			const uint64_t i = ib0 + ib;
			const uint64_t* src = i < i_end ? a.in_base + i * a.in_stride : nullptr;
#pragma unroll
			for (int j = 0; j < WPT / 2; j++) {
				const uint64_t w = w0 + 2 * tid + 512 * j;
				uint64_t v0 = 0, v1 = 0;
				if (src) {
					if (w + 1 < a.W && ((reinterpret_cast<uintptr_t>(src) & 15u) == 0)) {
						const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(src + w);
						v0 = v.x;
						v1 = v.y;
					} else {
						if (w < a.W) v0 = src[w];
						if (w + 1 < a.W) v1 = src[w + 1];
					}
				}
				d[ib][2 * j] = make_uint2((uint32_t)v0, (uint32_t)(v0 >> 32));
				d[ib][2 * j + 1] = make_uint2((uint32_t)v1, (uint32_t)(v1 >> 32));
			}
		}
		__syncthreads();  // previous batch finished reading tab
#pragma unroll
		for (int rr = 0; rr < RPT; rr++) {
			const int t = tid + 256 * rr;
			const int row = t >> 4, nlo = t & 15;
			if (row < ROWS) {
				const int k = row & 7;
				const int p = (row >> 3) % NP;
				const int ib = (row >> 3) / NP;
				const uint64_t i = ib0 + ib;
				uint64_t c[2] = {0, 0};
				if (i < i_end) {
					const uint64_t ra = r0 + 2 * p;
					if (ra < a.R) c[0] = a.coeff[ra * a.coeff_ld + i];
					if (ra + 1 < a.R) c[1] = a.coeff[(ra + 1) * a.coeff_ld + i];
				}
				uint64_t low[2], hb[2][4];
#pragma unroll
				for (int o = 0; o < 2; o++) {
					const uint64_t ck = gf64_mulx_pow_any(c[o], 8u * k);  // c * x^{8k}
					low[o] = gf64_mul_nibble(ck, (unsigned)nlo);
					hb[o][0] = gf64_mulx4(ck);  // basis of high(n_hi): c * x^{8k+4+b}
#pragma unroll
					for (int b = 1; b < 4; b++) hb[o][b] = gf64_mulx_pow(hb[o][b - 1], 1);
				}
				uint64_t hi0 = 0, hi1 = 0;  // high(n_hi) in Gray-code order
#pragma unroll
				for (int s = 0; s < 16; s++) {
					const int nhi = s ^ (s >> 1);
					if (s) {
						const int bit = __builtin_ctz(s);  // the basis element that changes
						hi0 ^= hb[0][bit];
						hi1 ^= hb[1][bit];
					}
					const uint64_t e0 = low[0] ^ hi0, e1 = low[1] ^ hi1;
					TAB8(ib, p, k, (nhi << 4) | nlo) =
					    make_uint4((uint32_t)e0, (uint32_t)(e0 >> 32), (uint32_t)e1, (uint32_t)(e1 >> 32));
				}
			}
		}
		__syncthreads();
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
#pragma unroll
			for (int j = 0; j < WPT; j++) {
				const uint2 x = d[ib][j];
#pragma unroll
				for (int kq = 0; kq < 8; kq += 2) {
					const uint32_t half = kq < 4 ? x.x : x.y;
					const unsigned n0 = (half >> (8 * (kq & 3))) & 255u;
					const unsigned n1 = (half >> (8 * ((kq + 1) & 3))) & 255u;
#pragma unroll
					for (int p = 0; p < NP; p++) {
						const uint4 v0 = TAB8(ib, p, kq, n0);
						const uint4 v1 = TAB8(ib, p, kq + 1, n1);
						uint4 c = acc[j][p];
						c.x = t8xor3(c.x, v0.x, v1.x);
						c.y = t8xor3(c.y, v0.y, v1.y);
						c.z = t8xor3(c.z, v0.z, v1.z);
						c.w = t8xor3(c.w, v0.w, v1.w);
						acc[j][p] = c;
					}
				}
			}
		}
	}
#undef TAB8
	// write out (direct or split-K partial)
#pragma unroll
	for (int r = 0; r < RT; r++) {
		const uint64_t rr = r0 + r;
		if (rr >= a.R) break;
		const int p = r >> 1;
#pragma unroll
		for (int j = 0; j < WPT / 2; j++) {
			const uint64_t w = w0 + 2 * tid + 512 * j;
#pragma unroll
			for (int u = 0; u < 2; u++) {
				if (w + u >= a.W) continue;
				const uint4 c = acc[2 * j + u][p];
				const uint64_t v = (r & 1) ? (((uint64_t)c.w << 32) | c.z) : (((uint64_t)c.y << 32) | c.x);
				if (a.partial) {
					a.partial[(g * a.R + rr) * a.W + w + u] = v;
				} else {
					uint64_t* dst = a.out_base + rr * a.out_stride + w + u;
					*dst = a.accumulate ? (*dst ^ v) : v;
				}
			}
		}
	}
}

template <int RT, int WPT, int IB>
static hipError_t launch_tab8_t(const MacArgs& a, uint64_t groups, hipStream_t s) {
	const uint64_t tile = 256ull * WPT;
	dim3 grid((unsigned)((a.W + tile - 1) / tile), (unsigned)groups, (unsigned)((a.R + RT - 1) / RT));
	hipLaunchKernelGGL((k_mac_tab8<RT, WPT, IB>), grid, dim3(256), 0, s, a);
	return hipGetLastError();
}

bool mac_tab8_supported(int rt, int wpt, int ib) {
	return (rt == 2 || rt == 4) && (wpt == 2 || wpt == 4 || (wpt == 8 && ib == 1)) && (ib == 1 || ib == 2);
}

hipError_t launch_mac_tab8(const MacArgs& a, uint64_t groups, int rt, int wpt, int ib, hipStream_t s) {
#define T8(R_, W_, I_) \
	if (rt == R_ && wpt == W_ && ib == I_) return launch_tab8_t<R_, W_, I_>(a, groups, s);
	T8(2, 2, 1) T8(2, 2, 2) T8(2, 4, 1) T8(2, 4, 2) T8(2, 8, 1)
	T8(4, 2, 1) T8(4, 2, 2) T8(4, 4, 1) T8(4, 4, 2) T8(4, 8, 1)
#undef T8
	return hipErrorInvalidValue;
}

}  // namespace parhip
