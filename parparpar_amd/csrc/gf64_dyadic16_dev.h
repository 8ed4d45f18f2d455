// gf64_dyadic16_dev.h -- the XOR-block Cauchy contraction over 16 x 16 blocks
// (device side), for 16 or more outputs.
//
// The 8 x 8 form (gf64_dyadic_dev.h) groups the points in aligned runs of 8.
// Runs of 16, X = 16b + t and Y = 16q + s with t, s in 0..15, make each
// 16 x 16 block a product in GF(2^64)[Z_2^4], whose subset convolution needs
// 3^4 = 81 field products per block: 0.32 per input and output against 27/64 =
// 0.42 for 8 x 8 blocks (a quarter fewer table reads at R = 16, 32, 64, ...).
//
// Tables: the 16 transformed coefficients g'_U of a block split by their bit 3
// into two sets of 8, F'_0 = (g'_0..g'_7) and F'_1 = (g'_8..g'_15), each
// served by the seven pair tables of the 8 x 8 form (56 KiB of LDS).  An input
// a'_T with bit 3 clear needs both sets, one with bit 3 set only set 0; the
// nibble offsets of an input are shared by the two sets' reads.  So a word of
// a block costs 84 table steps of 8 reads (80 x 8 ds_read_b128 + 4 x 8
// ds_read_b64), against 4 x 28 for the four 8 x 8 blocks it replaces.
// Product g'_S * a'_T goes to h'_(S | T); outputs stay in the subset basis
// over all blocks and are transformed back once before the store.
//
// Points that are not multiples of 16 get virtual points in front (zero
// inputs, outputs not stored), as in the 8 x 8 form.  The launch arguments are
// the same DyadicArgs: nblocks, blocks_per_group, pre_in, pre_out count in
// runs of 16; gp holds 16 words per (output group q, block b) at
// gp + (q*gp_ld + b)*16; lost masks (MASK launches) one u32 per block of 16
// points, bit t = virtual input 16b + t counts as zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf64_dyadic_dev.h"

namespace parhip {

// The 84 table steps of one word: (input subset T, half h, table set, pair
// table; -1 = 8-byte reads of slot a of table 4).  Steps of one (T, h) share
// the nibble offsets.
__device__ constexpr int kD16T[84] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  1,
                                      1,  1,  1,  1,  1,  1,  1,  2,  2,  2,  2,  2,  2,  2,  2,  3,  3,
                                      3,  3,  4,  4,  4,  4,  4,  4,  4,  4,  5,  5,  5,  5,  6,  6,  6,
                                      6,  7,  7,  7,  7,  8,  8,  8,  8,  8,  8,  8,  8,  9,  9,  9,  9,
                                      10, 10, 10, 10, 11, 11, 12, 12, 12, 12, 13, 13, 14, 14, 15, 15};
__device__ constexpr int kD16H[84] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 1,
                                      1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1,
                                      1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 1,
                                      1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 0, 1, 0, 1};
__device__ constexpr int kD16SG[84] = {0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0,
                                       1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1,
                                       0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0,
                                       0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__device__ constexpr int kD16TB[84] = {0,  0,  1,  1,  2,  2,  3, 3, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 4,
                                       4,  5,  5,  0,  0,  2,  2, 0, 0, 2, 2, 6, 6, 6, 6, 0, 0, 1, 1, 0, 0,
                                       1,  1,  4,  4,  4,  4,  0, 0, 0, 0, -1, -1, -1, -1, 0, 1, 2, 3, 0, 1,
                                       2,  3,  4,  5,  4,  5,  0, 2, 0, 2, 6, 6, 0, 1, 0, 1, 4, 4, 0, 0, -1, -1};
constexpr int kD16Steps = 84;
// LDS bytes of one table set (7 pair tables x 16 positions x 16 entries x 16 B)
constexpr int kD16SetBytes = kDyTables * 256 * 16;
// blocks per in-kernel coefficient chunk (16 coefficients each: 256 threads)
constexpr int kD16Chunk = 16;

// 4-bit superset transform of 16 words, in place (its own inverse).
__device__ __forceinline__ void d16_superset(uint2 (&x)[16]) {
#pragma unroll
	for (int bit = 1; bit < 16; bit <<= 1)
#pragma unroll
		for (int s = 0; s < 16; s++)
			if (!(s & bit)) {
				x[s].x ^= x[s | bit].x;
				x[s].y ^= x[s | bit].y;
			}
}

// The transformed coefficients of blocks [b0, b0 + 16) of output group q into
// coef[p*16 + U] (thread 16p + u: block b0 + p, point offset u; one inversion
// each, the 4-bit superset sums across the 16 lanes of a block).
__device__ __forceinline__ void d16_chunk_coeffs(const DyadicArgs& a, uint64_t* coef, uint64_t q,
                                                 uint64_t b0, uint64_t b_end) {
	const int tid = threadIdx.x;
	const int u = tid & 15;
	const uint64_t b = b0 + (uint64_t)(tid >> 4);
	uint64_t c = 0;
	if (b < b_end) {
		const uint64_t z = ((a.xbase + 16 * b) ^ (a.ybase + 16 * q)) ^ (uint64_t)u;
		c = z ? gf64_inv(z) : 1u;  // the reference's 0 -> 1 (src/par3_engine.cc:392-394)
	}
#pragma unroll
	for (int bit = 1; bit < 16; bit <<= 1) {
		const uint32_t lo = __shfl_xor((uint32_t)c, bit, 64);
		const uint32_t hi = __shfl_xor((uint32_t)(c >> 32), bit, 64);
		if (!(u & bit)) c ^= ((uint64_t)hi << 32) | lo;
	}
	coef[tid] = c;
}

// The four coefficients of the entries this thread builds for block b
// (thread (tbl, half, n), tbl < 7): slots a and b of pair table tbl in sets 0
// and 1; zeros past the group's blocks.
template <bool GP>
__device__ __forceinline__ void d16_load_coeffs(const DyadicArgs& a, const uint64_t* coef, uint64_t q,
                                                uint64_t b, uint64_t b_begin, uint64_t b_end,
                                                uint64_t (&c)[4]) {
	const int tid = threadIdx.x;
	c[0] = c[1] = c[2] = c[3] = 0;
	if (tid < kDyTables * 32 && b < b_end) {
		const int tbl = tid >> 5;
		const int sa = (0x00406420u >> (4 * tbl)) & 7u, sb = (0x04627531u >> (4 * tbl)) & 7u;
		const uint64_t* p = GP ? a.gp + (q * a.gp_ld + b) * 16 : coef + ((b - b_begin) % kD16Chunk) * 16;
		if (GP) {
			c[0] = dy_load(p + sa);
			c[1] = dy_load(p + sb);
			c[2] = dy_load(p + 8 + sa);
			c[3] = dy_load(p + 8 + sb);
		} else {
			c[0] = p[sa];
			c[1] = p[sb];
			c[2] = p[8 + sa];
			c[3] = p[8 + sb];
		}
	}
}

// Table build of one block from its four coefficients (both sets); thread
// (tbl, half, n) owns entry n of 8 of the 16 positions of pair table tbl.
__device__ __forceinline__ void d16_build(uint4* tab, const uint64_t (&c)[4]) {
	const int tid = threadIdx.x;
	if (tid < kDyTables * 32) {
		const int n = tid & 15;
		const int half = (tid >> 4) & 1;
		const int tbl = tid >> 5;
#pragma unroll
		for (int sg = 0; sg < 2; sg++) {
			uint64_t e0 = gf64_mul_nibble(c[2 * sg], (unsigned)n);
			uint64_t e1 = gf64_mul_nibble(c[2 * sg + 1], (unsigned)n);
			if (half) {
				e0 = gf64_mulx32(e0);
				e1 = gf64_mulx32(e1);
			}
			uint4* t = tab + sg * (kD16SetBytes / 16);
#pragma unroll
			for (int k = 0; k < 8; k++) {
				if (k) {
					e0 = gf64_mulx4(e0);
					e1 = gf64_mulx4(e1);
				}
				uint4 v;
				v.x = (uint32_t)e0;
				v.y = (uint32_t)(e0 >> 32);
				v.z = (uint32_t)e1;
				v.w = (uint32_t)(e1 >> 32);
				t[(tbl * 16 + 8 * half + k) * 16 + n] = v;
			}
		}
	}
}

// This thread's words w0 + tid + 256*j (j < WPT) of the 16 inputs of block b,
// zeros where the input is virtual, past the group's blocks or flagged in
// `lost` (loaded like the others and zeroed: its memory may hold anything), or
// the word is past the slice end (clamped indices, selects: straight-line
// loads).  FAST: full tiles, no word checks.
template <int WPT, bool FAST>
__device__ __forceinline__ void d16_load(const DyadicArgs& a, uint64_t b, uint64_t b_end, uint64_t w0,
                                         uint32_t lost, uint2 (&d)[16][WPT]) {
	const uint64_t wl = w0 + threadIdx.x;
#pragma unroll
	for (int t = 0; t < 16; t++) {
		const uint64_t iv = 16 * b + t - a.pre_in;  // wraps for a virtual point in front
		const bool valid = (b < b_end) & (iv < a.N);
		const uint64_t i = valid ? iv : 0;
		const uint32_t on = (valid & !((lost >> t) & 1u)) ? ~0u : 0u;
		const uint64_t* src = a.in_base + i * a.in_stride;
#pragma unroll
		for (int j = 0; j < WPT; j++) {
			const uint64_t w = wl + 256 * j;
			uint64_t v;
			if (FAST) {
				v = dy_load(src + w);
			} else {
				const uint64_t x = dy_load(src + (w < a.W ? w : a.W - 1));
				v = w < a.W ? x : 0;
			}
			d[t][j] = make_uint2((uint32_t)v & on, (uint32_t)(v >> 32) & on);
		}
	}
}

// Issue step I of a word: nibble offsets when the step starts a new (T, h),
// then its 8 reads into buffer I & 1.  (Template steps: an 84-step loop is
// past what the compiler unrolls, and indexing the step tables at run time
// would put the accumulators in scratch.)
template <int I>
__device__ __forceinline__ void d16_issue(const uint4* tab, const uint2 (&x)[16], uint32_t (&ad)[2][8],
                                          uint4 (&v)[2][8]) {
	constexpr int T = kD16T[I], h = kD16H[I], sg = kD16SG[I], tb = kD16TB[I];
	constexpr int grp = (T * 2 + h) & 1;
	constexpr bool first = I == 0 || kD16T[I > 0 ? I - 1 : 0] != T || kD16H[I > 0 ? I - 1 : 0] != h;
	if constexpr (first) dy_nib16x8(h ? x[T].y : x[T].x, ad[grp]);
	const char* base = reinterpret_cast<const char*>(tab) + sg * kD16SetBytes;
	if constexpr (tb >= 0) {
#pragma unroll
		for (int kk = 0; kk < 8; kk++)
			v[I & 1][kk] = *reinterpret_cast<const uint4*>(base + (tb * 16 + 8 * h + kk) * 256 + ad[grp][kk]);
	} else {
#pragma unroll
		for (int kk = 0; kk < 8; kk++) {
			const uint2 r = *reinterpret_cast<const uint2*>(base + (4 * 16 + 8 * h + kk) * 256 + ad[grp][kk]);
			v[I & 1][kk].x = r.x;
			v[I & 1][kk].y = r.y;
		}
	}
}

// Fold step I's reads into its outputs h'_(S | T).
template <int I>
__device__ __forceinline__ void d16_fold(const uint4 (&v)[2][8], uint2 (&acc)[16]) {
	constexpr int T = kD16T[I], sg = kD16SG[I], tb = kD16TB[I];
	constexpr int ua = (8 * sg) | T | (tb >= 0 ? kDySA[tb >= 0 ? tb : 0] : 7);
	const uint4* r = v[I & 1];
	uint2& ca = acc[ua];
	ca.x = dy_xor3(dy_xor3(dy_xor3(dy_xor3(ca.x, r[0].x, r[1].x), r[2].x, r[3].x), r[4].x, r[5].x), r[6].x, r[7].x);
	ca.y = dy_xor3(dy_xor3(dy_xor3(dy_xor3(ca.y, r[0].y, r[1].y), r[2].y, r[3].y), r[4].y, r[5].y), r[6].y, r[7].y);
	if constexpr (tb >= 0) {
		uint2& cb = acc[(8 * sg) | T | kDySB[tb]];
		cb.x = dy_xor3(dy_xor3(dy_xor3(dy_xor3(cb.x, r[0].z, r[1].z), r[2].z, r[3].z), r[4].z, r[5].z), r[6].z, r[7].z);
		cb.y = dy_xor3(dy_xor3(dy_xor3(dy_xor3(cb.y, r[0].w, r[1].w), r[2].w, r[3].w), r[4].w, r[5].w), r[6].w, r[7].w);
	}
}

// Steps I.. of a word: step I + 1's reads issued before step I is folded.
template <int I>
__device__ __forceinline__ void d16_steps(const uint4* tab, const uint2 (&x)[16], uint32_t (&ad)[2][8],
                                          uint4 (&v)[2][8], uint2 (&acc)[16]) {
	if constexpr (I + 1 < kD16Steps) d16_issue<I + 1>(tab, x, ad, v);
	__builtin_amdgcn_sched_barrier(0);
	d16_fold<I>(v, acc);
	__builtin_amdgcn_sched_barrier(0);
	if constexpr (I + 1 < kD16Steps) d16_steps<I + 1>(tab, x, ad, v, acc);
}

// All products of one word's block (x: its 16 inputs, transformed here; acc:
// its 16 outputs in the subset basis), one table step of reads ahead.
__device__ __forceinline__ void d16_word(const uint4* tab, uint2 (&x)[16], uint2 (&acc)[16]) {
	d16_superset(x);
	uint32_t ad[2][8];
	uint4 v[2][8];
	d16_issue<0>(tab, x, ad, v);
	d16_steps<0>(tab, x, ad, v, acc);
}

// One tile: words [w0, w0 + 256*WPT) of output group q (virtual recovery
// points 16q..16q+15), blocks [b_begin, b_end); tab: 2 x 7 x 256 uint4 of LDS;
// coef: kD16Chunk x 16 words of LDS (GP false).  Output modes as dy_tile.
template <int WPT, bool FAST, bool GP, bool MASK = false>
__device__ __forceinline__ void d16_tile(const DyadicArgs& a, uint4* tab, uint64_t* coef, uint64_t q,
                                         uint64_t w0, uint64_t b_begin, uint64_t b_end, uint64_t g,
                                         int out_mode) {
	uint2 acc[WPT][16];
#pragma unroll
	for (int j = 0; j < WPT; j++)
#pragma unroll
		for (int u = 0; u < 16; u++) acc[j][u] = make_uint2(0, 0);
	uint2 d[16][WPT];
	uint64_t c[4];
	if (!GP) {
		d16_chunk_coeffs(a, coef, q, b_begin, b_end);
		__syncthreads();
	}
	d16_load_coeffs<GP>(a, coef, q, b_begin, b_begin, b_end, c);
	d16_load<WPT, FAST>(a, b_begin, b_end, w0, dy_lost<MASK>(a, b_begin, b_end), d);
	for (uint64_t bb = b_begin; bb < b_end; bb++) {
		__syncthreads();
		d16_build(tab, c);
		__syncthreads();
		if (!GP && bb + 1 < b_end && (bb + 1 - b_begin) % kD16Chunk == 0) {
			// every read of the previous chunk is done (the build above used
			// its last values)
			d16_chunk_coeffs(a, coef, q, bb + 1, b_end);
			__syncthreads();
		}
		d16_load_coeffs<GP>(a, coef, q, bb + 1, b_begin, b_end, c);
		const uint32_t lost_next = dy_lost<MASK>(a, bb + 1, b_end);
#pragma unroll
		for (int j = 0; j < WPT; j++) {
			uint2 x[16];
#pragma unroll
			for (int t = 0; t < 16; t++) x[t] = d[t][j];
			d16_word(tab, x, acc[j]);
			// the next block's words j (the registers are free now)
			uint2 dn[16][1];
			d16_load<1, FAST>(a, bb + 1, b_end, w0 + 256 * j, lost_next, dn);
#pragma unroll
			for (int t = 0; t < 16; t++) d[t][j] = dn[t][0];
		}
	}

	// back to the output basis, then store the real rows of this group
#pragma unroll
	for (int j = 0; j < WPT; j++) d16_superset(acc[j]);
#pragma unroll
	for (int s = 0; s < 16; s++) {
		const uint64_t v = 16 * q + s;  // virtual recovery slot
		if (v < a.pre_out) continue;
		const uint64_t r = v - a.pre_out;
		if (r >= a.R) break;
		uint64_t* dst;
		bool xr;
		if (out_mode == DY_OUT_PARTIAL) {
			dst = a.partial + (g * a.R + r) * a.W;
			xr = false;
		} else {
			dst = a.out_base + r * a.out_stride;
			xr = a.accumulate != 0;
		}
#pragma unroll
		for (int j = 0; j < WPT; j++) {
			const uint64_t w = w0 + threadIdx.x + 256 * j;
			const uint64_t o = ((uint64_t)acc[j][s].y << 32) | acc[j][s].x;
			if (w >= a.W) continue;
			if (out_mode == DY_OUT_ATOMIC) {
				if (o) atomicXor((unsigned long long*)(dst + w), (unsigned long long)o);
			} else {
				dst[w] = xr ? (dst[w] ^ o) : o;
			}
		}
	}
}

}  // namespace parhip
