// gf64_dyadic32_dev.h -- the XOR-block Cauchy contraction over 32 x 32 blocks
// (device side), for 17 or more outputs.
//
// Runs of 32 points, X = 32b + t and Y = 32q + s with t, s in 0..31, make each
// 32 x 32 block a product in GF(2^64)[Z_2^5], whose subset convolution needs
// 3^5 = 243 field products per block: 0.24 per input and output, against
// 81/256 = 0.32 for 16 x 16 blocks and 27/64 = 0.42 for 8 x 8 blocks (a quarter
// fewer table reads than 16 x 16 blocks at R = 32 and 64).
//
// Tables: the 32 transformed coefficients g'_U of a block split by their bits
// 3-4 into four sets of 8, each served by the seven pair tables of the 8 x 8
// form (112 KiB of LDS: one work-group of 512 threads per CU, one word per
// thread).  LDS rows are ordered (half, set, table, position): the 8 positions
// a 32-bit input half reads of every set lie within 56 KiB of the half's base,
// which the nibble offsets of the second half carry in their upper bytes (LDS
// read offsets are 16-bit).  Input a'_T needs the sets S_hi disjoint from its
// own bits 3-4 (T_hi = 0: four sets, 1 or 2: two, 3: one) and, inside a set,
// the pair tables of its low bits as in the 8 x 8 form; the nibble offsets of
// an input half are shared by all its reads.  A word of a block costs 252 table steps
// of 8 reads (234 x 8 ds_read_b128 + 18 x 8 ds_read_b64).  Product
// g'_S * a'_T goes to h'_(S | T); outputs stay in the subset basis over all
// blocks and are transformed back once before the store.
//
// The inputs of the next block are loaded into an input's registers as soon
// as its first table step has taken its nibble offsets, so the loads of a
// block overlap the rest of its lookups.
//
// Points that are not multiples of 32 get virtual points in front (zero
// inputs, outputs not stored), as in the other forms.  DyadicArgs: nblocks,
// blocks_per_group, pre_in, pre_out count in runs of 32; gp holds 32 words per
// (output group q, block b) at gp + (q*gp_ld + b)*32; lost masks (MASK
// launches) one u32 per block of 32 points, bit t = virtual input 32b + t
// counts as zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf64_dyadic_dev.h"

namespace parhip {

constexpr int kD32Threads = 512;
constexpr int kD32Steps = 252;
// LDS bytes of one half of the tables (4 sets x 7 pair tables x 8 positions x
// 16 entries x 16 B)
constexpr int kD32HalfBytes = 4 * kDyTables * 8 * 256;
// blocks per in-kernel coefficient chunk (32 coefficients each: 512 threads)
constexpr int kD32Chunk = 16;
// table build tasks (set, pair table, half, entry): 4 x 7 x 2 x 16
constexpr int kD32Tasks = 4 * kDyTables * 32;

// The table steps of one word, in input order: (input subset T, half h, table
// set sg, pair table tb; -1 = 8-byte reads of slot a of table 4).  Steps of one
// (T, h) share the nibble offsets.
struct D32Steps {
	int T[kD32Steps], h[kD32Steps], sg[kD32Steps], tb[kD32Steps];
};
constexpr D32Steps d32_make_steps() {
	// pair tables of an input's low three bits (the 8 x 8 form's kDyTab)
	const int nt[8] = {4, 2, 2, 1, 2, 1, 1, 1};
	const int tabs[8][4] = {{0, 1, 2, 3}, {4, 5, 0, 0}, {0, 2, 0, 0}, {6, 0, 0, 0},
	                        {0, 1, 0, 0}, {4, 0, 0, 0}, {0, 0, 0, 0}, {-1, 0, 0, 0}};
	D32Steps s{};
	int i = 0;
	for (int T = 31; T >= 0; T--)
		for (int h = 0; h < 2; h++)
			for (int k = 0; k < nt[T & 7]; k++)
				for (int sg = 0; sg < 4; sg++)
					if (!(sg & (T >> 3))) {
						s.T[i] = T;
						s.h[i] = h;
						s.sg[i] = sg;
						s.tb[i] = tabs[T & 7][k];
						i++;
					}
	return s;
}
__device__ constexpr D32Steps kD32 = d32_make_steps();

// 5-bit superset transform of 32 words, in place (its own inverse).
__device__ __forceinline__ void d32_superset(uint2 (&x)[32]) {
#pragma unroll
	for (int bit = 1; bit < 32; bit <<= 1)
#pragma unroll
		for (int s = 0; s < 32; s++)
			if (!(s & bit)) {
				x[s].x ^= x[s | bit].x;
				x[s].y ^= x[s | bit].y;
			}
}

// The transformed coefficients of blocks [b0, b0 + 16) of output group q into
// coef[p*32 + U] (thread 32p + u: block b0 + p, point offset u; one inversion
// each, the 5-bit superset sums across the 32 lanes of a block).
__device__ __forceinline__ void d32_chunk_coeffs(const DyadicArgs& a, uint64_t* coef, uint64_t q,
                                                 uint64_t b0, uint64_t b_end) {
	const int tid = threadIdx.x;
	const int u = tid & 31;
	const uint64_t b = b0 + (uint64_t)(tid >> 5);
	uint64_t c = 0;
	if (b < b_end) {
		const uint64_t z = ((a.xbase + 32 * b) ^ (a.ybase + 32 * q)) ^ (uint64_t)u;
		c = z ? gf64_inv(z) : 1u;  // the reference's 0 -> 1 (src/par3_engine.cc:392-394)
	}
#pragma unroll
	for (int bit = 1; bit < 32; bit <<= 1) {
		const uint32_t lo = __shfl_xor((uint32_t)c, bit, 64);
		const uint32_t hi = __shfl_xor((uint32_t)(c >> 32), bit, 64);
		if (!(u & bit)) c ^= ((uint64_t)hi << 32) | lo;
	}
	coef[tid] = c;
}

// The coefficients of this thread's build tasks k = tid and tid + 512 (task
// (sg, tbl, half, n): slots a and b of pair table tbl of set sg); zeros past
// the group's blocks and for the tasks past the last.
template <bool GP>
__device__ __forceinline__ void d32_load_coeffs(const DyadicArgs& a, const uint64_t* coef, uint64_t q,
                                                uint64_t b, uint64_t b_begin, uint64_t b_end,
                                                uint64_t (&c)[4]) {
	c[0] = c[1] = c[2] = c[3] = 0;
	if (b >= b_end) return;
	const uint64_t* p = GP ? a.gp + (q * a.gp_ld + b) * 32 : coef + ((b - b_begin) % kD32Chunk) * 32;
#pragma unroll
	for (int r = 0; r < 2; r++) {
		const int k = threadIdx.x + kD32Threads * r;
		if (k < kD32Tasks) {
			const int sg = k / (kDyTables * 32);
			const int tbl = (k % (kDyTables * 32)) >> 5;
			const int sa = (0x00406420u >> (4 * tbl)) & 7u, sb = (0x04627531u >> (4 * tbl)) & 7u;
			if (GP) {
				c[2 * r] = dy_load(p + 8 * sg + sa);
				c[2 * r + 1] = dy_load(p + 8 * sg + sb);
			} else {
				c[2 * r] = p[8 * sg + sa];
				c[2 * r + 1] = p[8 * sg + sb];
			}
		}
	}
}

// Table build of one block: task k = (sg, tbl, half, n) writes entry n of 8 of
// the 16 positions of pair table tbl of set sg.
__device__ __forceinline__ void d32_build(uint4* tab, const uint64_t (&c)[4]) {
#pragma unroll
	for (int r = 0; r < 2; r++) {
		const int k = threadIdx.x + kD32Threads * r;
		if (k < kD32Tasks) {
			const int sg = k / (kDyTables * 32);
			const int rem = k % (kDyTables * 32);
			const int tbl = rem >> 5, half = (rem >> 4) & 1, n = rem & 15;
			uint64_t e0 = gf64_mul_nibble(c[2 * r], (unsigned)n);
			uint64_t e1 = gf64_mul_nibble(c[2 * r + 1], (unsigned)n);
			if (half) {
				e0 = gf64_mulx32(e0);
				e1 = gf64_mulx32(e1);
			}
			uint4* t = tab + half * (kD32HalfBytes / 16) + (sg * kDyTables + tbl) * 8 * 16;
#pragma unroll
			for (int kk = 0; kk < 8; kk++) {
				if (kk) {
					e0 = gf64_mulx4(e0);
					e1 = gf64_mulx4(e1);
				}
				uint4 v;
				v.x = (uint32_t)e0;
				v.y = (uint32_t)(e0 >> 32);
				v.z = (uint32_t)e1;
				v.w = (uint32_t)(e1 >> 32);
				t[kk * 16 + n] = v;
			}
		}
	}
}

// Input t of block b at word w: zero where the input is virtual, past the
// group's blocks or flagged in `lost` (loaded like the others and zeroed), or
// the word is past the slice end (clamped index and a select: straight-line
// loads).  FAST: full tiles, no word check.
template <bool FAST>
__device__ __forceinline__ uint2 d32_load(const DyadicArgs& a, uint64_t b, uint64_t b_end, uint64_t w,
                                          uint32_t lost, int t) {
	const uint64_t iv = 32 * b + t - a.pre_in;  // wraps for a virtual point in front
	const bool valid = (b < b_end) & (iv < a.N);
	const uint64_t i = valid ? iv : 0;
	const uint32_t on = (valid & !((lost >> t) & 1u)) ? ~0u : 0u;
	const uint64_t* src = a.in_base + i * a.in_stride;
	uint64_t v;
	if (FAST) {
		v = dy_load(src + w);
	} else {
		const uint64_t x = dy_load(src + (w < a.W ? w : a.W - 1));
		v = w < a.W ? x : 0;
	}
	return make_uint2((uint32_t)v & on, (uint32_t)(v >> 32) & on);
}

// Byte offsets 16 * n_k of the 8 nibbles of a 32-bit word written into byte 0
// of each ad[k], whose upper bytes are kept (SDWA dst_unused:UNUSED_PRESERVE):
// ad[k] = base | 16 * n_k for a base that is a multiple of 256.
__device__ __forceinline__ void d32_nib16x8(uint32_t x, uint32_t (&ad)[8]) {
#define D32_HI(M, R) asm("v_and_b32_sdwa %0, %1, %2 dst_sel:BYTE_0 dst_unused:UNUSED_PRESERVE src0_sel:BYTE_" #M " src1_sel:DWORD" : "+v"(R) : "v"(x), "s"(0xF0u))
#define D32_LO(M, R) asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:BYTE_0 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:BYTE_" #M : "+v"(R) : "v"(x))
	D32_LO(0, ad[0]);
	D32_HI(0, ad[1]);
	D32_LO(1, ad[2]);
	D32_HI(1, ad[3]);
	D32_LO(2, ad[4]);
	D32_HI(2, ad[5]);
	D32_LO(3, ad[6]);
	D32_HI(3, ad[7]);
#undef D32_HI
#undef D32_LO
}

// State of one word's pass over a block: the nibble offsets of the current
// input's first and second half (the second carrying its half's LDS base in
// its upper bytes), and two steps of reads.
struct D32Pass {
	uint32_t ad[2][8];
	uint4 v[2][8];
};

// Issue step I: nibble offsets when it starts a new (T, h), then its 8 reads
// into buffer I & 1.
template <int I>
__device__ __forceinline__ void d32_issue(const uint4* tab, const uint2 (&x)[32], D32Pass& P) {
	constexpr int T = kD32.T[I], h = kD32.h[I], sg = kD32.sg[I], tb = kD32.tb[I];
	// the first step of input T takes the offsets of both halves
	constexpr bool first = I == 0 || kD32.T[I > 0 ? I - 1 : 0] != T;
	if constexpr (first) {
		d32_nib16x8(x[T].x, P.ad[0]);
		d32_nib16x8(x[T].y, P.ad[1]);
	}
	const char* base = reinterpret_cast<const char*>(tab) + (sg * kDyTables) * 8 * 256;
	if constexpr (tb >= 0) {
#pragma unroll
		for (int kk = 0; kk < 8; kk++)
			P.v[I & 1][kk] = *reinterpret_cast<const uint4*>(base + (tb * 8 + kk) * 256 + P.ad[h][kk]);
	} else {
#pragma unroll
		for (int kk = 0; kk < 8; kk++) {
			const uint2 r = *reinterpret_cast<const uint2*>(base + (4 * 8 + kk) * 256 + P.ad[h][kk]);
			P.v[I & 1][kk].x = r.x;
			P.v[I & 1][kk].y = r.y;
		}
	}
}

// Fold step I's reads into its outputs h'_(S | T).
template <int I>
__device__ __forceinline__ void d32_fold(const D32Pass& P, uint2 (&acc)[32]) {
	constexpr int T = kD32.T[I], sg = kD32.sg[I], tb = kD32.tb[I];
	constexpr int ua = (8 * sg) | T | (tb >= 0 ? kDySA[tb >= 0 ? tb : 0] : 0);
	const uint4* r = P.v[I & 1];
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
// Once the first step of input T has taken its offsets, x[T] is free and
// receives input T of the next block (b_next); the inputs run from T = 31
// (two steps) to T = 0 (32 steps), so the last load of a pass is issued 32
// steps before its end.
template <int I, bool FAST>
__device__ __forceinline__ void d32_steps(const DyadicArgs& a, const uint4* tab, uint2 (&x)[32], D32Pass& P,
                                          uint2 (&acc)[32], uint64_t b_next, uint64_t b_end, uint64_t w,
                                          uint32_t lost_next) {
	if constexpr (I + 1 < kD32Steps) {
		d32_issue<I + 1>(tab, x, P);
		constexpr int T1 = kD32.T[I + 1];
		constexpr bool frees = kD32.T[I] != T1;  // first step of input T1
		if constexpr (frees) x[T1] = d32_load<FAST>(a, b_next, b_end, w, lost_next, T1);
	}
	__builtin_amdgcn_sched_barrier(0);
	d32_fold<I>(P, acc);
	__builtin_amdgcn_sched_barrier(0);
	if constexpr (I + 1 < kD32Steps) d32_steps<I + 1, FAST>(a, tab, x, P, acc, b_next, b_end, w, lost_next);
}

// One tile: words [w0, w0 + 512) of output group q (virtual recovery points
// 32q..32q+31), blocks [b_begin, b_end); tab: 4 x 7 x 256 uint4 of LDS; coef:
// kD32Chunk x 32 words of LDS (GP false).  Output modes as dy_tile.
template <bool FAST, bool GP, bool MASK>
__device__ __forceinline__ void d32_tile(const DyadicArgs& a, uint4* tab, uint64_t* coef, uint64_t q,
                                         uint64_t w0, uint64_t b_begin, uint64_t b_end, uint64_t g,
                                         int out_mode) {
	const uint64_t w = w0 + threadIdx.x;
	uint2 acc[32];
#pragma unroll
	for (int u = 0; u < 32; u++) acc[u] = make_uint2(0, 0);
	uint2 x[32];
	uint64_t c[4];
	if (!GP) {
		d32_chunk_coeffs(a, coef, q, b_begin, b_end);
		__syncthreads();
	}
	d32_load_coeffs<GP>(a, coef, q, b_begin, b_begin, b_end, c);
	{
		const uint32_t lost0 = dy_lost<MASK>(a, b_begin, b_end);
#pragma unroll
		for (int t = 0; t < 32; t++) x[t] = d32_load<FAST>(a, b_begin, b_end, w, lost0, t);
	}
	for (uint64_t bb = b_begin; bb < b_end; bb++) {
		__syncthreads();
		d32_build(tab, c);
		__syncthreads();
		if (!GP && bb + 1 < b_end && (bb + 1 - b_begin) % kD32Chunk == 0) {
			// every read of the previous chunk is done (the build above used
			// its last values)
			d32_chunk_coeffs(a, coef, q, bb + 1, b_end);
			__syncthreads();
		}
		d32_load_coeffs<GP>(a, coef, q, bb + 1, b_begin, b_end, c);
		const uint32_t lost_next = dy_lost<MASK>(a, bb + 1, b_end);
		d32_superset(x);
		D32Pass P;
#pragma unroll
		for (int kk = 0; kk < 8; kk++) {
			P.ad[0][kk] = 0;
			P.ad[1][kk] = kD32HalfBytes;
		}
		d32_issue<0>(tab, x, P);
		x[kD32.T[0]] = d32_load<FAST>(a, bb + 1, b_end, w, lost_next, kD32.T[0]);
		d32_steps<0, FAST>(a, tab, x, P, acc, bb + 1, b_end, w, lost_next);
	}

	// back to the output basis, then store the real rows of this group
	d32_superset(acc);
	if (w >= a.W) return;
#pragma unroll
	for (int s = 0; s < 32; s++) {
		const uint64_t v = 32 * q + s;  // virtual recovery slot
		if (v < a.pre_out) continue;
		const uint64_t r = v - a.pre_out;
		if (r >= a.R) break;
		const uint64_t o = ((uint64_t)acc[s].y << 32) | acc[s].x;
		if (out_mode == DY_OUT_PARTIAL) {
			a.partial[(g * a.R + r) * a.W + w] = o;
		} else if (out_mode == DY_OUT_ATOMIC) {
			if (o) atomicXor((unsigned long long*)(a.out_base + r * a.out_stride + w), (unsigned long long)o);
		} else {
			uint64_t* dst = a.out_base + r * a.out_stride + w;
			*dst = a.accumulate ? (*dst ^ o) : o;
		}
	}
}

}  // namespace parhip
