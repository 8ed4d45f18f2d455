// gf64_dyadic_dev.h -- device side of the XOR-block (dyadic) Cauchy
// contraction, shared by the plain kernel (gf64_dyadic.hip) and the persistent
// batch kernel (gf64_kernels.hip).
//
// The plain create multiplies by the Cauchy matrix of src/par3_engine.cc:358-429,
//   c(r, i) = f((firstInput + i) XOR (firstRecovery + r)),   f(z) = z ? 1/z : 1.
// Group the input points X = firstInput + i and the recovery points
// Y = firstRecovery + r into aligned runs of 8: X = 8b + t, Y = 8q + s with
// t, s in 0..7.  Then X XOR Y = (8b XOR 8q) XOR (t XOR s), so inside every
// 8 x 8 block (b, q) the coefficient depends on t XOR s only:
//   c = g_bq(t XOR s).
// Such a block is multiplication in the group algebra GF(2^64)[Z_2^3]
// (y_s = XOR_t g(t XOR s) d_t is the group convolution g * d).  In
// characteristic 2 that algebra is GF(2^64)[u1,u2,u3]/(u_k^2) with
// u_k = 1 + t_k, and in the u basis
//   a'_T = XOR_{S superset of T} a_S          (its own inverse: the same sums)
//   y'_U = XOR_{S, T disjoint, S | T = U} g'_S * d'_T     (subset convolution)
// which needs 3^3 = 27 field products per block instead of 8 x 8 = 64.  The
// transforms are XORs of words; the result is the same linear map, so the
// bytes are exactly those of the full product.  Outputs stay in the u basis
// over all input blocks (the back transform is linear and the same for every
// block) and are transformed once before the store.
//
// Products still go through nibble tables in LDS, two coefficients per 16-byte
// entry (one ds_read_b128 gives two products of one input nibble, and a
// 16-entry table is one 256-byte row: conflict-free).  Input d'_T needs the
// coefficients g'_S with S disjoint from T; they are served by seven pair
// tables per block:
//   P0 (0,1) P2 (2,3) P4 (4,5) P6 (6,7) Q0 (0,2) Q4 (4,6) X (0,4)
//   T = 0: P0 P2 P4 P6   T = 1: Q0 Q4   T = 2: P0 P4   T = 3: X
//   T = 4: P0 P2         T = 5: Q0      T = 6: P0      T = 7: first half of Q0
// so a word of a block costs 13 x 16 ds_read_b128 + 16 ds_read_b64 (3.5 KiB of
// LDS reads) against 32 x 16 ds_read_b128 (8 KiB) for the direct form.
//
// Points that are not multiples of 8 are handled with virtual points: the
// first block (and first output group) is padded in front with inputs that
// hold zeros (rows that are not stored), so every plain Cauchy call qualifies.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf64_dev.h"
#include "kernels.h"

namespace parhip {


#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const uint64_t dy_gu64;
typedef __attribute__((address_space(1))) const ulonglong2 dy_gu128;
__device__ __forceinline__ uint64_t dy_load(const uint64_t* p) { return *(dy_gu64*)p; }
__device__ __forceinline__ ulonglong2 dy_load2(const uint64_t* p) {
	dy_gu128* q = (dy_gu128*)p;
	ulonglong2 v;
	v.x = q->x;
	v.y = q->y;
	return v;
}
#else
__device__ __forceinline__ uint64_t dy_load(const uint64_t* p) { return *p; }
__device__ __forceinline__ ulonglong2 dy_load2(const uint64_t* p) {
	return *reinterpret_cast<const ulonglong2*>(p);
}
#endif

__device__ __forceinline__ uint32_t dy_xor3(uint32_t a, uint32_t b, uint32_t c) {
	return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// The seven pair tables: coefficient subsets (slot a, slot b).
__device__ constexpr int kDySA[7] = {0, 2, 4, 6, 0, 4, 0};
__device__ constexpr int kDySB[7] = {1, 3, 5, 7, 2, 6, 4};
// Tables read for input T (T = 7 reads slot a of table 4 only).
__device__ constexpr int kDyNT[8] = {4, 2, 2, 1, 2, 1, 1, 1};
__device__ constexpr int kDyTab[8][4] = {{0, 1, 2, 3}, {4, 5, 0, 0}, {0, 2, 0, 0}, {6, 0, 0, 0},
                                         {0, 1, 0, 0}, {4, 0, 0, 0}, {0, 0, 0, 0}, {4, 0, 0, 0}};
constexpr int kDyTables = 7;

// a'_T = XOR over supersets S of T of a_S (3-bit masks), in place; also the
// inverse transform.
__device__ __forceinline__ void dy_superset(uint2 (&x)[8]) {
#pragma unroll
	for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
		for (int s = 0; s < 8; s++)
			if (!(s & bit)) {
				x[s].x ^= x[s | bit].x;
				x[s].y ^= x[s | bit].y;
			}
}

// Byte offsets 16 * n_k of the 8 nibbles n_k of a 32-bit word (table rows of
// 16-byte entries): SDWA byte selects, one VALU op per offset.
//   odd nibble 2m+1:  byte_m & 0xF0
//   even nibble 2m:   (byte_m << 4) & 0xFF   (dst_sel BYTE_0, upper bits zeroed)
__device__ __forceinline__ void dy_nib16x8(uint32_t x, uint32_t (&ad)[8]) {
#define DY_HI(M, R) asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_" #M " src1_sel:DWORD" : "=v"(R) : "v"(x), "s"(0xF0u))
#define DY_LO(M, R) asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:BYTE_0 dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_" #M : "=v"(R) : "v"(x))
	DY_LO(0, ad[0]);
	DY_HI(0, ad[1]);
	DY_LO(1, ad[2]);
	DY_HI(1, ad[3]);
	DY_LO(2, ad[4]);
	DY_HI(2, ad[5]);
	DY_LO(3, ad[6]);
	DY_HI(3, ad[7]);
#undef DY_HI
#undef DY_LO
}

// c * x^32: the 32 bits shifted out fold back times x^4 + x^3 + x + 1 (at
// most 36 bits, no second fold).
__device__ __forceinline__ uint64_t gf64_mulx32(uint64_t c) {
	const uint64_t h = c >> 32;
	return (c << 32) ^ h ^ (h << 1) ^ (h << 3) ^ (h << 4);
}

// Coefficients computed in the kernel (a.gp null) come in chunks of
// kDyChunk blocks, kept in LDS.
constexpr int kDyChunk = 32;

// The transformed coefficients of blocks [b0, b0 + kDyChunk) of output group q
// into coef (thread 8p + u: block b0 + p, slot u; one inversion each, the
// superset sums across the 8 lanes of a block).  Every thread takes part.
__device__ __forceinline__ void dy_chunk_coeffs(const DyadicArgs& a, uint64_t* coef, uint64_t q,
                                                uint64_t b0, uint64_t b_end) {
	const int tid = threadIdx.x;
	if (tid < kDyChunk * 8) {
		const int u = tid & 7;
		const uint64_t b = b0 + (uint64_t)(tid >> 3);
		uint64_t c = 0;
		if (b < b_end) {
			const uint64_t z = ((a.xbase + 8 * b) ^ (a.ybase + 8 * q)) ^ (uint64_t)u;
			c = z ? gf64_inv(z) : 1u;  // the reference's 0 -> 1 (src/par3_engine.cc:392-394)
		}
#pragma unroll
		for (int bit = 1; bit < 8; bit <<= 1) {
			const uint32_t lo = __shfl_xor((uint32_t)c, bit, 64);
			const uint32_t hi = __shfl_xor((uint32_t)(c >> 32), bit, 64);
			if (!(u & bit)) c ^= ((uint64_t)hi << 32) | lo;
		}
		coef[tid] = c;
	}
}

// The two coefficients of the table entries this thread builds for block b
// (thread (tbl, half, n) with tbl < 7), 0 past the group's blocks: from the
// coefficient array, or from the chunk in LDS (b_begin: the tile's first block).
template <bool GP>
__device__ __forceinline__ void dy_load_coeffs(const DyadicArgs& a, const uint64_t* coef, uint64_t q,
                                               uint64_t b, uint64_t b_begin, uint64_t b_end,
                                               uint64_t& c0, uint64_t& c1) {
	const int tid = threadIdx.x;
	c0 = 0;
	c1 = 0;
	if (tid < kDyTables * 32 && b < b_end) {
		// kDySA[tbl], kDySB[tbl] from packed nibbles (an array indexed by a
		// lane value would be a global load, whose wait is also a wait for
		// every input load issued before it)
		const int tbl = tid >> 5;
		const int sa = (0x00406420u >> (4 * tbl)) & 7u, sb = (0x04627531u >> (4 * tbl)) & 7u;
		if (GP) {
			const uint64_t* gp = a.gp + (q * a.gp_ld + b) * 8;
			c0 = dy_load(gp + sa);
			c1 = dy_load(gp + sb);
		} else {
			const uint64_t* cp = coef + ((b - b_begin) % kDyChunk) * 8;
			c0 = cp[sa];
			c1 = cp[sb];
		}
	}
}

// The block's mask of inputs that count as zero (MASK launches only; the
// block index is the same for the whole work-group, so this is one scalar).
template <bool MASK>
__device__ __forceinline__ uint32_t dy_lost(const DyadicArgs& a, uint64_t b, uint64_t b_end) {
	if (!MASK) return 0u;
	// a scalar load (constant address space): the mask is read-only here
	typedef const __attribute__((address_space(4))) uint32_t cu32;
	return ((cu32*)(a.lost))[b < b_end ? b : 0];  // (past the end: unused)
}

// Inputs of block b that count: bit t set when virtual slot t is a real input
// (not a virtual point in front, not past the group's blocks or past N) and is
// not flagged in `lost` (a lost slice is loaded like the others and zeroed: its
// memory may hold anything).  One scalar per block.
__device__ __forceinline__ uint32_t dy_live(const DyadicArgs& a, uint64_t b, uint64_t b_end, uint32_t lost) {
	uint32_t m = 0;
#pragma unroll
	for (int t = 0; t < 8; t++) {
		// wraps to a huge index for a virtual point in front
		const uint64_t i = 8 * b + t - a.pre_in;
		m |= (uint32_t)(i < a.N) << t;
	}
	return (b < b_end ? m : 0u) & ~lost;
}

// This thread's word pair JJ (words w0 + 2*tid + 512*JJ, +1) of the 8 inputs of
// block b, zeroed where the input does not count (live, from dy_live) or the
// word is past the slice end.  Indices are clamped into range and the zeros
// applied with selects, so the loads are straight-line code (a per-input
// condition on the load itself makes the compiler spill the whole pipeline).
// FAST (full, 16-byte aligned tiles) loads each pair with one 16-byte load.
template <int WPT, int JJ, bool FAST>
__device__ __forceinline__ void dy_load_pair(const DyadicArgs& a, uint64_t b, uint64_t b_end,
                                             uint64_t w0, uint32_t live, uint2 (&d)[8][WPT]) {
	const int tid = threadIdx.x;
	const uint64_t w = w0 + 2 * tid + 512 * JJ;
#pragma unroll
	for (int t = 0; t < 8; t++) {
		// real input of virtual slot t of block b (the index wraps for a
		// virtual point in front); no short-circuit, so no branch
		const uint64_t iv = 8 * b + t - a.pre_in;
		const uint64_t i = ((b < b_end) & (iv < a.N)) ? iv : 0;
		const uint32_t on = 0u - ((live >> t) & 1u);
		uint64_t v0, v1;
		if (FAST) {
			const ulonglong2 x = dy_load2(a.in_base + i * a.in_stride + w);
			v0 = x.x;
			v1 = x.y;
		} else {
			// ragged or unaligned tiles: two 8-byte loads at clamped word indices
			const uint64_t* src = a.in_base + i * a.in_stride;
			const uint64_t wa = w < a.W ? w : a.W - 1, wb = w + 1 < a.W ? w + 1 : a.W - 1;
			const uint64_t x0 = dy_load(src + wa), x1 = dy_load(src + wb);
			v0 = w < a.W ? x0 : 0;
			v1 = w + 1 < a.W ? x1 : 0;
		}
		d[t][2 * JJ] = make_uint2((uint32_t)v0 & on, (uint32_t)(v0 >> 32) & on);
		d[t][2 * JJ + 1] = make_uint2((uint32_t)v1 & on, (uint32_t)(v1 >> 32) & on);
	}
}

// All products of one word's block: x = the block's 8 input words, transformed
// to the subset basis here; acc = the 8 outputs in the subset basis.
__device__ __forceinline__ void dy_word(const uint4* tab, uint2 (&x)[8], uint2 (&acc)[8]) {
	dy_superset(x);
#pragma unroll
	for (int T = 0; T < 8; T++) {
#pragma unroll
		for (int h = 0; h < 2; h++) {
			const uint32_t src = h ? x[T].y : x[T].x;
			uint32_t ad[8];
			dy_nib16x8(src, ad);
			if (T == 7) {
				// slot a of table Q0 (g'_0): 8-byte reads
				uint2 v[8];
#pragma unroll
				for (int kk = 0; kk < 8; kk++)
					v[kk] = *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(tab) +
					                                        (4 * 16 + 8 * h + kk) * 256 + ad[kk]);
				uint2 c = acc[7];
				c.x = dy_xor3(dy_xor3(dy_xor3(dy_xor3(c.x, v[0].x, v[1].x), v[2].x, v[3].x), v[4].x, v[5].x), v[6].x, v[7].x);
				c.y = dy_xor3(dy_xor3(dy_xor3(dy_xor3(c.y, v[0].y, v[1].y), v[2].y, v[3].y), v[4].y, v[5].y), v[6].y, v[7].y);
				acc[7] = c;
				__builtin_amdgcn_sched_barrier(0);
			} else {
#pragma unroll
				for (int uu = 0; uu < 4; uu++) {
					if (uu >= kDyNT[T]) break;
					const int tbl = kDyTab[T][uu];
					const int ua = kDySA[tbl] | T, ub = kDySB[tbl] | T;
					uint4 v[8];
#pragma unroll
					for (int kk = 0; kk < 8; kk++)
						v[kk] = *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(tab) +
						                                        (tbl * 16 + 8 * h + kk) * 256 + ad[kk]);
					uint2 ca = acc[ua], cb = acc[ub];
					ca.x = dy_xor3(dy_xor3(dy_xor3(dy_xor3(ca.x, v[0].x, v[1].x), v[2].x, v[3].x), v[4].x, v[5].x), v[6].x, v[7].x);
					ca.y = dy_xor3(dy_xor3(dy_xor3(dy_xor3(ca.y, v[0].y, v[1].y), v[2].y, v[3].y), v[4].y, v[5].y), v[6].y, v[7].y);
					cb.x = dy_xor3(dy_xor3(dy_xor3(dy_xor3(cb.x, v[0].z, v[1].z), v[2].z, v[3].z), v[4].z, v[5].z), v[6].z, v[7].z);
					cb.y = dy_xor3(dy_xor3(dy_xor3(dy_xor3(cb.y, v[0].w, v[1].w), v[2].w, v[3].w), v[4].w, v[5].w), v[6].w, v[7].w);
					acc[ua] = ca;
					acc[ub] = cb;
					// one table's 8 reads in flight at a time: deeper hoisting
					// runs out of registers at two waves per SIMD
					__builtin_amdgcn_sched_barrier(0);
				}
			}
		}
	}
}

// The table steps of one word, in order: (T, half h, pair table; -1 = T = 7's
// 8-byte reads of slot a of table 4).  A step is 8 reads of one table at the
// nibble offsets of x'_T's half h.
__device__ constexpr int kDyStepT[28] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2,
                                         2, 2, 3, 3, 4, 4, 4, 4, 5, 5, 6, 6, 7, 7};
__device__ constexpr int kDyStepH[28] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 0,
                                         1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 0, 1, 0, 1};
__device__ constexpr int kDyStepTab[28] = {0, 1, 2, 3, 0, 1, 2, 3, 4, 5, 4, 5, 0, 2,
                                           0, 2, 6, 6, 0, 1, 0, 1, 4, 4, 0, 0, -1, -1};
constexpr int kDySteps = 28;

// Issue step i of the pair (word i / 28): nibble offsets when the step starts
// a new (word, T, h) group, then its 8 reads into buffer i & 1.
__device__ __forceinline__ void dy_issue(const uint4* tab, int i, const uint2 (&x)[2][8],
                                         uint32_t (&ad)[2][8], uint4 (&v)[2][8]) {
	const int wd = i / kDySteps, s = i % kDySteps;
	const int T = kDyStepT[s], h = kDyStepH[s], tb = kDyStepTab[s];
	const int grp = (wd * 16 + T * 2 + h) & 1;
	const bool first = s == 0 || kDyStepT[s - 1] != T || kDyStepH[s - 1] != h;
	if (first) dy_nib16x8(h ? x[wd][T].y : x[wd][T].x, ad[grp]);
	const char* base = reinterpret_cast<const char*>(tab);
	if (tb >= 0) {
#pragma unroll
		for (int kk = 0; kk < 8; kk++)
			v[i & 1][kk] = *reinterpret_cast<const uint4*>(base + (tb * 16 + 8 * h + kk) * 256 + ad[grp][kk]);
	} else {
#pragma unroll
		for (int kk = 0; kk < 8; kk++) {
			const uint2 r = *reinterpret_cast<const uint2*>(base + (4 * 16 + 8 * h + kk) * 256 + ad[grp][kk]);
			v[i & 1][kk].x = r.x;
			v[i & 1][kk].y = r.y;
		}
	}
}

// Fold step i's reads into its accumulators (two outputs; T = 7: one).
template <int WPT>
__device__ __forceinline__ void dy_fold(int i, int j0, const uint4 (&v)[2][8], uint2 (&acc)[WPT][8]) {
	const int wd = i / kDySteps, s = i % kDySteps;
	const int T = kDyStepT[s], tb = kDyStepTab[s];
	const uint4* r = v[i & 1];
	uint2& ca = acc[j0 + wd][tb >= 0 ? (kDySA[tb] | T) : 7];
	ca.x = dy_xor3(dy_xor3(dy_xor3(dy_xor3(ca.x, r[0].x, r[1].x), r[2].x, r[3].x), r[4].x, r[5].x), r[6].x, r[7].x);
	ca.y = dy_xor3(dy_xor3(dy_xor3(dy_xor3(ca.y, r[0].y, r[1].y), r[2].y, r[3].y), r[4].y, r[5].y), r[6].y, r[7].y);
	if (tb >= 0) {
		uint2& cb = acc[j0 + wd][kDySB[tb] | T];
		cb.x = dy_xor3(dy_xor3(dy_xor3(dy_xor3(cb.x, r[0].z, r[1].z), r[2].z, r[3].z), r[4].z, r[5].z), r[6].z, r[7].z);
		cb.y = dy_xor3(dy_xor3(dy_xor3(dy_xor3(cb.y, r[0].w, r[1].w), r[2].w, r[3].w), r[4].w, r[5].w), r[6].w, r[7].w);
	}
}

// All products of a word pair's block (x: the pair's 8 input words each,
// transformed here; acc[j0], acc[j0 + 1]: their outputs in the subset basis),
// software-pipelined one table ahead: step i + 1's reads are issued before
// step i's are folded, so each wave keeps two tables' reads in flight.
template <int WPT>
__device__ __forceinline__ void dy_pair_words(const uint4* tab, uint2 (&x)[2][8], int j0,
                                              uint2 (&acc)[WPT][8]) {
	dy_superset(x[0]);
	dy_superset(x[1]);
	uint32_t ad[2][8];
	uint4 v[2][8];
	dy_issue(tab, 0, x, ad, v);
#pragma unroll
	for (int i = 0; i < 2 * kDySteps; i++) {
		if (i + 1 < 2 * kDySteps) dy_issue(tab, i + 1, x, ad, v);
		__builtin_amdgcn_sched_barrier(0);
		dy_fold<WPT>(i, j0, v, acc);
		__builtin_amdgcn_sched_barrier(0);
	}
}

// Block bb's products of word pair JJ, then the pair's inputs of block bb + 1
// into the same registers.  Every index is a compile-time constant, so d and
// acc stay in VGPRs.
template <int WPT, int JJ, bool FAST, bool PIPE>
__device__ __forceinline__ void dy_pair_step(const DyadicArgs& a, const uint4* tab, uint64_t bb,
                                             uint64_t b_end, uint64_t w0, uint32_t live_next,
                                             uint2 (&d)[8][WPT], uint2 (&acc)[WPT][8]) {
	uint2 x[2][8];
#pragma unroll
	for (int t = 0; t < 8; t++) {
		x[0][t] = d[t][2 * JJ];
		x[1][t] = d[t][2 * JJ + 1];
	}
	if (PIPE && WPT == 4) {
		// two tables in flight (64 VGPRs of reads): fits the 256 VGPRs of two
		// waves per SIMD, not the 128 of WPT = 2's four
		dy_pair_words<WPT>(tab, x, 2 * JJ, acc);
	} else {
		dy_word(tab, x[0], acc[2 * JJ]);
		dy_word(tab, x[1], acc[2 * JJ + 1]);
	}
	dy_load_pair<WPT, JJ, FAST>(a, bb + 1, b_end, w0, live_next, d);
}

// Table build of one block from the coefficients c0, c1 of this thread's pair
// table: thread (tbl, half, n) owns entry n of 8 of the 16 positions of one
// pair table; entry(k, n) = x^4 * entry(k-1, n), the second half starting from
// x^32 * entry(0, n).  The 16 lanes of a (table, half) write 16 consecutive
// entries of one row per store.
__device__ __forceinline__ void dy_build(uint4* tab, uint64_t c0, uint64_t c1) {
	const int tid = threadIdx.x;
	if (tid < kDyTables * 32) {
		// thread (tbl, half, n): entry n of positions 8*half .. 8*half+7
		const int n = tid & 15;
		const int half = (tid >> 4) & 1;
		const int tbl = tid >> 5;
		uint64_t e0 = gf64_mul_nibble(c0, (unsigned)n);
		uint64_t e1 = gf64_mul_nibble(c1, (unsigned)n);
		if (half) {
			e0 = gf64_mulx32(e0);
			e1 = gf64_mulx32(e1);
		}
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
			tab[(tbl * 16 + 8 * half + k) * 16 + n] = v;
		}
	}
}

// One tile: words [w0, w0 + 256*WPT) of output group q (virtual recovery
// points 8q..8q+7), input blocks [b_begin, b_end); `tab` is 7 x 256 uint4 of
// LDS.  out_mode: DY_OUT_DIRECT stores (or XORs when a.accumulate) into the
// outputs, DY_OUT_PARTIAL into partial row g, DY_OUT_ATOMIC XORs atomically
// (a job's input range split over tiles of one persistent launch; XOR is
// order-independent, so the bytes equal the unsplit result).
enum { DY_OUT_DIRECT = 0, DY_OUT_PARTIAL = 1, DY_OUT_ATOMIC = 2 };

//
// GP: coefficients from a.gp; otherwise computed here in chunks into `coef`
// (a template parameter, so the in-kernel form has no global coefficient loads
// whose waits would also wait for the input loads issued before them).
// PIPE (WPT = 4): table reads pipelined one table ahead (dy_pair_words).
// Measured on 1024 x 1 MiB: 4 % fewer cycles at R = 8; at R >= 16 the cycle
// count is the same and the clock lower (the chip's power limit), so the plan
// takes it for one output group only.
// (Two table buffers, block bb + 1 built beside block bb's reads with one
// barrier per block, measured 3-5 % slower at 1024 x 1 MiB, R = 8 and 24.)
template <int WPT, bool FAST, bool MASK, bool GP, bool PIPE = false>
__device__ __forceinline__ void dy_tile(const DyadicArgs& a, uint4* tab, uint64_t* coef, uint64_t q,
                                        uint64_t w0, uint64_t b_begin, uint64_t b_end, uint64_t g,
                                        int out_mode) {
	static_assert(WPT == 2 || WPT == 4, "WPT");
	const int tid = threadIdx.x;
	uint2 acc[WPT][8];
#pragma unroll
	for (int j = 0; j < WPT; j++)
#pragma unroll
		for (int u = 0; u < 8; u++) acc[j][u] = make_uint2(0, 0);

	// Software pipeline without a second buffer: a word pair's inputs of the
	// next block are loaded into its registers as soon as this block's
	// products of the pair are done, so the HBM latency hides behind the other
	// pairs' lookups and the next table build.
	uint2 d[8][WPT];
	uint64_t c0, c1;
	if (!GP) {
		dy_chunk_coeffs(a, coef, q, b_begin, b_end);
		__syncthreads();
	}
	dy_load_coeffs<GP>(a, coef, q, b_begin, b_begin, b_end, c0, c1);
	const uint32_t live0 = dy_live(a, b_begin, b_end, dy_lost<MASK>(a, b_begin, b_end));
	dy_load_pair<WPT, 0, FAST>(a, b_begin, b_end, w0, live0, d);
	if (WPT == 4) dy_load_pair<WPT, (WPT == 4 ? 1 : 0), FAST>(a, b_begin, b_end, w0, live0, d);
	for (uint64_t bb = b_begin; bb < b_end; bb++) {
		// ---- tables: thread (tbl, half, n) owns entry n of 8 of the 16
		// positions of one pair table; entry(k, n) = x^4 * entry(k-1, n), the
		// second half starting from x^32 * entry(0, n).  The 16 lanes of a
		// (table, half) write 16 consecutive entries of one row per store.
		__syncthreads();
		dy_build(tab, c0, c1);
		__syncthreads();
		if (!GP && bb + 1 < b_end && (bb + 1 - b_begin) % kDyChunk == 0) {
			// (every read of the previous chunk is done: its last values were
			// used by the build above)
			dy_chunk_coeffs(a, coef, q, bb + 1, b_end);
			__syncthreads();
		}
		dy_load_coeffs<GP>(a, coef, q, bb + 1, b_begin, b_end, c0, c1);
		const uint32_t live = dy_live(a, bb + 1, b_end, dy_lost<MASK>(a, bb + 1, b_end));
		dy_pair_step<WPT, 0, FAST, PIPE>(a, tab, bb, b_end, w0, live, d, acc);
		if (WPT == 4) dy_pair_step<WPT, (WPT == 4 ? 1 : 0), FAST, PIPE>(a, tab, bb, b_end, w0, live, d, acc);
	}

	// ---- back to the output basis, then store the real rows of this group
#pragma unroll
	for (int j = 0; j < WPT; j++) dy_superset(acc[j]);
#pragma unroll
	for (int s = 0; s < 8; s++) {
		const uint64_t v = 8 * q + s;  // virtual recovery slot
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
		for (int jj = 0; jj < WPT / 2; jj++) {
			const uint64_t w = w0 + 2 * tid + 512 * jj;
			const uint2 p0 = acc[2 * jj][s], p1 = acc[2 * jj + 1][s];
			uint64_t o0 = ((uint64_t)p0.y << 32) | p0.x;
			uint64_t o1 = ((uint64_t)p1.y << 32) | p1.x;
			if (out_mode == DY_OUT_ATOMIC) {
				if (w < a.W && o0) atomicXor((unsigned long long*)(dst + w), (unsigned long long)o0);
				if (w + 1 < a.W && o1)
					atomicXor((unsigned long long*)(dst + w + 1), (unsigned long long)o1);
			} else if (w + 1 < a.W && ((reinterpret_cast<uintptr_t>(dst + w) & 15u) == 0)) {
				ulonglong2* qd = reinterpret_cast<ulonglong2*>(dst + w);
				if (xr) {
					const ulonglong2 o = *qd;
					o0 ^= o.x;
					o1 ^= o.y;
				}
				ulonglong2 val;
				val.x = o0;
				val.y = o1;
				*qd = val;
			} else {
				if (w < a.W) dst[w] = xr ? (dst[w] ^ o0) : o0;
				if (w + 1 < a.W) dst[w + 1] = xr ? (dst[w + 1] ^ o1) : o1;
			}
		}
	}
}

}  // namespace parhip
