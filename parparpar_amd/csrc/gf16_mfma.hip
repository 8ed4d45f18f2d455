// gf16_mfma.hip -- PAR2 GF(2^16) contraction as a GF(2) bit-matrix product on MFMA.
//
// Multiplication by a coefficient c is a 16x16 matrix over GF(2): column j is
// c * x^j (mod 0x1100B).  For one 16-bit word w,
//     out_bits[(e, o)] = parity( sum_{(i, j)} M_{c(e,i)}[o][j] * bit_j(in_i[w]) ),
// an integer matrix product (rows = 16 R, K = 16 N, columns = words) whose
// parity is the GF(2) result -- the same scheme as gf64_mfma.hip with 16x16
// blocks.  k_gf16_rows prepares the matrix rows of every coefficient once,
// two inputs per 32-bit word: rows[(e*P + ip)*16 + o] = Row_o(c(e, 2ip)) |
// Row_o(c(e, 2ip+1)) << 16, where bit j of Row_o(c) is bit o of c * x^j.
// The main kernel stages them in LDS, expanded to 0/1 matrix elements:
//   - fp4:  v_mfma_scale_f32_32x32x64_f8f6f4, e2m1 elements 0 / 0.5 with block
//           scale 2 (products 0 or 1, the fp32 sum is an exact count);
//           K = 64 = 4 inputs: lane L = row + 32h holds input pair h, its 32
//           bits placed with the fp4 expansion of gf64_mfma.hip;
//   - int8: v_mfma_i32_32x32x32_i8, K = 32 = 2 inputs: lane h holds one input's
//           16 bits, one byte each.
// The words are expanded the same way, so A and B use one K order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "gf16_dev.h"
#include "kernels.h"

namespace parhip {

namespace {
typedef int g16_v4i __attribute__((ext_vector_type(4)));
typedef int g16_v8i __attribute__((ext_vector_type(8)));
typedef int g16_v16i __attribute__((ext_vector_type(16)));
typedef float g16_v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t g16_spread4_i8(uint32_t x) {  // 4 low bits -> 4 bytes of 0/1
	return ((x & 0xFu) * 0x00204081u) & 0x01010101u;
}

__device__ __forceinline__ uint4 g16_expand_i8(uint32_t x) {  // 16 bits -> 16 bytes
	return make_uint4(g16_spread4_i8(x), g16_spread4_i8(x >> 4), g16_spread4_i8(x >> 8),
	                  g16_spread4_i8(x >> 12));
}

__device__ __forceinline__ uint4 g16_expand_fp4(uint32_t x) {  // bit v+4q -> nibble q of dword v
	return make_uint4(x & 0x11111111u, (x >> 1) & 0x11111111u, (x >> 2) & 0x11111111u,
	                  (x >> 3) & 0x11111111u);
}
}  // namespace

// rows[(e*P + ip)*16 + o], P = ceil(N/2); inputs past N contribute zero rows.
__global__ void k_gf16_rows(const uint16_t* __restrict__ coeff, uint64_t ld, uint64_t N, uint64_t R,
                            uint32_t* __restrict__ rows) {
	const uint64_t P = (N + 1) / 2;
	const uint64_t total = R * P * 16;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t o = idx & 15, q = idx >> 4;
		const uint64_t e = q / P, ip = q - e * P;
		uint32_t word = 0;
#pragma unroll
		for (int half = 0; half < 2; half++) {
			const uint64_t i = 2 * ip + half;
			uint32_t p = i < N ? coeff[e * ld + i] : 0u;  // c * x^j, j = 0..15
			uint32_t row = 0;
#pragma unroll
			for (int j = 0; j < 16; j++) {
				row |= ((p >> o) & 1u) << j;
				p = gf16_mulx(p);
			}
			word |= row << (16 * half);
		}
		rows[idx] = word;
	}
}

// One work-group: 4 waves; lane `col` of a wave owns NB consecutive words
// (wbase + NB*col + nb, one 32-wide column block per nb), so a lane reads them
// with one NB*16-bit load per input and writes them with one store per output.
// OT output tiles of 32 rows (2 outputs each); an input range that is a
// multiple of 4.  Stages of 4*IB inputs: the next stage's row words and input
// words are loaded while the current stage feeds MFMA, its A operands go to
// the other half of a double-buffered LDS ring, one barrier per stage.  Loads
// use clamped, in-bounds addresses: words past W and outputs past R only reach
// accumulator columns / rows that are never stored, and inputs past the range
// get zero rows when they are stored to LDS.
template <int OT, int NB, int IB, bool FP4, int OCC>
__global__ __launch_bounds__(256, OCC) void k_gf16_mfma(Gf16MfmaArgs a) {
	constexpr int KS = FP4 ? 1 : 2;             // MFMA K steps per 4 inputs
	constexpr int LPT = (IB * OT * 64 + 255) / 256;  // row words loaded per thread per stage
	static_assert((IB * OT * 64) % 256 == 0, "whole waves stage row words");
	static_assert(NB == 2 || NB == 4, "a lane's words fit one 32- or 64-bit access");
	typedef typename std::conditional<NB == 4, uint64_t, uint32_t>::type xw_t;
	__shared__ uint4 sA[2][IB][OT][KS][64];

	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wave = tid >> 6;
	const int col = lane & 31;
	const int h = lane >> 5;
	// XCD-aware order (as in gf64_mfma.hip): runs of 8 * Z ids cover 8 word
	// blocks x all Z output tiles, so the tiles that read the same words run on
	// one XCD together.
	const uint64_t Z = (a.R + 2 * OT - 1) / (2 * OT);
	const uint64_t L = blockIdx.x;
	const uint64_t rem = L % (8 * Z);
	const uint64_t xblk = (L / (8 * Z)) * 8 + (rem & 7);
	if (xblk * 4 * 32 * NB >= a.W) return;  // padding of the last run (whole work-group)
	const uint64_t wbase = (xblk * 4 + wave) * (32ull * NB);
	const uint64_t w0 = wbase + (uint64_t)NB * col;  // this lane's first word
	const uint64_t r0 = (rem >> 3) * (2 * OT);
	const uint64_t g = blockIdx.y;
	const uint64_t i_begin = g * a.in_per_group;
	uint64_t i_end = i_begin + a.in_per_group;
	if (i_end > a.N) i_end = a.N;
	const uint64_t P = (a.N + 1) / 2;
	// whole NB-word accesses: aligned rows and a lane range inside the slice
	const bool vec_in = ((reinterpret_cast<uintptr_t>(a.in_base) & (2 * NB - 1)) == 0) &&
	                    (a.in_stride % NB) == 0 && w0 + NB <= a.W;

	typedef typename std::conditional<FP4, g16_v16f, g16_v16i>::type acc_t;
	acc_t acc[OT][NB];
#pragma unroll
	for (int t = 0; t < OT; t++)
#pragma unroll
		for (int nb = 0; nb < NB; nb++)
#pragma unroll
			for (int e = 0; e < 16; e++) acc[t][nb][e] = 0;

	uint32_t pre[LPT];      // row words of the next stage
	xw_t xp[4 * IB];        // input words of the next stage (NB words per input)
	xw_t xc[4 * IB];        // input words of the current stage
	const uint64_t i_last = i_end - 1;

	auto load_stage = [&](uint64_t ibase) {
#pragma unroll
		for (int k = 0; k < LPT; k++) {
			const int j = tid + 256 * k;  // (ib, t, pair, row)
			const int ib = j / (OT * 64), t = (j >> 6) % OT, rho = j & 31, pr = (j >> 5) & 1;
			uint64_t e = r0 + 2 * t + (rho >> 4);
			e = e < a.R ? e : a.R - 1;
			uint64_t ip = (ibase + 4 * ib) / 2 + pr;
			ip = ip < P ? ip : P - 1;
			pre[k] = a.rows[(e * P + ip) * 16 + (rho & 15)];
		}
#pragma unroll
		for (int u = 0; u < 4 * IB; u++) {
			const uint64_t i = ibase + u;
			const uint64_t ic = i < i_end ? i : i_last;
			const uint16_t* src = a.in_base + ic * a.in_stride;
			if (vec_in) {
				xp[u] = *reinterpret_cast<const xw_t*>(src + w0);
			} else {
				xw_t v = 0;
#pragma unroll
				for (int nb = 0; nb < NB; nb++) {
					const uint64_t w = w0 + nb;
					v |= (xw_t)src[w < a.W ? w : a.W - 1] << (16 * nb);
				}
				xp[u] = v;
			}
		}
	};
	auto store_stage = [&](int buf, uint64_t ibase) {
#pragma unroll
		for (int k = 0; k < LPT; k++) {
			const int j = tid + 256 * k;
			const int ib = j / (OT * 64), t = (j >> 6) % OT, rho = j & 31, pr = (j >> 5) & 1;
			const uint64_t i0 = ibase + 4 * ib + 2 * pr;  // inputs of this pair: i0, i0 + 1
			uint32_t wv = pre[k] & (i0 < i_end ? 0xFFFFu : 0u);
			wv |= i0 + 1 < i_end ? (pre[k] & 0xFFFF0000u) : 0u;
			if (FP4) {
				sA[buf][ib][t][0][rho + 32 * pr] = g16_expand_fp4(wv);
			} else {
				sA[buf][ib][t][pr][rho] = g16_expand_i8(wv & 0xFFFFu);
				sA[buf][ib][t][pr][rho + 32] = g16_expand_i8(wv >> 16);
			}
		}
	};

	if (i_begin < i_end) {  // uniform over the work-group
		int buf = 0;
		load_stage(i_begin);
		store_stage(0, i_begin);
		__syncthreads();
		for (uint64_t ibase = i_begin; ibase < i_end; ibase += 4 * IB) {
#pragma unroll
			for (int u = 0; u < 4 * IB; u++) xc[u] = xp[u];
			load_stage(ibase + 4 * IB);
#pragma unroll
			for (int ib = 0; ib < IB; ib++) {
				// B operands of 4 inputs: fp4 lane h = inputs 2h, 2h+1; int8 step s,
				// lane h = input 2s + h
				typedef typename std::conditional<FP4, g16_v8i, g16_v4i>::type bop_t;
				bop_t bf[KS][NB];
#pragma unroll
				for (int nb = 0; nb < NB; nb++) {
					const uint32_t e0 = (uint32_t)(xc[4 * ib + 0] >> (16 * nb)) & 0xFFFFu;
					const uint32_t e1 = (uint32_t)(xc[4 * ib + 1] >> (16 * nb)) & 0xFFFFu;
					const uint32_t e2 = (uint32_t)(xc[4 * ib + 2] >> (16 * nb)) & 0xFFFFu;
					const uint32_t e3 = (uint32_t)(xc[4 * ib + 3] >> (16 * nb)) & 0xFFFFu;
					if constexpr (FP4) {
						const uint4 ex = g16_expand_fp4(h ? (e2 | (e3 << 16)) : (e0 | (e1 << 16)));
						bf[0][nb][0] = (int)ex.x;
						bf[0][nb][1] = (int)ex.y;
						bf[0][nb][2] = (int)ex.z;
						bf[0][nb][3] = (int)ex.w;
#pragma unroll
						for (int v = 4; v < 8; v++) bf[0][nb][v] = 0;
					} else {
						const uint4 x0 = g16_expand_i8(h ? e1 : e0);
						const uint4 x1 = g16_expand_i8(h ? e3 : e2);
						bf[0][nb][0] = (int)x0.x;
						bf[0][nb][1] = (int)x0.y;
						bf[0][nb][2] = (int)x0.z;
						bf[0][nb][3] = (int)x0.w;
						bf[KS - 1][nb][0] = (int)x1.x;
						bf[KS - 1][nb][1] = (int)x1.y;
						bf[KS - 1][nb][2] = (int)x1.z;
						bf[KS - 1][nb][3] = (int)x1.w;
					}
				}
#pragma unroll
				for (int t = 0; t < OT; t++) {
#pragma unroll
					for (int s = 0; s < KS; s++) {
						const uint4 av = sA[buf][ib][t][s][lane];
						if constexpr (FP4) {
							g16_v8i af;
							af[0] = (int)av.x;
							af[1] = (int)av.y;
							af[2] = (int)av.z;
							af[3] = (int)av.w;
#pragma unroll
							for (int v = 4; v < 8; v++) af[v] = 0;
#pragma unroll
							for (int nb = 0; nb < NB; nb++)
								acc[t][nb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
								    af, bf[0][nb], acc[t][nb], 4, 4, 0, 128, 0, 128);
						} else {
							g16_v4i af;
							af[0] = (int)av.x;
							af[1] = (int)av.y;
							af[2] = (int)av.z;
							af[3] = (int)av.w;
#pragma unroll
							for (int nb = 0; nb < NB; nb++)
								acc[t][nb] =
								    __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[s][nb], acc[t][nb], 0, 0, 0);
						}
					}
				}
			}
			store_stage(buf ^ 1, ibase + 4 * IB);
			__syncthreads();
			buf ^= 1;
		}
	}
	// parity -> 16-bit words.  Lane (col, h) holds rows 8(j/4) + 4h + (j%4) of
	// each 32-row tile: j < 8 are bits of output 2t, j >= 8 of output 2t + 1.
	// After the exchange with lane col + 32, lane half h writes output 2t + h.
#pragma unroll
	for (int t = 0; t < OT; t++) {
		xw_t ov = 0;
#pragma unroll
		for (int nb = 0; nb < NB; nb++) {
			uint32_t b0 = 0, b1 = 0;
#pragma unroll
			for (int j = 0; j < 8; j++) {
				const int o = 8 * (j >> 2) + 4 * h + (j & 3);
				b0 |= ((uint32_t)acc[t][nb][j] & 1u) << o;
				b1 |= ((uint32_t)acc[t][nb][j + 8] & 1u) << o;
			}
			b0 |= (uint32_t)__shfl_xor((int)b0, 32);
			b1 |= (uint32_t)__shfl_xor((int)b1, 32);
			ov |= (xw_t)(h ? b1 : b0) << (16 * nb);
		}
		const uint64_t e = r0 + 2 * t + h;
		if (e >= a.R) continue;
		uint16_t* dst;
		bool xr;
		if (a.partial) {
			dst = a.partial + (g * a.R + e) * a.W;
			xr = false;
		} else {
			dst = a.out_base + e * a.out_stride;
			xr = a.accumulate != 0;
		}
		if (w0 + NB <= a.W && ((reinterpret_cast<uintptr_t>(dst + w0) & (2 * NB - 1)) == 0)) {
			xw_t* q = reinterpret_cast<xw_t*>(dst + w0);
			*q = xr ? (*q ^ ov) : ov;
		} else {
#pragma unroll
			for (int nb = 0; nb < NB; nb++) {
				const uint64_t w = w0 + nb;
				const uint16_t v = (uint16_t)(ov >> (16 * nb));
				if (w < a.W) dst[w] = xr ? (uint16_t)(dst[w] ^ v) : v;
			}
		}
	}
}

// The fp4 kernel for full tiles (W a multiple of 256 words, 4-byte aligned
// rows at an even stride, 32-bit offsets; launch_gf16_mfma checks them): the
// same tiles, K order and staging as k_gf16_mfma<4, 2, 2, fp4, 2>, with the
// address work taken out of the loop.  A stage's input rows and row words are
// a uniform (scalar) base plus a per-lane offset fixed for the whole launch,
// and a lane loads only the two inputs of its B operand (lane half h: inputs
// 2h, 2h+1 of every 4), one 32-bit load each, so a B operand is one byte
// permute plus the fp4 expansion.  Stages that run past the input range (the
// last one) take the clamped form.
template <int IB>
__global__ __launch_bounds__(256, 2) void k_gf16_fp4f(Gf16MfmaArgs a) {
	constexpr int OT = 4, NB = 2;
	constexpr int LPT = IB * OT * 64 / 256;  // row words staged per thread per stage
	static_assert((IB * OT * 64) % 256 == 0, "whole waves stage row words");
	__shared__ uint4 sA[2][IB][OT][64];

	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wave = tid >> 6;
	const int col = lane & 31;
	const int h = lane >> 5;
	const uint64_t Z = (a.R + 2 * OT - 1) / (2 * OT);
	const uint64_t L = blockIdx.x;
	const uint64_t rem = L % (8 * Z);
	const uint64_t xblk = (L / (8 * Z)) * 8 + (rem & 7);
	if (xblk * 4 * 32 * NB >= a.W) return;  // padding of the last run (whole work-group)
	const uint64_t w0 = (xblk * 4 + wave) * (32ull * NB) + (uint64_t)NB * col;
	const uint64_t r0 = (rem >> 3) * (2 * OT);
	const uint64_t g = blockIdx.y;
	const uint64_t i_begin = g * a.in_per_group;
	uint64_t i_end = i_begin + a.in_per_group;
	if (i_end > a.N) i_end = a.N;
	const uint64_t P = (a.N + 1) / 2;

	// per-lane offsets, fixed for the launch: input 2h (+ in_stride for 2h+1)
	// word w0, in bytes; row words (e - r0, pair pr, bit o) in words
	const uint32_t xoff = (uint32_t)(((uint64_t)(2 * h) * a.in_stride + w0) * 2);
	const uint32_t xstep = (uint32_t)(a.in_stride * 2);
	uint32_t roff[LPT];
#pragma unroll
	for (int k = 0; k < LPT; k++) {
		const int j = tid + 256 * k;  // (ib, t, pair, row)
		const int ib = j / (OT * 64), t = (j >> 6) % OT, rho = j & 31, pr = (j >> 5) & 1;
		uint64_t e = r0 + 2 * t + (rho >> 4);
		e = e < a.R ? e : a.R - 1;
		roff[k] = (uint32_t)(((e - r0) * P + 2 * ib + pr) * 16 + (rho & 15));
	}

	g16_v16f acc[OT][NB];
#pragma unroll
	for (int t = 0; t < OT; t++)
#pragma unroll
		for (int nb = 0; nb < NB; nb++)
#pragma unroll
			for (int e = 0; e < 16; e++) acc[t][nb][e] = 0;

	// two stages of loads in flight: stage s computes from registers loaded
	// two stages before (set s & 1), while the loads of stage s + 2 refill
	// that set and stage s + 1's rows, loaded a stage earlier, are staged
	uint32_t pre2[2][LPT];     // row words
	uint32_t xp2[2][2 * IB];   // this lane's input words
	uint32_t pre[LPT];
	uint32_t xc[2 * IB];

	// a full stage: every input of [ibase, ibase + 4 IB) is in range
	auto load_full = [&](uint64_t ibase, uint32_t (&pr)[LPT], uint32_t (&xp)[2 * IB]) {
		const uint32_t* rbase = a.rows + (r0 * P + ibase / 2) * 16;
		const char* xbase = reinterpret_cast<const char*>(a.in_base + ibase * a.in_stride);
#pragma unroll
		for (int k = 0; k < LPT; k++) pr[k] = rbase[roff[k]];
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
			const char* xb = xbase + (uint64_t)(4 * ib) * a.in_stride * 2;
			xp[2 * ib] = *reinterpret_cast<const uint32_t*>(xb + xoff);
			xp[2 * ib + 1] = *reinterpret_cast<const uint32_t*>(xb + xoff + xstep);
		}
	};
	auto store_stage = [&](int buf, uint64_t ibase, bool full, const uint32_t (&rw)[LPT]) {
#pragma unroll
		for (int k = 0; k < LPT; k++) {
			const int j = tid + 256 * k;
			const int ib = j / (OT * 64), t = (j >> 6) % OT, rho = j & 31, pr = (j >> 5) & 1;
			uint32_t wv = rw[k];
			if (!full) {
				const uint64_t i0 = ibase + 4 * ib + 2 * pr;  // inputs of this pair: i0, i0 + 1
				wv = (i0 < i_end ? wv & 0xFFFFu : 0u) | (i0 + 1 < i_end ? wv & 0xFFFF0000u : 0u);
			}
			sA[buf][ib][t][rho + 32 * pr] = g16_expand_fp4(wv);
		}
	};
	// B operands of K step ib: word nb of inputs 2h, 2h+1 is one byte
	// permute plus the fp4 expansion
	auto bops = [&](int ib, g16_v8i (&bf)[NB]) {
#pragma unroll
		for (int nb = 0; nb < NB; nb++) {
			const uint32_t pw = __builtin_amdgcn_perm(xc[2 * ib + 1], xc[2 * ib], nb ? 0x07060302u : 0x05040100u);
			const uint4 ex = g16_expand_fp4(pw);
			bf[nb][0] = (int)ex.x;
			bf[nb][1] = (int)ex.y;
			bf[nb][2] = (int)ex.z;
			bf[nb][3] = (int)ex.w;
#pragma unroll
			for (int v = 4; v < 8; v++) bf[nb][v] = 0;
		}
	};
	// K step ib's MFMAs with step ib + 1's B operands computed between them:
	// the expansion (VALU) runs while the matrix core works on this step
	// (sched_group_barrier: each MFMA followed by two VALU ops)
	auto compute = [&](int buf) {
		g16_v8i bfa[NB], bfb[NB];
		bops(0, bfa);
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
			g16_v8i(&cur)[NB] = (ib & 1) ? bfb : bfa;
			g16_v8i(&nxt)[NB] = (ib & 1) ? bfa : bfb;
			uint4 av[OT];
#pragma unroll
			for (int t = 0; t < OT; t++) av[t] = sA[buf][ib][t][lane];
			if (ib + 1 < IB) bops(ib + 1 < IB ? ib + 1 : ib, nxt);
#pragma unroll
			for (int t = 0; t < OT; t++) {
				g16_v8i af;
				af[0] = (int)av[t].x;
				af[1] = (int)av[t].y;
				af[2] = (int)av[t].z;
				af[3] = (int)av[t].w;
#pragma unroll
				for (int v = 4; v < 8; v++) af[v] = 0;
#pragma unroll
				for (int nb = 0; nb < NB; nb++)
					acc[t][nb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(af, cur[nb], acc[t][nb], 4, 4, 0, 128,
					                                                             0, 128);
			}
			__builtin_amdgcn_sched_group_barrier(0x100, OT, 0);  // this step's A reads first
#pragma unroll
			for (int k = 0; k < OT * NB; k++) {
				__builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
				__builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // two VALU
			}
			__builtin_amdgcn_sched_barrier(0);
		}
	};

	// full stages in the loop; a last stage that runs past the input range
	// after it, with clamped loads (its rows past the range staged as zeros)
	const uint64_t full_end = i_begin + (i_end > i_begin ? (i_end - i_begin) / (4 * IB) * (4 * IB) : 0);
	int buf = 0;
	if (full_end > i_begin) {  // uniform over the work-group
		constexpr uint64_t SP = 4 * IB;  // inputs per stage
		// a stage past the last full one loads the last one again (harmless:
		// its staging is never read), so the body has no branch for the
		// compiler to sink the MFMAs past
		auto clampst = [&](uint64_t ib) { return ib < full_end ? ib : full_end - SP; };
		load_full(i_begin, pre2[0], xp2[0]);
		load_full(clampst(i_begin + SP), pre2[1], xp2[1]);
		store_stage(0, i_begin, true, pre2[0]);
		__syncthreads();
		// one stage with its loads in set S: compute from set S, refill S
		// with stage + 2, stage stage + 1 from set S ^ 1
		auto step = [&](uint64_t ibase, auto S) {
			constexpr int cs = decltype(S)::value;
#pragma unroll
			for (int u = 0; u < 2 * IB; u++) xc[u] = xp2[cs][u];
			load_full(clampst(ibase + 2 * SP), pre2[cs], xp2[cs]);
			compute(buf);
			store_stage(buf ^ 1, clampst(ibase + SP), true, pre2[cs ^ 1]);
			__syncthreads();
			buf ^= 1;
		};
		for (uint64_t ibase = i_begin; ibase < full_end; ibase += 2 * SP) {
			step(ibase, std::integral_constant<int, 0>());
			if (ibase + SP >= full_end) break;
			step(ibase + SP, std::integral_constant<int, 1>());
		}
	}
	if (full_end < i_end) {
		const uint64_t ibase = full_end;
		const uint64_t i_last = i_end - 1;
#pragma unroll
		for (int k = 0; k < LPT; k++) {
			const int j = tid + 256 * k;
			const int ib = j / (OT * 64), t = (j >> 6) % OT, rho = j & 31, pr = (j >> 5) & 1;
			uint64_t e = r0 + 2 * t + (rho >> 4);
			e = e < a.R ? e : a.R - 1;
			uint64_t ip = ibase / 2 + 2 * ib + pr;
			ip = ip < P ? ip : P - 1;
			pre[k] = a.rows[(e * P + ip) * 16 + (rho & 15)];
		}
#pragma unroll
		for (int ib = 0; ib < IB; ib++)
#pragma unroll
			for (int v = 0; v < 2; v++) {
				uint64_t i = ibase + 4 * ib + 2 * h + v;
				i = i < i_end ? i : i_last;
				xc[2 * ib + v] = *reinterpret_cast<const uint32_t*>(a.in_base + i * a.in_stride + w0);
			}
		store_stage(buf, ibase, false, pre);
		__syncthreads();
		compute(buf);
	}
	// parity -> 16-bit words (as k_gf16_mfma): lane (col, h) holds rows
	// 8(j/4) + 4h + (j%4) of each 32-row tile; after the exchange with lane
	// col + 32, lane half h writes output 2t + h, words w0, w0 + 1.
#pragma unroll
	for (int t = 0; t < OT; t++) {
		uint32_t ov = 0;
#pragma unroll
		for (int nb = 0; nb < NB; nb++) {
			uint32_t b0 = 0, b1 = 0;
#pragma unroll
			for (int j = 0; j < 8; j++) {
				const int o = 8 * (j >> 2) + 4 * h + (j & 3);
				b0 |= ((uint32_t)acc[t][nb][j] & 1u) << o;
				b1 |= ((uint32_t)acc[t][nb][j + 8] & 1u) << o;
			}
			b0 |= (uint32_t)__shfl_xor((int)b0, 32);
			b1 |= (uint32_t)__shfl_xor((int)b1, 32);
			ov |= (h ? b1 : b0) << (16 * nb);
		}
		const uint64_t e = r0 + 2 * t + h;
		if (e >= a.R) continue;
		uint16_t* dst;
		bool xr;
		if (a.partial) {
			dst = a.partial + (g * a.R + e) * a.W;
			xr = false;
		} else {
			dst = a.out_base + e * a.out_stride;
			xr = a.accumulate != 0;
		}
		if ((reinterpret_cast<uintptr_t>(dst + w0) & 3u) == 0) {
			uint32_t* q = reinterpret_cast<uint32_t*>(dst + w0);
			*q = xr ? (*q ^ ov) : ov;
		} else {
			dst[w0] = xr ? (uint16_t)(dst[w0] ^ ov) : (uint16_t)ov;
			dst[w0 + 1] = xr ? (uint16_t)(dst[w0 + 1] ^ (ov >> 16)) : (uint16_t)(ov >> 16);
		}
	}
}

hipError_t launch_gf16_rows(const uint16_t* coeff, uint64_t ld, uint64_t N, uint64_t R, uint32_t* rows,
                            hipStream_t s) {
	const uint64_t total = R * ((N + 1) / 2) * 16;
	if (!total) return hipSuccess;
	uint64_t blocks = (total + 255) / 256;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_gf16_rows, dim3((unsigned)blocks), dim3(256), 0, s, coeff, ld, N, R, rows);
	return hipGetLastError();
}

template <int OT, int NB, int IB, bool FP4, int OCC>
static hipError_t launch_g16m_t(const Gf16MfmaArgs& a, uint64_t groups, hipStream_t s) {
	const uint64_t span = 4ull * 32 * NB;
	const uint64_t nx = (a.W + span - 1) / span, Z = (a.R + 2 * OT - 1) / (2 * OT);
	const uint64_t blocks = (nx + 7) / 8 * 8 * Z;
	if (blocks > 0x7fffffffu || groups > 65535) return hipErrorInvalidValue;
	hipLaunchKernelGGL((k_gf16_mfma<OT, NB, IB, FP4, OCC>), dim3((unsigned)blocks, (unsigned)groups, 1),
	                   dim3(256), 0, s, a);
	return hipGetLastError();
}

// The (OT, NB) pair selects the variant; (4, 2) runs the two-waves-per-SIMD form.
bool gf16_mfma_supported(int ot, int nb, bool /*fp4*/) {
	return (ot == 4 && (nb == 4 || nb == 2)) || (ot == 8 && nb == 2);
}

// Whether the full-tile fp4 kernel takes the launch: 256-word tiles, 4-byte
// aligned input rows at an even stride, the per-lane offsets in 32 bits, and
// stages of 8 inputs that never straddle split-K rows.
static bool fp4f_ok(const Gf16MfmaArgs& a) {
	const uint64_t P = (a.N + 1) / 2;
	return a.W % 256 == 0 && (reinterpret_cast<uintptr_t>(a.in_base) & 3u) == 0 && a.in_stride % 2 == 0 &&
	       (3 * a.in_stride + a.W) * 2 < (1ull << 32) && 8 * P * 16 + 64 < (1ull << 32) &&
	       a.in_per_group % 8 == 0;
}

hipError_t launch_gf16_mfma(const Gf16MfmaArgs& a, uint64_t groups, int ot, int nb, int ib, bool fp4,
                            hipStream_t s) {
	if (fp4 && ot == 4 && nb == 2 && fp4f_ok(a) && (ib == 2 || ib == 4)) {
		const uint64_t span = 4ull * 32 * 2;
		const uint64_t nx = (a.W + span - 1) / span, Z = (a.R + 7) / 8;
		const uint64_t blocks = (nx + 7) / 8 * 8 * Z;
		if (blocks > 0x7fffffffu || groups > 65535) return hipErrorInvalidValue;
		if (ib == 4 && a.in_per_group % 16 == 0)
			hipLaunchKernelGGL((k_gf16_fp4f<4>), dim3((unsigned)blocks, (unsigned)groups, 1), dim3(256), 0, s, a);
		else
			hipLaunchKernelGGL((k_gf16_fp4f<2>), dim3((unsigned)blocks, (unsigned)groups, 1), dim3(256), 0, s, a);
		return hipGetLastError();
	}
	if (ot == 4 && nb == 4) return fp4 ? launch_g16m_t<4, 4, 4, true, 1>(a, groups, s)
	                                   : launch_g16m_t<4, 4, 4, false, 1>(a, groups, s);
	if (ot == 4 && nb == 2) return fp4 ? launch_g16m_t<4, 2, 2, true, 2>(a, groups, s)
	                                   : launch_g16m_t<4, 2, 2, false, 2>(a, groups, s);
	if (ot == 8 && nb == 2) return fp4 ? launch_g16m_t<8, 2, 4, true, 1>(a, groups, s)
	                                   : launch_g16m_t<8, 2, 4, false, 1>(a, groups, s);
	return hipErrorInvalidValue;
}

}  // namespace parhip
