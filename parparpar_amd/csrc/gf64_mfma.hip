// gf64_mfma.hip -- GF(2^64) contraction as a GF(2) bit-matrix product on MFMA.
//
// Multiplication by a coefficient c is a 64x64 matrix over GF(2): column j is
// c * x^j.  Stacking R outputs and N inputs, the contraction for one word w is
//     out_bits[(r,o)] = parity( sum_{(i,j)} M_{c(r,i)}[o][j] * bit_j(D[i][w]) )
// an integer matrix product whose parity is the GF(2) result.  Rows of M_c
// (Row_o, bit j = bit o of c*x^j) are prepared once per coefficient by
// k_gf64_rows with 64 wave ballots (a 64x64 bit transpose), then the main
// kernel feeds MFMA with bits expanded to 0/1 elements:
//   - int8:  v_mfma_i32_32x32x32_i8, parity = bit 0 of the int32 accumulator;
//   - fp4:   v_mfma_scale_f32_32x32x64_f8f6f4 with e2m1 elements 0 / 0.5 and
//            block scale 2, so every product is 0 or 1 and the fp32 sum is the
//            exact count (< 2^24), parity = bit 0 of its integer value.
// Operand layout (verified on the device): lane l holds row/column l%32 and
// K block l/32; the K order inside a lane is free as long as A and B agree,
// so it is chosen to make bit expansion a shift and a mask.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "gf64_dev.h"
#include "kernels.h"

namespace parhip {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// rows[(r*N + i)*64 + o] = Row_o(c(r,i)): one wave per coefficient.
__global__ __launch_bounds__(256) void k_gf64_rows(const uint64_t* __restrict__ coeff, uint64_t ld,
                                                   uint64_t N, uint64_t R,
                                                   uint64_t* __restrict__ rows) {
	const uint64_t lane = threadIdx.x & 63;
	const uint64_t total = N * R;
	for (uint64_t q = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < total;
	     q += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
		const uint64_t r = q / N, i = q - r * N;
		const uint64_t c = coeff[r * ld + i];
		const uint64_t p = gf64_mulx_pow(c, (unsigned)lane);  // column `lane` of M_c
		uint64_t mine = 0;
#pragma unroll 8
		for (int o = 0; o < 64; o++) {
			const uint64_t m = __ballot((p >> o) & 1u);  // bit j = bit o of c*x^j
			if ((int)lane == o) mine = m;
		}
		rows[q * 64 + lane] = mine;
	}
}

__device__ __forceinline__ uint32_t spread4_i8(uint32_t x) {  // 4 low bits -> 4 bytes of 0/1
	return ((x & 0xFu) * 0x00204081u) & 0x01010101u;
}

__device__ __forceinline__ uint4 expand_fp4(uint32_t x) {  // bit v+4q -> nibble q of dword v
	return make_uint4(x & 0x11111111u, (x >> 1) & 0x11111111u, (x >> 2) & 0x11111111u,
	                  (x >> 3) & 0x11111111u);
}

// One work-group: 4 waves x NB column blocks of 32 words, RT outputs, an input
// range.  The A operands (expanded coefficient rows) are shared by the 4 waves:
// they are staged IB inputs at a time into a double-buffered LDS ring, the
// next stage's global loads (rows and input words) in flight while the current
// one feeds MFMA; one barrier per stage.
//   fp4:  A fragment (r, t): lane L = col + 32h holds bits 32h + v + 4q of row
//         32t + col as nibble q of dword v;
//   int8: A fragment (r, t, s): bits 32s + 16h + 4v + q as byte q of dword v.
template <int RT, int NB, int IB, bool FP4, int OCC, int PD = 4>
__global__ __launch_bounds__(256, OCC) void k_mac_mfma(MfmaArgs a) {
	constexpr int KS = FP4 ? 1 : 2;          // K steps of the MFMA per input
	constexpr int FR = RT * 2 * KS;          // A fragments per input
	// coefficient rows staged per thread; with fewer rows than threads (RT=2)
	// two waves stage the same rows and store identical values, which keeps
	// the loop free of branches
	constexpr int RPT = (RT * 64 + 255) / 256;
	static_assert(RT * 64 % 256 == 0 || 256 % (RT * 64) == 0, "whole waves stage coefficient rows");
	__shared__ uint4 sA[2][IB][FR][64];
	typedef typename std::conditional<FP4, uint32_t, uint64_t>::type xw_t;

	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wave = tid >> 6;
	const int col = lane & 31;  // word within a block / output row within a tile
	const int h = lane >> 5;    // K block
	// XCD-aware order: work-groups are dispatched round-robin over the 8 XCDs,
	// so ids L and L + 8 share an XCD's L2.  Each run of 8 * Z ids covers 8
	// word blocks x all Z output tiles, tile-major: the Z tiles that read the
	// same input words run on one XCD at the same time and the input is
	// fetched from HBM once.
	const uint64_t Z = (a.R + RT - 1) / RT;
	const uint64_t L = blockIdx.x;
	const uint64_t rem = L % (8 * Z);
	const uint64_t xblk = (L / (8 * Z)) * 8 + (rem & 7);
	if (xblk * 4 * 32 * NB >= a.W) return;  // padding of the last run (whole work-group)
	const uint64_t wbase = (xblk * 4 + wave) * (32ull * NB);
	const uint64_t r0 = (rem >> 3) * RT;
	const uint64_t g = blockIdx.y;
	const uint64_t i_begin = g * a.in_per_group;
	uint64_t i_end = i_begin + a.in_per_group;
	if (i_end > a.N) i_end = a.N;

	typedef typename std::conditional<FP4, v16f, v16i>::type acc_t;
	acc_t acc[RT][2][NB];
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int t = 0; t < 2; t++)
#pragma unroll
			for (int nb = 0; nb < NB; nb++)
#pragma unroll
				for (int e = 0; e < 16; e++) acc[r][t][nb][e] = 0;

	uint64_t pre[IB][RPT];  // coefficient rows of the next stage
	xw_t xp[IB][NB];        // input words of the next stage
	xw_t xc[IB][NB];        // input words of the current stage

	// Loads are unconditional from clamped, in-bounds addresses.  Words past W
	// and outputs past R only reach accumulator columns/rows that are never
	// written, so they need no masking; an input past the range (clamped to the
	// last one) is cancelled by storing its coefficient rows as zero.  The
	// mask is applied when the rows are stored to LDS, at the end of the
	// stage, so the loads stay in flight under the MFMAs.  The hot loop is
	// straight-line code and the accumulators stay in AGPRs.
	const uint64_t i_last = i_end - 1;
	auto load_stage = [&](uint64_t ibase) {
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
			const uint64_t i = ibase + ib;
			const uint64_t ic = i < i_end ? i : i_last;
#pragma unroll
			for (int k = 0; k < RPT; k++) {
				const int j = (tid + 256 * k) % (RT * 64);
				const uint64_t rr = r0 + (j >> 6);
				pre[ib][k] = a.rows[((rr < a.R ? rr : a.R - 1) * a.rows_ld + ic) * 64 + (j & 63)];
			}
			const uint64_t* src = a.in_base + ic * a.in_stride;
#pragma unroll
			for (int nb = 0; nb < NB; nb++) {
				const uint64_t w = wbase + nb * 32 + col;
				const uint64_t wc = w < a.W ? w : a.W - 1;
				if (FP4)
					xp[ib][nb] = (xw_t)reinterpret_cast<const uint32_t*>(src + wc)[h];
				else
					xp[ib][nb] = (xw_t)src[wc];
			}
		}
	};
	auto store_stage = [&](int buf, uint64_t ibase) {
#pragma unroll
		for (int ib = 0; ib < IB; ib++)
#pragma unroll
			for (int k = 0; k < RPT; k++) {
				const int j = (tid + 256 * k) % (RT * 64);
				{
					const int r = j >> 6, o = j & 63, t = o >> 5, c = o & 31;
					const uint64_t row = (ibase + ib < i_end) ? pre[ib][k] : 0u;
					if (FP4) {
						sA[buf][ib][r * 2 + t][c] = expand_fp4((uint32_t)row);
						sA[buf][ib][r * 2 + t][c + 32] = expand_fp4((uint32_t)(row >> 32));
					} else {
#pragma unroll
						for (int s = 0; s < 2; s++)
#pragma unroll
							for (int hh = 0; hh < 2; hh++) {
								const uint32_t b = (uint32_t)(row >> (32 * s + 16 * hh));
								sA[buf][ib][(r * 2 + t) * 2 + s][c + 32 * hh] =
								    make_uint4(spread4_i8(b), spread4_i8(b >> 4), spread4_i8(b >> 8),
								               spread4_i8(b >> 12));
							}
					}
				}
			}
	};

	if (i_begin < i_end) {  // uniform: the whole work-group has the same range
		int buf = 0;
		load_stage(i_begin);
		store_stage(0, i_begin);
		__syncthreads();
		for (uint64_t ibase = i_begin; ibase < i_end; ibase += IB) {
#pragma unroll
			for (int ib = 0; ib < IB; ib++)
#pragma unroll
				for (int nb = 0; nb < NB; nb++) xc[ib][nb] = xp[ib][nb];
			load_stage(ibase + IB);  // past the range: clamped loads, stored as zero
			// B operands of the stage, then the MFMA sequence over its NF A
			// fragments with a rolling window of PD fragments read ahead from
			// LDS; the sched barriers keep that order so the LDS latency hides
			// behind the MFMAs already issued.
			constexpr int NF = IB * FR;
			typedef typename std::conditional<FP4, v8i, v4i>::type bop_t;
			bop_t bf[IB][KS][NB];
#pragma unroll
			for (int ib = 0; ib < IB; ib++)
#pragma unroll
				for (int nb = 0; nb < NB; nb++) {
					if (FP4) {
						const uint4 e = expand_fp4((uint32_t)xc[ib][nb]);
						bf[ib][0][nb][0] = (int)e.x;
						bf[ib][0][nb][1] = (int)e.y;
						bf[ib][0][nb][2] = (int)e.z;
						bf[ib][0][nb][3] = (int)e.w;
						if constexpr (FP4)
#pragma unroll
							for (int v = 4; v < 8; v++) bf[ib][0][nb][v] = 0;
					} else {
#pragma unroll
						for (int s = 0; s < KS; s++)
#pragma unroll
							for (int v = 0; v < 4; v++)
								bf[ib][s][nb][v] = (int)spread4_i8(
								    (uint32_t)((uint64_t)xc[ib][nb] >> (32 * s + 16 * h + 4 * v)));
					}
				}
			uint4 win[PD];
#pragma unroll
			for (int f = 0; f < PD; f++) win[f] = sA[buf][f / FR][f % FR][lane];
#pragma unroll
			for (int f = 0; f < NF; f++) {
				const int ib = f / FR, q = f % FR;
				const int r = q / (2 * KS), t = (q / KS) & 1, sk = q % KS;
				const uint4 av = win[f % PD];
				if (f + PD < NF) win[f % PD] = sA[buf][(f + PD) / FR][(f + PD) % FR][lane];
				if constexpr (FP4) {
					v8i af;
					af[0] = (int)av.x;
					af[1] = (int)av.y;
					af[2] = (int)av.z;
					af[3] = (int)av.w;
#pragma unroll
					for (int v = 4; v < 8; v++) af[v] = 0;
#pragma unroll
					for (int nb = 0; nb < NB; nb++)
						acc[r][t][nb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
						    af, bf[ib][0][nb], acc[r][t][nb], 4, 4, 0, 128, 0, 128);
				} else {
					v4i af;
					af[0] = (int)av.x;
					af[1] = (int)av.y;
					af[2] = (int)av.z;
					af[3] = (int)av.w;
#pragma unroll
					for (int nb = 0; nb < NB; nb++)
						acc[r][t][nb] =
						    __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[ib][sk][nb], acc[r][t][nb], 0, 0, 0);
				}
				__builtin_amdgcn_sched_barrier(0);
			}
			store_stage(buf ^ 1, ibase + IB);
			__syncthreads();
			buf ^= 1;
		}
	}
	// parity -> bits of output word (rows o = 32t + 8(j/4) + 4h + (j%4))
#pragma unroll
	for (int r = 0; r < RT; r++) {
		const uint64_t rr = r0 + r;
#pragma unroll
		for (int nb = 0; nb < NB; nb++) {
			uint64_t bits = 0;
#pragma unroll
			for (int t = 0; t < 2; t++)
#pragma unroll
				for (int j = 0; j < 16; j++) {
					const uint32_t cnt = (uint32_t)acc[r][t][nb][j];  // exact count
					bits |= (uint64_t)(cnt & 1u) << (32 * t + 8 * (j >> 2) + 4 * h + (j & 3));
				}
			const uint64_t other = __shfl_xor(bits, 32);
			bits |= other;
			const uint64_t w = wbase + nb * 32 + col;
			if (h == 0 && rr < a.R && w < a.W) {
				if (a.partial)
					a.partial[(g * a.R + rr) * a.W + w] = bits;
				else {
					uint64_t* dst = a.out_base + rr * a.out_stride + w;
					*dst = a.accumulate ? (*dst ^ bits) : bits;
				}
			}
		}
	}
}

hipError_t launch_gf64_rows(const uint64_t* coeff, uint64_t ld, uint64_t N, uint64_t R,
                            uint64_t* rows, hipStream_t s) {
	const uint64_t waves = N * R;
	if (!waves) return hipSuccess;
	uint64_t blocks = (waves + 3) / 4;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_gf64_rows, dim3((unsigned)blocks), dim3(256), 0, s, coeff, ld, N, R, rows);
	return hipGetLastError();
}

template <int RT, int NB, int IB, bool FP4, int OCC>
static hipError_t launch_mfma_t(const MfmaArgs& a, uint64_t groups, hipStream_t s) {
	const uint64_t span = 4ull * 32 * NB;
	const uint64_t nx = (a.W + span - 1) / span, Z = (a.R + RT - 1) / RT;
	const uint64_t blocks = (nx + 7) / 8 * 8 * Z;  // runs of 8 word blocks x Z tiles
	if (blocks > 0x7fffffffu || groups > 65535) return hipErrorInvalidValue;
	dim3 grid((unsigned)blocks, (unsigned)groups, 1);
	hipLaunchKernelGGL((k_mac_mfma<RT, NB, IB, FP4, OCC>), grid, dim3(256), 0, s, a);
	return hipGetLastError();
}

// (RT, NB, IB, fp4, waves per SIMD): 256 accumulators run one wave per SIMD,
// 128 accumulators two
#define PARHIP_MFMA_VARIANTS(X) \
	X(4, 2, 4, true, 1)          \
	X(4, 2, 2, true, 1)          \
	X(4, 1, 2, true, 2)          \
	X(2, 2, 2, true, 2)          \
	X(2, 2, 4, true, 2)          \
	X(8, 1, 2, true, 1)          \
	X(4, 2, 2, false, 1)         \
	X(4, 1, 2, false, 2)         \
	X(2, 2, 2, false, 2)

bool mfma_supported(int rt, int nb, int ib, bool fp4) {
#define X(R_, N_, I_, F_, O_) \
	if (rt == R_ && nb == N_ && ib == I_ && fp4 == F_) return true;
	PARHIP_MFMA_VARIANTS(X)
#undef X
	return false;
}

hipError_t launch_mac_mfma(const MfmaArgs& a, uint64_t groups, int rt, int nb, int ib, bool fp4,
                           hipStream_t s) {
#define X(R_, N_, I_, F_, O_) \
	if (rt == R_ && nb == N_ && ib == I_ && fp4 == F_) return launch_mfma_t<R_, N_, I_, F_, O_>(a, groups, s);
	PARHIP_MFMA_VARIANTS(X)
#undef X
	return hipErrorInvalidValue;
}

}  // namespace parhip
