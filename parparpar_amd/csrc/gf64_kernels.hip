// gf64_kernels.hip -- GF(2^64) recovery-slice kernels for gfx950 (MI355X).
//
// The hot operation is the contraction
//     R[r][w] ^= XOR_i c(r,i) * D[i][w]          (64-bit LE words, GF(2^64))
// behind ComputeRecoveryBlocks* / SolveAndReconstruct / the region ops of the
// reference (src/par3_engine.cc:784-911, gf64/gf64_region_scalar.c:8-91).
//
// Strategy "nibble tables" (k_mac_tab): a work-group owns a tile of words and
// a tile of RT outputs.  For every input it stages, per coefficient, sixteen
// 16-entry split tables T_k[n] = c * (n x^{4k}) in LDS, two outputs interleaved
// per 16-byte entry, so one ds_read_b128 serves two outputs and a 16-entry
// table occupies exactly one 256-byte LDS row (conflict-free for any mix of
// nibble values across the wave).  Each word of input is read from HBM once
// (128-bit loads) and applied to all RT accumulators held in VGPRs.  Splitting
// the input range over work-groups (split-K) gives enough work-groups for 256
// CUs; partial results are XOR-reduced by k_xor_reduce.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gf64_dev.h"
#include "gf64_dyadic16_dev.h"
#include "gf64_dyadic32_dev.h"
#include "gf64_dyadic_dev.h"
#include "kernels.h"

namespace parhip {

// --------------------------------------------------------------------------
// Cauchy coefficients (src/par3_engine.cc:358-429):
//   M[r][c] = inv((firstInput + c) XOR (firstRecovery + r)), 0 -> 1.
// --------------------------------------------------------------------------
__global__ void k_cauchy(uint64_t* __restrict__ out, uint64_t N, uint64_t R, uint64_t fi,
                         uint64_t fr) {
	uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint64_t total = N * R;
	for (; idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
		uint64_t r = idx / N, c = idx - r * N;
		uint64_t d = (fi + c) ^ (fr + r);
		out[idx] = gf64_inv(d ? d : 1);
	}
}

// Same coefficients for an arbitrary list of input points and recovery points
// (repair: rows = available recovery ids, columns = missing or known inputs).
__global__ void k_cauchy_points(uint64_t* __restrict__ out, const uint64_t* __restrict__ xs,
                                uint64_t nx, const uint64_t* __restrict__ ys, uint64_t ny) {
	uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint64_t total = nx * ny;
	for (; idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
		uint64_t r = idx / nx, c = idx - r * nx;
		uint64_t d = xs[c] ^ ys[r];
		out[idx] = gf64_inv(d ? d : 1);
	}
}

// --------------------------------------------------------------------------
// Nibble-table contraction.
// --------------------------------------------------------------------------
// Global-memory views: the engine only hands device (HBM) pointers to the
// kernels; loading through address space 1 keeps these off the flat path
// (flat loads share the LDS wait counter and serialise against ds_read).
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const uint64_t gu64;
typedef __attribute__((address_space(1))) const ulonglong2 gu128;
__device__ __forceinline__ uint64_t gload(const uint64_t* p) { return *(gu64*)p; }
__device__ __forceinline__ ulonglong2 gload2(const uint64_t* p) {
	gu128* q = (gu128*)p;
	ulonglong2 v;
	v.x = q->x;
	v.y = q->y;
	return v;
}
#else
__device__ __forceinline__ uint64_t gload(const uint64_t* p) { return *p; }
__device__ __forceinline__ ulonglong2 gload2(const uint64_t* p) {
	return *reinterpret_cast<const ulonglong2*>(p);
}
#endif

// a ^ b ^ c in one VALU op (gfx950 v_bitop3_b32, truth table 0x96)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
	return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// One tile of the contraction: words [w0, w0 + 256*WPT), outputs [r0, r0+RT),
// inputs [i_begin, i_end).  `tab` is IB*(RT/2)*256 uint4 of LDS.  out_mode:
//   OUT_DIRECT  store (or XOR when a.accumulate) into the outputs,
//   OUT_PARTIAL store into a.partial row g (split-K, reduced afterwards),
//   OUT_ATOMIC  64-bit atomic XOR into the outputs (split-K inside one
//               persistent launch; XOR is order-independent, so the result is
//               deterministic).
enum { OUT_DIRECT = 0, OUT_PARTIAL = 1, OUT_ATOMIC = 2 };

// Loads of one batch of IB inputs for mac_tab_tile: this thread's words of
// each input (zeros past the input range or the slice end) and the two
// coefficients of the table entry it builds (thread t < IB*NP*16 owns entry
// (ib, p, n) = (t/16/NP, t/16%NP, t%16)).  The coefficient loads are issued
// first, so a wait for them leaves the data loads in flight.  `fast`: strided,
// 16-byte aligned slices and a tile that lies inside the slice (no per-lane checks).
template <int RT, int WPT, int IB>
__device__ __forceinline__ void mac_tab_load(const MacArgs& a, uint64_t ib0, uint64_t i_end,
                                             uint64_t w0, uint64_t r0, bool fast,
                                             uint2 (&d)[IB][WPT], uint64_t& c0, uint64_t& c1) {
	constexpr int NP = RT / 2;
	const int tid = threadIdx.x;
	c0 = 0;
	c1 = 0;
	if (tid < IB * NP * 16) {
		const int p = (tid >> 4) % NP;
		const int ib = (tid >> 4) / NP;
		const uint64_t i = ib0 + ib;
		if (i < i_end) {
			const uint64_t ra = r0 + 2 * p, rb = ra + 1;
			if (ra < a.R) c0 = gload(a.coeff + ra * a.coeff_ld + i);
			if (rb < a.R) c1 = gload(a.coeff + rb * a.coeff_ld + i);
		}
	}
#pragma unroll
	for (int ib = 0; ib < IB; ib++) {
		const uint64_t i = ib0 + ib;
#pragma unroll
		for (int j = 0; j < WPT / 2; j++) {
			const uint64_t w = w0 + 2 * tid + 512 * j;
			uint64_t v0 = 0, v1 = 0;
			if (i < i_end) {
				if (fast) {
					const ulonglong2 v = gload2(a.in_base + i * a.in_stride + w);
					v0 = v.x;
					v1 = v.y;
				} else {
					const uint64_t* src = a.in_ptrs ? a.in_ptrs[i] : a.in_base + i * a.in_stride;
					if (w + 1 < a.W && ((reinterpret_cast<uintptr_t>(src) & 15u) == 0)) {
						const ulonglong2 v = gload2(src + w);
						v0 = v.x;
						v1 = v.y;
					} else {
						if (w < a.W) v0 = gload(src + w);
						if (w + 1 < a.W) v1 = gload(src + w + 1);
					}
				}
			}
			d[ib][2 * j] = make_uint2((uint32_t)v0, (uint32_t)(v0 >> 32));
			d[ib][2 * j + 1] = make_uint2((uint32_t)v1, (uint32_t)(v1 >> 32));
		}
	}
}

template <int RT, int WPT, int IB>
__device__ __forceinline__ void mac_tab_tile(const MacArgs& a, uint4* tab, uint64_t w0,
                                             uint64_t r0, uint64_t i_begin, uint64_t i_end,
                                             uint64_t g, int out_mode) {
	static_assert(RT % 2 == 0 && RT >= 2 && RT <= 16, "RT");
	static_assert(WPT % 2 == 0 && WPT >= 2 && WPT <= 8, "WPT");
	constexpr int NP = RT / 2;  // output pairs
	static_assert(IB * NP * 16 <= 256, "one table entry per thread");
	// tab[((ib*NP + p)*16 + k)*16 + n] = (c_{2p} * n x^{4k}, c_{2p+1} * n x^{4k})
#define TAB(ib, p, k, n) tab[(((ib) * NP + (p)) * 16 + (k)) * 16 + (n)]
	const int tid = threadIdx.x;
	const bool fast = !a.in_ptrs && ((reinterpret_cast<uintptr_t>(a.in_base) & 15u) == 0) &&
	                  (a.in_stride & 1) == 0 && w0 + 256ull * WPT <= a.W;

	// word positions owned by this thread: pairs (2*tid + 512*j, +1);
	// acc[j][p] = (lo, hi) of output 2p and (lo, hi) of output 2p+1
	uint4 acc[WPT][NP];
#pragma unroll
	for (int j = 0; j < WPT; j++)
#pragma unroll
		for (int p = 0; p < NP; p++) acc[j][p] = make_uint4(0, 0, 0, 0);

	// Software pipeline: the loads of batch b+1 are issued before the lookups
	// of batch b, so their HBM latency is hidden behind LDS work; the table
	// build of a batch then never waits for global memory.
	uint2 d[IB][WPT];
	uint64_t c0, c1;
	mac_tab_load<RT, WPT, IB>(a, i_begin, i_end, w0, r0, fast, d, c0, c1);
	for (uint64_t ib0 = i_begin; ib0 < i_end; ib0 += IB) {
		// ---- build the split tables for IB inputs x RT outputs.  Thread
		// (ib, p, n) owns entry n of every position k of its table pair:
		// entry(k, n) = n(x) * c * x^{4k} = x^4 * entry(k-1, n), so it starts from
		// the 4-bit product n*c and steps by x^4.  In each store the 16 lanes
		// of a group write 16 consecutive entries of one 256-byte row, i.e. all
		// 64 banks once: the build is free of LDS bank conflicts.
		__syncthreads();  // previous batch finished reading tab
		if (tid < IB * NP * 16) {
			const int n = tid & 15;
			const int p = (tid >> 4) % NP;
			const int ib = (tid >> 4) / NP;
			uint64_t e0 = gf64_mul_nibble(c0, (unsigned)n);
			uint64_t e1 = gf64_mul_nibble(c1, (unsigned)n);
#pragma unroll
			for (int k = 0; k < 16; k++) {
				if (k) {
					e0 = gf64_mulx4(e0);
					e1 = gf64_mulx4(e1);
				}
				uint4 v;
				v.x = (uint32_t)e0;
				v.y = (uint32_t)(e0 >> 32);
				v.z = (uint32_t)e1;
				v.w = (uint32_t)(e1 >> 32);
				TAB(ib, p, k, n) = v;
			}
		}
		__syncthreads();
		// ---- next batch's loads, in flight during this batch's lookups
		uint2 dn[IB][WPT];
		uint64_t n0, n1;
		mac_tab_load<RT, WPT, IB>(a, ib0 + IB, i_end, w0, r0, fast, dn, n0, n1);
		// ---- lookups: four nibble positions per step, folded with 3-input XORs
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
#pragma unroll
			for (int j = 0; j < WPT; j++) {
				const uint2 x = d[ib][j];
#pragma unroll
				for (int kq = 0; kq < 16; kq += 4) {
					unsigned n[4];
#pragma unroll
					for (int u = 0; u < 4; u++) {
						const int k = kq + u;
						const uint32_t half = k < 8 ? x.x : x.y;
						n[u] = (half >> (4 * (k & 7))) & 15u;
					}
#pragma unroll
					for (int p = 0; p < NP; p++) {
						const uint4 v0 = TAB(ib, p, kq + 0, n[0]);
						const uint4 v1 = TAB(ib, p, kq + 1, n[1]);
						const uint4 v2 = TAB(ib, p, kq + 2, n[2]);
						const uint4 v3 = TAB(ib, p, kq + 3, n[3]);
						uint4 c = acc[j][p];
						c.x = xor3(xor3(c.x, v0.x, v1.x), v2.x, v3.x);
						c.y = xor3(xor3(c.y, v0.y, v1.y), v2.y, v3.y);
						c.z = xor3(xor3(c.z, v0.z, v1.z), v2.z, v3.z);
						c.w = xor3(xor3(c.w, v0.w, v1.w), v2.w, v3.w);
						acc[j][p] = c;
					}
				}
			}
		}
#pragma unroll
		for (int ib = 0; ib < IB; ib++)
#pragma unroll
			for (int j = 0; j < WPT; j++) d[ib][j] = dn[ib][j];
		c0 = n0;
		c1 = n1;
	}
#undef TAB

	// ---- write out
#pragma unroll
	for (int r = 0; r < RT; r++) {
		const uint64_t rr = r0 + r;
		if (rr >= a.R) break;
		uint64_t* dst;
		bool xr;
		if (out_mode == OUT_PARTIAL) {
			dst = a.partial + (g * a.R + rr) * a.W;
			xr = false;
		} else {
			dst = a.out_ptrs ? a.out_ptrs[rr] : a.out_base + rr * a.out_stride;
			xr = a.accumulate != 0;
		}
		const int p = r >> 1;
#pragma unroll
		for (int j = 0; j < WPT / 2; j++) {
			const uint64_t w = w0 + 2 * tid + 512 * j;
			const uint4 c0 = acc[2 * j][p], c1 = acc[2 * j + 1][p];
			uint64_t o0 = (r & 1) ? (((uint64_t)c0.w << 32) | c0.z) : (((uint64_t)c0.y << 32) | c0.x);
			uint64_t o1 = (r & 1) ? (((uint64_t)c1.w << 32) | c1.z) : (((uint64_t)c1.y << 32) | c1.x);
			if (out_mode == OUT_ATOMIC) {
				if (w < a.W && o0) atomicXor((unsigned long long*)(dst + w), (unsigned long long)o0);
				if (w + 1 < a.W && o1)
					atomicXor((unsigned long long*)(dst + w + 1), (unsigned long long)o1);
			} else if (w + 1 < a.W && ((reinterpret_cast<uintptr_t>(dst) & 15u) == 0)) {
				ulonglong2* q = reinterpret_cast<ulonglong2*>(dst + w);
				if (xr) {
					const ulonglong2 o = *q;
					o0 ^= o.x;
					o1 ^= o.y;
				}
				ulonglong2 v;
				v.x = o0;
				v.y = o1;
				*q = v;
			} else {
				if (w < a.W) dst[w] = xr ? (dst[w] ^ o0) : o0;
				if (w + 1 < a.W) dst[w + 1] = xr ? (dst[w + 1] ^ o1) : o1;
			}
		}
	}
}

// Occupancy asked of the register allocator: 3 waves per SIMD where the LDS
// holds 3 work-groups per CU (one wave per SIMD each; 160 KiB of LDS per CU),
// otherwise no constraint (LDS already limits residency to 1-2 work-groups).
template <int RT, int IB>
struct TabWaves {
	static constexpr int by_lds = 160 / (IB * RT * 2);
	static constexpr int value = by_lds >= 3 ? 3 : 1;
};

template <int RT, int WPT, int IB>
__global__ __launch_bounds__(256, (TabWaves<RT, IB>::value)) void k_mac_tab(MacArgs a) {
	__shared__ uint4 tab[IB * (RT / 2) * 256];
	const uint64_t g = blockIdx.y;
	const uint64_t i_begin = g * a.in_per_group;
	uint64_t i_end = i_begin + a.in_per_group;
	if (i_end > a.N) i_end = a.N;
	mac_tab_tile<RT, WPT, IB>(a, tab, (uint64_t)blockIdx.x * 256 * WPT, (uint64_t)blockIdx.z * RT,
	                          i_begin, i_end, g, a.partial ? OUT_PARTIAL : OUT_DIRECT);
}

// --------------------------------------------------------------------------
// Batched create: one persistent kernel drains a device queue of tiles that
// span many independent jobs (uneven N, slice size, R).  Each work-group pulls
// the next tile index with one atomic, runs the tile with the launch variant
// the autotuner picked for its job, and loops until the queue is empty (every
// wave of the group sees the same index, so all of them leave together).
// --------------------------------------------------------------------------
template <int RT, int WPT, int IB>
__device__ __forceinline__ void run_batch_tile(const BatchJob& jb, const BatchTile& tl, uint4* tab) {
	MacArgs a{};
	a.in_base = jb.in;
	a.in_stride = jb.in_stride;
	a.out_base = jb.out;
	a.out_stride = jb.out_stride;
	a.coeff = jb.coeff;
	a.coeff_ld = jb.N;
	a.N = jb.N;
	a.R = jb.R;
	a.W = jb.W;
	a.accumulate = 0;
	mac_tab_tile<RT, WPT, IB>(a, tab, tl.w0, tl.r0, tl.i0, tl.i1, 0, tl.mode);
}

template <int WPT>
__device__ __forceinline__ void run_batch_dy(const BatchJob& jb, const BatchTile& tl, uint4* tab) {
	DyadicArgs a{};
	a.in_base = jb.in;
	a.in_stride = jb.in_stride;
	a.out_base = jb.out;
	a.out_stride = jb.out_stride;
	a.gp = jb.gp;
	a.gp_ld = jb.nblocks;
	a.N = jb.N;
	a.R = jb.R;
	a.W = jb.W;
	a.pre_in = jb.pre_in;
	a.pre_out = jb.pre_out;
	a.nblocks = jb.nblocks;
	a.accumulate = 0;
	dy_tile<WPT, true, false, true>(a, tab, nullptr, tl.r0, tl.w0, tl.i0, tl.i1, 0,
	                   tl.mode == OUT_ATOMIC ? DY_OUT_ATOMIC : DY_OUT_DIRECT);
}

// A 16 x 16 XOR-block tile (gf64_dyadic16_dev.h): 256 words, output group r0
// of 16 points, blocks [i0, i1) of 16 points; its 56 KiB of tables fit `tab`.
__device__ __forceinline__ void run_batch_dy16(const BatchJob& jb, const BatchTile& tl, uint4* tab) {
	DyadicArgs a{};
	a.in_base = jb.in;
	a.in_stride = jb.in_stride;
	a.out_base = jb.out;
	a.out_stride = jb.out_stride;
	a.gp = jb.gp;
	a.gp_ld = jb.nblocks;
	a.N = jb.N;
	a.R = jb.R;
	a.W = jb.W;
	a.pre_in = jb.pre_in;
	a.pre_out = jb.pre_out;
	a.nblocks = jb.nblocks;
	a.accumulate = 0;
	d16_tile<1, true, true>(a, tab, nullptr, tl.r0, tl.w0, tl.i0, tl.i1, 0,
	                        tl.mode == OUT_ATOMIC ? DY_OUT_ATOMIC : DY_OUT_DIRECT);
}

// A 32 x 32 XOR-block tile (gf64_dyadic32_dev.h): 512 words, output group r0
// of 32 points, blocks [i0, i1) of 32 points; 112 KiB of tables in `tab`.
__device__ __forceinline__ void run_batch_dy32(const BatchJob& jb, const BatchTile& tl, uint4* tab) {
	DyadicArgs a{};
	a.in_base = jb.in;
	a.in_stride = jb.in_stride;
	a.out_base = jb.out;
	a.out_stride = jb.out_stride;
	a.gp = jb.gp;
	a.gp_ld = jb.nblocks;
	a.N = jb.N;
	a.R = jb.R;
	a.W = jb.W;
	a.pre_in = jb.pre_in;
	a.pre_out = jb.pre_out;
	a.nblocks = jb.nblocks;
	a.accumulate = 0;
	d32_tile<true, true, false>(a, tab, nullptr, tl.r0, tl.w0, tl.i0, tl.i1, 0,
	                            tl.mode == OUT_ATOMIC ? DY_OUT_ATOMIC : DY_OUT_DIRECT);
}

// The 32 x 32 family: one work-group of 512 threads per CU, the same queue
// protocol as k_batch_persistent.
__global__ __launch_bounds__(kD32Threads, 1) void k_batch_persistent32(const BatchJob* __restrict__ jobs,
                                                                   const BatchTile* __restrict__ tiles,
                                                                   uint32_t ntiles, uint32_t* counter) {
	__shared__ uint4 tab[4 * kDyTables * 256];  // four sets of seven pair tables
	__shared__ uint32_t t_next;
	for (;;) {
		if (threadIdx.x == 0) t_next = atomicAdd(counter, 1u);
		__syncthreads();
		const uint32_t t = t_next;
		__syncthreads();
		if (t >= ntiles) break;
		const BatchTile tl = tiles[t];
		run_batch_dy32(jobs[tl.job], tl, tab);
	}
}

// One family of tiles (batch_family): FAM 0 the table variants, 1 the 8 x 8
// XOR-block tiles, 2 the 16 x 16 ones (the 32 x 32 ones: k_batch_persistent32).
template <int FAM>
__global__ __launch_bounds__(256, 2) void k_batch_persistent(const BatchJob* __restrict__ jobs,
                                                          const BatchTile* __restrict__ tiles,
                                                          uint32_t ntiles, uint32_t* counter) {
	__shared__ uint4 tab[4096];  // 64 KiB: the largest table set of the family's variants
	__shared__ uint32_t t_next;
	for (;;) {
		if (threadIdx.x == 0) t_next = atomicAdd(counter, 1u);
		__syncthreads();
		const uint32_t t = t_next;
		__syncthreads();
		if (t >= ntiles) break;
		const BatchTile tl = tiles[t];
		const BatchJob jb = jobs[tl.job];
		if constexpr (FAM == 0) {
			switch (tl.cfg) {
				case BATCH_CFG_8_4_4: run_batch_tile<8, 4, 4>(jb, tl, tab); break;
				case BATCH_CFG_8_2_4: run_batch_tile<8, 2, 4>(jb, tl, tab); break;
				case BATCH_CFG_16_2_2: run_batch_tile<16, 2, 2>(jb, tl, tab); break;
				case BATCH_CFG_4_4_4: run_batch_tile<4, 4, 4>(jb, tl, tab); break;
				default: break;
			}
		} else if constexpr (FAM == 1) {
			if (tl.cfg == BATCH_CFG_DY4)
				run_batch_dy<4>(jb, tl, tab);
			else
				run_batch_dy<2>(jb, tl, tab);
		} else {
			run_batch_dy16(jb, tl, tab);
		}
	}
}

__global__ void k_reset_counters(uint32_t* counters) {
	if (blockIdx.x == 0 && threadIdx.x < kBatchFamilies) counters[threadIdx.x] = 0;
}

// Cauchy matrices of every job of a batch: thread -> (job, r, c) through the
// prefix offsets jobs[j].coeff_off (sorted), binary search over jobs.
__global__ void k_cauchy_batch(const BatchJob* __restrict__ jobs, uint32_t njobs, uint64_t total) {
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		uint32_t lo = 0, hi = njobs - 1;
		while (lo < hi) {
			const uint32_t mid = (lo + hi + 1) >> 1;
			if (jobs[mid].coeff_off <= idx) lo = mid; else hi = mid - 1;
		}
		const BatchJob& jb = jobs[lo];
		const uint64_t e = idx - jb.coeff_off;
		const uint64_t r = e / jb.N, c = e - r * jb.N;
		const uint64_t d = (jb.fi + c) ^ (jb.fr + r);
		jb.coeff[e] = gf64_inv(d ? d : 1);
	}
}

// XOR-block coefficients of every dyadic job of a batch: 8 threads per
// 8-word unit, job found by binary search over jobs[j].gp_off; the same values
// as k_dyadic_coeffs (bits 3: a unit is one (q, block) pair),
// k_dyadic16_coeffs (bits 4: units 2p, 2p+1 are pair p, from an even unit, so
// a pair's 16 threads are 16-aligned in the work-group) and k_dyadic32_coeffs
// (bits 5: units 4p..4p+3 are pair p, from a multiple of 4: 32-aligned).
__global__ __launch_bounds__(256) void k_dyadic_coeffs_batch(const BatchJob* __restrict__ jobs,
                                                             uint32_t njobs, uint64_t total) {
	__shared__ uint64_t gs[256];
	const int u = threadIdx.x & 7;
	for (uint64_t base_unit = (uint64_t)blockIdx.x * 32; base_unit < total;
	     base_unit += (uint64_t)gridDim.x * 32) {
		const uint64_t idx = base_unit + (threadIdx.x >> 3);
		uint64_t c = 0;
		uint32_t lo = 0;
		int bits = 3;
		if (idx < total) {
			uint32_t hi = njobs - 1;
			while (lo < hi) {
				const uint32_t mid = (lo + hi + 1) >> 1;
				if (jobs[mid].gp_off <= idx) lo = mid; else hi = mid - 1;
			}
			const BatchJob& jb = jobs[lo];
			const uint64_t e = idx - jb.gp_off;
			bits = (int)jb.bits;
			uint64_t z;
			if (bits == 5) {
				const uint64_t pr = e >> 2, q = pr / jb.nblocks, b = pr - q * jb.nblocks;
				z = (((jb.fi & ~31ull) + 32 * b) ^ ((jb.fr & ~31ull) + 32 * q)) ^ (uint64_t)(8 * (e & 3) + u);
			} else if (bits == 4) {
				const uint64_t pr = e >> 1, q = pr / jb.nblocks, b = pr - q * jb.nblocks;
				z = (((jb.fi & ~15ull) + 16 * b) ^ ((jb.fr & ~15ull) + 16 * q)) ^ (uint64_t)(8 * (e & 1) + u);
			} else {
				const uint64_t q = e / jb.nblocks, b = e - q * jb.nblocks;
				z = (((jb.fi & ~7ull) + 8 * b) ^ ((jb.fr & ~7ull) + 8 * q)) ^ (uint64_t)u;
			}
			c = z ? gf64_inv(z) : 1u;
		}
		gs[threadIdx.x] = c;
		__syncthreads();
		uint64_t acc = 0;
		if (bits == 5) {
			const int uu = threadIdx.x & 31;
			for (int t = 0; t < 32; t++)
				if ((t & uu) == uu) acc ^= gs[(threadIdx.x & ~31) + t];
		} else if (bits == 4) {
			const int uu = threadIdx.x & 15;
#pragma unroll
			for (int t = 0; t < 16; t++)
				if ((t & uu) == uu) acc ^= gs[(threadIdx.x & ~15) + t];
		} else {
#pragma unroll
			for (int t = 0; t < 8; t++)
				if ((t & u) == u) acc ^= gs[(threadIdx.x & ~7) + t];
		}
		if (idx < total) {
			const BatchJob& jb = jobs[lo];
			jb.gp[(idx - jb.gp_off) * 8 + u] = acc;
		}
		__syncthreads();
	}
}

// Zero the outputs of the jobs that are accumulated with atomics (a kernel,
// not a memset node, so the sequence can be captured in a graph and replayed).
__global__ void k_zero_outputs(const BatchJob* __restrict__ jobs, const uint32_t* __restrict__ ids,
                               uint32_t nids) {
	for (uint32_t k = blockIdx.y; k < nids; k += gridDim.y) {
		const BatchJob& jb = jobs[ids[k]];
		const uint64_t total = jb.R * jb.W;
		for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
		     idx += (uint64_t)gridDim.x * blockDim.x) {
			const uint64_t r = idx / jb.W, w = idx - r * jb.W;
			jb.out[r * jb.out_stride + w] = 0;
		}
	}
}

// Split-K partial reduction.  Pass: dst[h][r][w] = XOR_{g in chunk h} src[g][r][w],
// F partials per chunk, one thread per (h, r, w) -- plenty of parallelism even
// when R*W is small and G is large (many small slices).
__global__ void k_xor_fold(const uint64_t* __restrict__ src, uint64_t G, uint64_t RW, uint64_t F,
                           uint64_t* __restrict__ dst) {
	const uint64_t H = (G + F - 1) / F;
	const uint64_t total = H * RW;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t h = idx / RW, e = idx - h * RW;
		const uint64_t g0 = h * F, g1 = (g0 + F < G) ? g0 + F : G;
		uint64_t s = 0;
		for (uint64_t g = g0; g < g1; g++) s ^= src[g * RW + e];
		dst[idx] = s;
	}
}

// Final: out[r][w] = (accumulate ? out[r][w] : 0) ^ XOR_g partial[g][r][w] (G small)
__global__ void k_xor_reduce(const uint64_t* __restrict__ partial, uint64_t G, uint64_t R,
                             uint64_t W, uint64_t* out_base, uint64_t out_stride,
                             uint64_t* const* out_ptrs, int accumulate) {
	const uint64_t total = R * W;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t r = idx / W, w = idx - r * W;
		uint64_t s = 0;
		for (uint64_t g = 0; g < G; g++) s ^= partial[(g * R + r) * W + w];
		uint64_t* dst = out_ptrs ? out_ptrs[r] : out_base + r * out_stride;
		dst[w] = accumulate ? (dst[w] ^ s) : s;
	}
}

// out[w] = in[w] * coeff[w % n]  (gf64/gf64_region_scalar.c:14-18, cycling coefficients)
__global__ void k_region_mul_arr(uint64_t* __restrict__ out, const uint64_t* __restrict__ in,
                                 const uint64_t* __restrict__ coeff, uint64_t len,
                                 uint64_t n_coeff) {
	for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < len;
	     w += (uint64_t)gridDim.x * blockDim.x)
		out[w] = gf64_mul(in[w], coeff[w % n_coeff]);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
hipError_t launch_cauchy(uint64_t* out, uint64_t N, uint64_t R, uint64_t fi, uint64_t fr,
                         hipStream_t s) {
	uint64_t total = N * R;
	if (total == 0) return hipSuccess;
	uint64_t blocks = (total + 255) / 256;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_cauchy, dim3((unsigned)blocks), dim3(256), 0, s, out, N, R, fi, fr);
	return hipGetLastError();
}

hipError_t launch_cauchy_points(uint64_t* out, const uint64_t* xs, uint64_t nx, const uint64_t* ys,
                                uint64_t ny, hipStream_t s) {
	uint64_t total = nx * ny;
	if (total == 0) return hipSuccess;
	uint64_t blocks = (total + 255) / 256;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_cauchy_points, dim3((unsigned)blocks), dim3(256), 0, s, out, xs, nx, ys,
	                   ny);
	return hipGetLastError();
}

template <int RT, int WPT, int IB>
static hipError_t launch_tab_rt(const MacArgs& a, uint64_t groups, hipStream_t s) {
	const uint64_t tiles = (a.W + 256 * WPT - 1) / (256 * WPT);
	const uint64_t rtiles = (a.R + RT - 1) / RT;
	dim3 grid((unsigned)tiles, (unsigned)groups, (unsigned)rtiles);
	hipLaunchKernelGGL((k_mac_tab<RT, WPT, IB>), grid, dim3(256), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_mac_tab(const MacArgs& a, uint64_t groups, int rt, int wpt, int ib,
                          hipStream_t s) {
#define PARHIP_TAB_CASE(RT_, WPT_, IB_) \
	if (rt == RT_ && wpt == WPT_ && ib == IB_) return launch_tab_rt<RT_, WPT_, IB_>(a, groups, s);
	PARHIP_TAB_CASE(2, 2, 4) PARHIP_TAB_CASE(2, 4, 4)
	PARHIP_TAB_CASE(4, 2, 4) PARHIP_TAB_CASE(4, 4, 4) PARHIP_TAB_CASE(4, 2, 2) PARHIP_TAB_CASE(4, 4, 2)
	PARHIP_TAB_CASE(8, 2, 4) PARHIP_TAB_CASE(8, 4, 4) PARHIP_TAB_CASE(8, 2, 2) PARHIP_TAB_CASE(8, 4, 2)
	PARHIP_TAB_CASE(8, 2, 1) PARHIP_TAB_CASE(8, 4, 1)
	PARHIP_TAB_CASE(16, 2, 2) PARHIP_TAB_CASE(16, 4, 2) PARHIP_TAB_CASE(16, 2, 1)
#undef PARHIP_TAB_CASE
	return hipErrorInvalidValue;
}

bool mac_tab_supported(int rt, int wpt, int ib) {
	static const int cfg[][3] = {{2, 2, 4}, {2, 4, 4}, {4, 2, 4}, {4, 4, 4}, {4, 2, 2}, {4, 4, 2},
	                             {8, 2, 4}, {8, 4, 4}, {8, 2, 2}, {8, 4, 2}, {8, 2, 1}, {8, 4, 1},
	                             {16, 2, 2}, {16, 4, 2}, {16, 2, 1}};
	for (auto& c : cfg)
		if (c[0] == rt && c[1] == wpt && c[2] == ib) return true;
	return false;
}

// Reduce G partials into the outputs.  `partial` holds G x R x W words followed
// by a tail of ceil(G/16) x R x W words (xor_reduce_scratch_words); folding
// passes of 16 ping-pong between the two regions until at most 16 remain, then
// k_xor_reduce writes (or XORs into) the outputs.  Stream order separates the
// passes, so a region is only rewritten after the pass reading it finished.
uint64_t xor_reduce_scratch_words(uint64_t G, uint64_t R, uint64_t W) {
	const uint64_t F = 16;
	return (G + (G > F ? (G + F - 1) / F : 0)) * R * W;
}

hipError_t launch_xor_reduce(uint64_t* partial, uint64_t G, uint64_t R, uint64_t W,
                             uint64_t* out_base, uint64_t out_stride, uint64_t* const* out_ptrs,
                             int accumulate, hipStream_t s) {
	const uint64_t RW = R * W;
	if (RW == 0) return hipSuccess;
	const uint64_t F = 16;
	uint64_t* front = partial;
	uint64_t* tail = partial + G * RW;
	uint64_t* src = front;
	while (G > F) {
		uint64_t* dst = (src == front) ? tail : front;
		const uint64_t H = (G + F - 1) / F;
		uint64_t blocks = (H * RW + 255) / 256;
		if (blocks > 65536) blocks = 65536;
		hipLaunchKernelGGL(k_xor_fold, dim3((unsigned)blocks), dim3(256), 0, s, src, G, RW, F, dst);
		hipError_t e = hipGetLastError();
		if (e != hipSuccess) return e;
		src = dst;
		G = H;
	}
	uint64_t blocks = (RW + 255) / 256;
	if (blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(k_xor_reduce, dim3((unsigned)blocks), dim3(256), 0, s, src, G, R, W,
	                   out_base, out_stride, out_ptrs, accumulate);
	return hipGetLastError();
}

hipError_t launch_batch(const BatchJob* jobs, uint32_t njobs, const BatchTile* tiles,
                        const uint32_t* fam_tiles, uint32_t* counters, const uint32_t* zero_ids,
                        uint32_t nzero, uint64_t coeff_total, uint64_t dy_pairs, int blocks,
                        hipStream_t s) {
	if (coeff_total) {
		uint64_t cb = (coeff_total + 255) / 256;
		if (cb > 65536) cb = 65536;
		hipLaunchKernelGGL(k_cauchy_batch, dim3((unsigned)cb), dim3(256), 0, s, jobs, njobs,
		                   coeff_total);
	}
	if (dy_pairs) {
		uint64_t cb = (dy_pairs + 31) / 32;
		if (cb > 65536) cb = 65536;
		hipLaunchKernelGGL(k_dyadic_coeffs_batch, dim3((unsigned)cb), dim3(256), 0, s, jobs, njobs,
		                   dy_pairs);
	}
	if (nzero) {
		hipLaunchKernelGGL(k_zero_outputs, dim3(64, nzero < 1024 ? nzero : 1024), dim3(256), 0, s,
		                   jobs, zero_ids, nzero);
	}
	uint32_t ntiles = 0;
	for (int f = 0; f < kBatchFamilies && fam_tiles; f++) ntiles += fam_tiles[f];
	if (ntiles && blocks > 0) {
		hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters);
		uint32_t off = 0;
		for (int f = 0; f < kBatchFamilies; f++) {
			const uint32_t n = fam_tiles[f];
			if (!n) continue;
			if (f == 3) {
				const int b32 = blocks > 1 ? blocks / 2 : 1;
				hipLaunchKernelGGL(k_batch_persistent32, dim3((unsigned)b32), dim3(kD32Threads), 0, s, jobs,
				                   tiles + off, n, counters + f);
			} else {
				void (*k)(const BatchJob*, const BatchTile*, uint32_t, uint32_t*) =
				    f == 0 ? k_batch_persistent<0> : (f == 1 ? k_batch_persistent<1> : k_batch_persistent<2>);
				hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(256), 0, s, jobs, tiles + off, n, counters + f);
			}
			off += n;
		}
	}
	return hipGetLastError();
}

hipError_t launch_region_mul_arr(uint64_t* out, const uint64_t* in, const uint64_t* coeff,
                                 uint64_t len, uint64_t n_coeff, hipStream_t s) {
	if (len == 0) return hipSuccess;
	uint64_t blocks = (len + 255) / 256;
	if (blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(k_region_mul_arr, dim3((unsigned)blocks), dim3(256), 0, s, out, in, coeff,
	                   len, n_coeff);
	return hipGetLastError();
}

}  // namespace parhip
