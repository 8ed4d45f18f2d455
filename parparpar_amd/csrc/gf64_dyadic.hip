// gf64_dyadic.hip -- the GF(2^64) Cauchy contraction through the XOR-block
// structure of its coefficient matrix (gfx950).
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

// Loads of one batch of IB blocks: this thread's words of the 8 inputs of each
// block (zeros for virtual inputs, past the group's blocks or past the slice
// end) and the two coefficients of the table entries it builds.
template <int WPT, int IB>
__device__ __forceinline__ void dy_load_batch(const DyadicArgs& a, uint64_t q, uint64_t bb,
                                              uint64_t b_end, uint64_t w0, bool fast,
                                              uint2 (&d)[IB][8][WPT], uint64_t& c0, uint64_t& c1) {
	const int tid = threadIdx.x;
	c0 = 0;
	c1 = 0;
	if (tid < IB * kDyTables * 16) {
		const int ib = tid / (kDyTables * 16);
		const int tbl = (tid - ib * kDyTables * 16) >> 4;
		const uint64_t b = bb + ib;
		if (b < b_end) {
			const uint64_t* gp = a.gp + (q * a.gp_ld + b) * 8;
			c0 = dy_load(gp + kDySA[tbl]);
			c1 = dy_load(gp + kDySB[tbl]);
		}
	}
#pragma unroll
	for (int ib = 0; ib < IB; ib++) {
		const uint64_t b = bb + ib;
#pragma unroll
		for (int t = 0; t < 8; t++) {
			// real input of virtual slot t of block b
			const uint64_t v = 8 * b + t;
			const bool valid = b < b_end && v >= a.pre_in && v - a.pre_in < a.N;
			const uint64_t i = v - a.pre_in;
#pragma unroll
			for (int jj = 0; jj < WPT / 2; jj++) {
				const uint64_t w = w0 + 2 * tid + 512 * jj;
				uint64_t v0 = 0, v1 = 0;
				if (valid) {
					const uint64_t* src = a.in_base + i * a.in_stride;
					if (fast) {
						const ulonglong2 x = dy_load2(src + w);
						v0 = x.x;
						v1 = x.y;
					} else if (w + 1 < a.W && ((reinterpret_cast<uintptr_t>(src) & 15u) == 0)) {
						const ulonglong2 x = dy_load2(src + w);
						v0 = x.x;
						v1 = x.y;
					} else {
						if (w < a.W) v0 = dy_load(src + w);
						if (w + 1 < a.W) v1 = dy_load(src + w + 1);
					}
				}
				d[ib][t][2 * jj] = make_uint2((uint32_t)v0, (uint32_t)(v0 >> 32));
				d[ib][t][2 * jj + 1] = make_uint2((uint32_t)v1, (uint32_t)(v1 >> 32));
			}
		}
	}
}

// Work-group: words [w0, w0 + 256*WPT) of output group q (virtual recovery
// points 8q..8q+7), input blocks [b_begin, b_end) of split-K row g.
template <int WPT, int IB>
__global__ __launch_bounds__(256, 2) void k_mac_dyadic(DyadicArgs a) {
	static_assert(WPT == 2 || WPT == 4, "WPT");
	static_assert(IB * kDyTables * 16 <= 256, "one table column per thread");
	// tab[((ib*7 + tbl)*16 + k)*16 + n] = (g'_SA * n x^{4k}, g'_SB * n x^{4k})
	__shared__ uint4 tab[IB * kDyTables * 256];
#define DYTAB(ib, tbl, k, n) tab[(((ib) * kDyTables + (tbl)) * 16 + (k)) * 16 + (n)]
	const int tid = threadIdx.x;
	const uint64_t q = blockIdx.z;
	const uint64_t g = blockIdx.y;
	const uint64_t w0 = (uint64_t)blockIdx.x * 256 * WPT;
	const uint64_t b_begin = g * a.blocks_per_group;
	uint64_t b_end = b_begin + a.blocks_per_group;
	if (b_end > a.nblocks) b_end = a.nblocks;
	const bool fast = ((reinterpret_cast<uintptr_t>(a.in_base) & 15u) == 0) && (a.in_stride & 1) == 0 &&
	                  w0 + 256ull * WPT <= a.W;

	uint2 acc[WPT][8];
#pragma unroll
	for (int j = 0; j < WPT; j++)
#pragma unroll
		for (int u = 0; u < 8; u++) acc[j][u] = make_uint2(0, 0);

	uint2 d[IB][8][WPT];
	uint64_t c0, c1;
	dy_load_batch<WPT, IB>(a, q, b_begin, b_end, w0, fast, d, c0, c1);
	for (uint64_t bb = b_begin; bb < b_end; bb += IB) {
		// ---- tables: thread (ib, tbl, n) owns entry n of all 16 positions of
		// one pair table; entry(k, n) = x^4 * entry(k-1, n).  The 16 lanes of
		// a table write 16 consecutive entries of one row per store.
		__syncthreads();
		if (tid < IB * kDyTables * 16) {
			const int n = tid & 15;
			const int ib = tid / (kDyTables * 16);
			const int tbl = (tid - ib * kDyTables * 16) >> 4;
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
				DYTAB(ib, tbl, k, n) = v;
			}
		}
		__syncthreads();
		// ---- next batch's loads, in flight during this batch's lookups
		uint2 dn[IB][8][WPT];
		uint64_t n0, n1;
		dy_load_batch<WPT, IB>(a, q, bb + IB, b_end, w0, fast, dn, n0, n1);
		// ---- lookups
#pragma unroll
		for (int ib = 0; ib < IB; ib++) {
#pragma unroll
			for (int j = 0; j < WPT; j++) {
				uint2 x[8];
#pragma unroll
				for (int t = 0; t < 8; t++) x[t] = d[ib][t][j];
				dy_superset(x);
#pragma unroll
				for (int T = 0; T < 8; T++) {
#pragma unroll
					for (int kq = 0; kq < 16; kq += 4) {
						unsigned nib[4];
#pragma unroll
						for (int u = 0; u < 4; u++) {
							const int k = kq + u;
							const uint32_t half = k < 8 ? x[T].x : x[T].y;
							nib[u] = (half >> (4 * (k & 7))) & 15u;
						}
						if (T == 7) {
							// slot a of table Q0 (g'_0): 8-byte reads
							const uint2* t2 = reinterpret_cast<const uint2*>(tab);
							const uint2 v0 = t2[2 * (((ib * kDyTables + 4) * 16 + kq + 0) * 16 + nib[0])];
							const uint2 v1 = t2[2 * (((ib * kDyTables + 4) * 16 + kq + 1) * 16 + nib[1])];
							const uint2 v2 = t2[2 * (((ib * kDyTables + 4) * 16 + kq + 2) * 16 + nib[2])];
							const uint2 v3 = t2[2 * (((ib * kDyTables + 4) * 16 + kq + 3) * 16 + nib[3])];
							uint2 c = acc[j][7];
							c.x = dy_xor3(dy_xor3(c.x, v0.x, v1.x), v2.x, v3.x);
							c.y = dy_xor3(dy_xor3(c.y, v0.y, v1.y), v2.y, v3.y);
							acc[j][7] = c;
						} else {
#pragma unroll
							for (int uu = 0; uu < 4; uu++) {
								if (uu >= kDyNT[T]) break;
								const int tbl = kDyTab[T][uu];
								const int ua = kDySA[tbl] | T, ub = kDySB[tbl] | T;
								const uint4 v0 = DYTAB(ib, tbl, kq + 0, nib[0]);
								const uint4 v1 = DYTAB(ib, tbl, kq + 1, nib[1]);
								const uint4 v2 = DYTAB(ib, tbl, kq + 2, nib[2]);
								const uint4 v3 = DYTAB(ib, tbl, kq + 3, nib[3]);
								uint2 ca = acc[j][ua], cb = acc[j][ub];
								ca.x = dy_xor3(dy_xor3(ca.x, v0.x, v1.x), v2.x, v3.x);
								ca.y = dy_xor3(dy_xor3(ca.y, v0.y, v1.y), v2.y, v3.y);
								cb.x = dy_xor3(dy_xor3(cb.x, v0.z, v1.z), v2.z, v3.z);
								cb.y = dy_xor3(dy_xor3(cb.y, v0.w, v1.w), v2.w, v3.w);
								acc[j][ua] = ca;
								acc[j][ub] = cb;
							}
						}
					}
				}
			}
		}
#pragma unroll
		for (int ib = 0; ib < IB; ib++)
#pragma unroll
			for (int t = 0; t < 8; t++)
#pragma unroll
				for (int j = 0; j < WPT; j++) d[ib][t][j] = dn[ib][t][j];
		c0 = n0;
		c1 = n1;
	}
#undef DYTAB

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
		if (a.partial) {
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
			if (w + 1 < a.W && ((reinterpret_cast<uintptr_t>(dst + w) & 15u) == 0)) {
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

// Transformed coefficients of every (output group q, input block b):
//   g(u) = f(X_b XOR Y_q XOR u), X_b = xbase + 8b, Y_q = ybase + 8q,
//   out[(q*nblocks + b)*8 + S] = XOR over supersets T of S of g(T).
__global__ void k_dyadic_coeffs(uint64_t* __restrict__ out, uint64_t nblocks, uint64_t Q,
                                uint64_t xbase, uint64_t ybase) {
	const uint64_t total = nblocks * Q;
	for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t q = idx / nblocks, b = idx - q * nblocks;
		const uint64_t base = (xbase + 8 * b) ^ (ybase + 8 * q);
		uint2 x[8];
#pragma unroll
		for (int u = 0; u < 8; u++) {
			const uint64_t z = base ^ (uint64_t)u;
			const uint64_t c = z ? gf64_inv(z) : 1u;
			x[u] = make_uint2((uint32_t)c, (uint32_t)(c >> 32));
		}
		dy_superset(x);
#pragma unroll
		for (int u = 0; u < 8; u++) out[idx * 8 + u] = ((uint64_t)x[u].y << 32) | x[u].x;
	}
}

hipError_t launch_dyadic_coeffs(uint64_t* out, uint64_t nblocks, uint64_t Q, uint64_t xbase,
                                uint64_t ybase, hipStream_t s) {
	const uint64_t total = nblocks * Q;
	if (!total) return hipSuccess;
	uint64_t blocks = (total + 255) / 256;
	if (blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_dyadic_coeffs, dim3((unsigned)blocks), dim3(256), 0, s, out, nblocks, Q,
	                   xbase, ybase);
	return hipGetLastError();
}

bool mac_dyadic_supported(int wpt, int ib) {
	return (wpt == 4 && ib == 1) || (wpt == 2 && (ib == 1 || ib == 2));
}

hipError_t launch_mac_dyadic(const DyadicArgs& a, uint64_t groups, int wpt, int ib, hipStream_t s) {
	const uint64_t Q = (a.pre_out + a.R + 7) / 8;
	const uint64_t tiles = (a.W + 256ull * wpt - 1) / (256ull * wpt);
	if (!tiles || !groups || !Q) return hipSuccess;
	if (tiles > 0x7fffffffu || groups > 65535 || Q > 65535) return hipErrorInvalidValue;
	dim3 grid((unsigned)tiles, (unsigned)groups, (unsigned)Q);
	if (wpt == 4 && ib == 1)
		hipLaunchKernelGGL((k_mac_dyadic<4, 1>), grid, dim3(256), 0, s, a);
	else if (wpt == 2 && ib == 1)
		hipLaunchKernelGGL((k_mac_dyadic<2, 1>), grid, dim3(256), 0, s, a);
	else if (wpt == 2 && ib == 2)
		hipLaunchKernelGGL((k_mac_dyadic<2, 2>), grid, dim3(256), 0, s, a);
	else
		return hipErrorInvalidValue;
	return hipGetLastError();
}

}  // namespace parhip
