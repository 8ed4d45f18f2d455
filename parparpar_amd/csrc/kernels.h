// kernels.h -- launch interface between the C-ABI host code and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace parhip {

// One contraction R[r][w] (^)= XOR_{i in [0,N)} coeff[r*coeff_ld + i] * in_i[w].
// Inputs and outputs are either strided views of one allocation or arrays of
// device pointers (scattered slices, e.g. the known blocks of a repair).
struct MacArgs {
	const uint64_t* in_base;
	uint64_t in_stride;               // words between consecutive inputs
	const uint64_t* const* in_ptrs;   // if non-null, overrides in_base/in_stride
	uint64_t* out_base;
	uint64_t out_stride;
	uint64_t* const* out_ptrs;        // if non-null, overrides out_base/out_stride
	const uint64_t* coeff;            // row-major, coeff[r*coeff_ld + i]
	uint64_t coeff_ld;
	uint64_t N, R, W;                 // inputs, outputs, words per slice
	uint64_t in_per_group;            // split-K: inputs per work-group row
	uint64_t* partial;                // if non-null: G x R x W partials, reduced afterwards
	int accumulate;                   // XOR into existing outputs (when partial == nullptr)
};

// Batched create (persistent kernel over a device tile queue).
enum BatchCfg : uint16_t {
	BATCH_CFG_8_4_4 = 0,   // RT=8,  WPT=4, IB=4
	BATCH_CFG_8_2_4 = 1,   // RT=8,  WPT=2, IB=4
	BATCH_CFG_MFMA = 2,    // GF(2) bit-matrix on MFMA (fp4): job-level launches, no tiles
	BATCH_CFG_16_2_2 = 3,  // RT=16, WPT=2, IB=2
	BATCH_CFG_4_4_4 = 4,   // RT=4,  WPT=4, IB=4
	BATCH_CFG_DY4 = 5,     // XOR-block (dyadic) tiles, WPT=4 (full 1024-word tiles)
	BATCH_CFG_DY2 = 6,     // XOR-block (dyadic) tiles, WPT=2 (full 512-word tiles)
	BATCH_CFG_DY16 = 7,    // 16 x 16 XOR-block tiles, WPT=1 (full 256-word tiles)
	BATCH_CFG_DY32 = 8,    // 32 x 32 XOR-block tiles, 512 threads x 1 word (full 512-word tiles)
	BATCH_CFG_COUNT = 9
};
struct BatchJob {
	const uint64_t* in;
	uint64_t in_stride;
	uint64_t* out;
	uint64_t out_stride;
	uint64_t* coeff;      // R x N, built by the batch's Cauchy kernel (table / MFMA jobs)
	uint64_t N, R, W;
	uint64_t fi, fr;
	uint64_t coeff_off;   // first global index of this job's coefficients (0 of them for dyadic jobs)
	// dyadic jobs: Q x nblocks x 2^bits XOR-block coefficients (k_dyadic_coeffs
	// layout; bits 4: k_dyadic16_coeffs layout; bits 5: k_dyadic32_coeffs
	// layout); nblocks, Q, pre_in, pre_out count in runs of 2^bits points
	uint64_t* gp;
	uint64_t nblocks, Q, pre_in, pre_out;
	// first global 8-word unit of this job's coefficients (Q*nblocks units,
	// 2*Q*nblocks from an even unit for bits 4, 4*Q*nblocks from a multiple
	// of 4 for bits 5; 0 for other jobs)
	uint64_t gp_off;
	uint64_t bits;        // 3: 8 x 8 blocks, 4: 16 x 16 blocks, 5: 32 x 32 blocks
};
// Table tiles: outputs [r0, r0+RT), inputs [i0, i1).  Dyadic tiles: output
// group r0, input blocks [i0, i1) (of 8, 16 or 32 points).  Words [w0, w0 + 256*WPT)
// (32 x 32 tiles: [w0, w0 + 512)).
struct BatchTile {
	uint32_t job;
	uint16_t cfg;         // BatchCfg
	uint16_t mode;        // OUT_DIRECT (0) or OUT_ATOMIC (2)
	uint64_t r0;
	uint64_t w0, i0, i1;
};

// The batch's launch sequence: coefficient matrices (coeff_total words over
// the table / MFMA jobs), XOR-block coefficients (dy_pairs 8-word units over
// the dyadic jobs), zeroing of the atomically accumulated jobs' outputs,
// queue reset, persistent kernel.
// Tiles come grouped by family (table tiles, 8 x 8 XOR-block tiles, 16 x 16
// XOR-block tiles, 32 x 32 XOR-block tiles; fam_tiles[f] of each, in that
// order), and each family runs as its own persistent kernel over its own queue
// counter (counters[f]): one kernel holding every variant would size its
// registers for the largest and spill in the others.  `blocks` is the
// persistent grid of the 256-thread families (two work-groups per CU); the
// 32 x 32 family runs half as many work-groups of 512 threads (one per CU,
// 112 KiB of LDS each).
constexpr int kBatchFamilies = 4;
inline int batch_family(int cfg) {
	if (cfg == BATCH_CFG_DY32) return 3;
	return cfg == BATCH_CFG_DY16 ? 2 : ((cfg == BATCH_CFG_DY4 || cfg == BATCH_CFG_DY2) ? 1 : 0);
}
hipError_t launch_batch(const BatchJob* jobs, uint32_t njobs, const BatchTile* tiles,
                        const uint32_t* fam_tiles, uint32_t* counters, const uint32_t* zero_ids,
                        uint32_t nzero, uint64_t coeff_total, uint64_t dy_pairs, int blocks,
                        hipStream_t s);

hipError_t launch_cauchy(uint64_t* out, uint64_t N, uint64_t R, uint64_t fi, uint64_t fr,
                         hipStream_t s);

// Cauchy contraction through its XOR-block structure (gf64_dyadic.hip):
// virtual input points are firstInput - pre_in + v, v = 8*block + t; virtual
// recovery points firstRecovery - pre_out + 8*q + s.  gp holds, per (q, block),
// the 8 coefficients of the block in the subset basis (k_dyadic_coeffs).
struct DyadicArgs {
	const uint64_t* in_base;
	uint64_t in_stride;        // words between consecutive inputs
	uint64_t* out_base;
	uint64_t out_stride;
	const uint64_t* gp;        // coefficients of (q, block b): gp + (q*gp_ld + b)*8
	uint64_t gp_ld;
	uint64_t N, R, W;
	uint64_t pre_in, pre_out;  // firstInput & 7, firstRecovery & 7
	uint64_t nblocks;          // (pre_in + N + 7) / 8
	uint64_t blocks_per_group; // split-K: blocks per work-group row
	uint64_t* partial;         // if non-null: G x R x W partials, reduced afterwards
	int accumulate;
	// if non-null: bit t of lost[b] set = the input at virtual slot 8b+t
	// counts as zero whatever its memory holds (the known inputs of a repair)
	const uint32_t* lost;
	// gp null: the kernel computes the coefficients itself from the points,
	// X_b = xbase + 8b and Y_q = ybase + 8q (multiples of 8)
	uint64_t xbase, ybase;
	// split-K rows XOR their results into the outputs with atomics (zeroed
	// first unless accumulating) instead of writing partials
	int atomic_out;
};
hipError_t launch_zero_rows(uint64_t* out, uint64_t rows, uint64_t words, uint64_t stride,
                            hipStream_t s);
hipError_t launch_dyadic_coeffs(uint64_t* out, uint64_t nblocks, uint64_t Q, uint64_t xbase,
                                uint64_t ybase, hipStream_t s);
bool mac_dyadic_supported(int wpt, int ib);
hipError_t launch_mac_dyadic(const DyadicArgs& a, uint64_t groups, int wpt, int ib, hipStream_t s);
// The 16 x 16 form: blocks and output groups of 16 points, gp with 16 words
// per (q, b), lost masks one u32 per block of 16 (with gp only); wpt 1 or 2.
hipError_t launch_dyadic16_coeffs(uint64_t* out, uint64_t nblocks, uint64_t Q, uint64_t xbase,
                                  uint64_t ybase, hipStream_t s);
hipError_t launch_mac_dyadic16(const DyadicArgs& a, uint64_t groups, int wpt, hipStream_t s);
// The 32 x 32 form: blocks and output groups of 32 points, tiles of 512 words
// (one word per thread, 512 threads), gp with 32 words per (q, b), lost masks
// one u32 per block of 32 (with gp only).
hipError_t launch_dyadic32_coeffs(uint64_t* out, uint64_t nblocks, uint64_t Q, uint64_t xbase,
                                  uint64_t ybase, hipStream_t s);
hipError_t launch_mac_dyadic32(const DyadicArgs& a, uint64_t groups, hipStream_t s);
hipError_t launch_cauchy_points(uint64_t* out, const uint64_t* xs, uint64_t nx,
                                const uint64_t* ys, uint64_t ny, hipStream_t s);
hipError_t launch_mac_tab(const MacArgs& a, uint64_t groups, int rt, int wpt, int ib,
                          hipStream_t s);
bool mac_tab_supported(int rt, int wpt, int ib);
uint64_t xor_reduce_scratch_words(uint64_t G, uint64_t R, uint64_t W);
hipError_t launch_xor_reduce(uint64_t* partial, uint64_t G, uint64_t R, uint64_t W,
                             uint64_t* out_base, uint64_t out_stride, uint64_t* const* out_ptrs,
                             int accumulate, hipStream_t s);
hipError_t launch_region_mul_arr(uint64_t* out, const uint64_t* in, const uint64_t* coeff,
                                 uint64_t len, uint64_t n_coeff, hipStream_t s);

// PAR2 GF(2^16) contraction: out[e][w] (^)= XOR_i coeff[e*coeff_ld + i] * in_i[w]
// over 16-bit words (the last word of each staged slice may be a checksum word).
struct Gf16Args {
	const uint16_t* in_base;
	uint64_t in_stride;  // words
	uint16_t* out_base;
	uint64_t out_stride; // words
	const uint16_t* coeff;
	uint64_t coeff_ld;
	uint64_t N, R, W;
	uint64_t in_per_group;
	uint16_t* partial;
	int accumulate;
};

// GF(2^16) contraction as a GF(2) bit-matrix on MFMA (gf16_mfma.hip).
struct Gf16MfmaArgs {
	const uint16_t* in_base;
	uint64_t in_stride;   // words
	uint16_t* out_base;
	uint64_t out_stride;  // words
	const uint32_t* rows; // k_gf16_rows: (R x ceil(N/2) x 16) row words
	uint64_t N, R, W;
	uint64_t in_per_group;  // multiple of 4
	uint16_t* partial;      // if non-null: G x R x W partials
	int accumulate;
};
hipError_t launch_gf16_rows(const uint16_t* coeff, uint64_t ld, uint64_t N, uint64_t R, uint32_t* rows,
                            hipStream_t s);
bool gf16_mfma_supported(int ot, int nb, bool fp4);
hipError_t launch_gf16_mfma(const Gf16MfmaArgs& a, uint64_t groups, int ot, int nb, int ib, bool fp4,
                            hipStream_t s);

hipError_t launch_gf16_coeffs(uint16_t* coeff, const uint16_t* input_nums, uint64_t N,
                              const uint16_t* exps, uint64_t R, const uint16_t* exp_tab,
                              const uint16_t* inlog_tab, hipStream_t s);
hipError_t launch_gf16_tab(const Gf16Args& a, uint64_t groups, int ng, hipStream_t s);
uint64_t gf16_reduce_scratch_words(uint64_t G, uint64_t R, uint64_t W);
hipError_t launch_gf16_reduce(uint16_t* partial, uint64_t G, uint64_t R, uint64_t W,
                              uint16_t* out_base, uint64_t out_stride, int accumulate,
                              hipStream_t s);

// GF(2) bit-matrix contraction on MFMA.
struct MfmaArgs {
	const uint64_t* in_base;
	uint64_t in_stride;
	uint64_t* out_base;
	uint64_t out_stride;
	const uint64_t* rows;  // (r*rows_ld + i)*64 + o : Row_o of coefficient (r, i)
	uint64_t rows_ld;
	uint64_t N, R, W;
	uint64_t in_per_group;
	uint64_t* partial;
	int accumulate;
};
// 8-bit split tables (gf64_tab8.hip): RT in {2, 4}, WPT in {2, 4, 8}, IB in {1, 2}
// (IB = 1 with WPT = 8).
bool mac_tab8_supported(int rt, int wpt, int ib);
hipError_t launch_mac_tab8(const MacArgs& a, uint64_t groups, int rt, int wpt, int ib, hipStream_t s);

hipError_t launch_gf64_rows(const uint64_t* coeff, uint64_t ld, uint64_t N, uint64_t R,
                            uint64_t* rows, hipStream_t s);
hipError_t launch_mac_mfma(const MfmaArgs& a, uint64_t groups, int rt, int nb, int ib, bool fp4,
                           hipStream_t s);
bool mfma_supported(int rt, int nb, int ib, bool fp4);

// PAR2 repair matrix (gf16_inv.hip): in-place Gauss-Jordan of an n x n
// block over GF(2^16), one pivot per launch; *bad = first zero pivot or -1.
hipError_t launch_gf16_invert(uint16_t* a, uint16_t* b, uint32_t n, const uint16_t* exp_tab,
                              const uint16_t* log_tab, int* bad, uint16_t** result, hipStream_t s);

// PAR3 recovery packets (packets.hip): ws holds 2 x num x chunks x 8 words.
uint64_t b3_chunks_of(uint64_t block_size);
hipError_t launch_rec_packets(const uint8_t* rec, uint64_t isid, uint64_t first, uint64_t num,
                              uint64_t bs, uint8_t* out, uint32_t* ws, hipStream_t s);

// Repair (closed-form Cauchy inverse and helpers).
hipError_t launch_cauchy_inverse(const uint64_t* ys, const uint64_t* xs, uint64_t k,
                                 uint64_t* alpha, uint64_t* beta, uint64_t* inv, hipStream_t s);
hipError_t launch_repair_coeffs(uint64_t* out, uint64_t ld, const uint64_t* xs_known, uint64_t nk,
                                const uint64_t* ys, uint64_t k, hipStream_t s);
hipError_t launch_slice_ptrs(const uint64_t* base, uint64_t stride, const uint64_t* idx, uint64_t n,
                             const uint64_t** out, hipStream_t s);

}  // namespace parhip
