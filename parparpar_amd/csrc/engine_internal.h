// engine_internal.h -- shared host-side helpers of the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace parhip {

struct Plan {
	int method;            // PARHIP_METHOD_TABLE / PARHIP_METHOD_MFMA
	int rt, wpt, ib;       // outputs per work-group, words per thread, inputs per LDS batch
	uint64_t groups;       // split-K work-group rows
	uint64_t in_per_group; // inputs per row
};

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int current_device();
int current_method();
// Split-K group count forced by parhip_tune_override (0 = automatic).
int64_t tune_groups_override();
// Tile shape forced by parhip_tune_override (0 fields = automatic).
void tune_shape_override(int* rt, int* wpt, int* ib);
int device_cus(int dev);
int host_stream(hipStream_t* out);
int dev_alloc(void** p, size_t bytes, hipStream_t s);
void dev_free(void* p, hipStream_t s);
int coeff_acquire(uint64_t n, uint64_t r, uint64_t fi, uint64_t fr, hipStream_t s,
                  const uint64_t** out);
// s: the stream of the call (coefficients built for that call alone are freed on it)
void coeff_release(const uint64_t* d, hipStream_t s);
void coeff_cache_clear();
Plan plan_contraction(uint64_t N, uint64_t R, uint64_t W, int dev);
size_t plan_workspace_bytes(const Plan& p, uint64_t R, uint64_t W);
int run_contraction(MacArgs a, hipStream_t s);
// The nibble-table strategy whatever this thread's parhip_set_method says.
int run_contraction_tables(MacArgs a, hipStream_t s);

// The GF(2) bit-matrix strategy on MFMA: launch shape, split-K groups and the
// workspace it needs (bit-matrix rows, then split-K partials), and the launch
// sequence on a caller-provided workspace.  `overrides`: honour this thread's
// parhip_tune_override.
struct MfmaPlan {
	bool fp4;
	int rt, nb, ib;
	uint64_t groups, per;
	size_t rows_bytes, part_bytes;
};
int plan_mfma(uint64_t N, uint64_t R, uint64_t W, bool fp4, bool overrides, MfmaPlan* p);
hipError_t enqueue_mfma(const MacArgs& m, const MfmaPlan& p, void* ws, hipStream_t s);

// The XOR-block (dyadic) form of the plain Cauchy contraction (gf64_dyadic.hip).
struct DyadicPlan {
	// bits 3: 8 x 8 blocks (k_mac_dyadic); 4: 16 x 16 blocks (k_mac_dyadic16);
	// 5: 32 x 32 blocks (k_mac_dyadic32, tiles of 512 words).
	// pre_in, pre_out, nblocks, Q count in runs of 2^bits points.
	int bits = 3;
	uint64_t pre_in, pre_out, nblocks, Q;
	int wpt, ib;
	uint64_t groups, blocks_per_group;
	size_t part_bytes;  // split-K partials + fold tail (0 when groups == 1)
};
// max_part_bytes: the largest split-K partial buffer the caller can give
// (fewer split-K rows below it).
// Whether 16 x 16 XOR blocks need fewer table reads than 8 x 8 ones for R
// outputs from point fr (or this thread's method forces either).
bool dyadic16_better(uint64_t R, uint64_t fr);
// The block size (log2 points: 3, 4 or 5, at most max_bits) that needs the
// fewest table reads for R outputs from point fr, or the one this thread's
// method forces.
int dyadic_block_bits(uint64_t R, uint64_t fr, int max_bits);
// allow16: the caller can run the 16 x 16 form (coefficients computed in the
// kernel or in the 16-word layout); allow32 (with allow16): also the 32 x 32
// form (coefficients in the kernel or in the 32-word layout); the plan takes
// the block size that needs the fewest table reads.
// force_bits 3 / 4 / 5: that block size whatever the rule says (workspace
// bounds).
DyadicPlan plan_dyadic(uint64_t N, uint64_t R, uint64_t W, uint64_t fi, uint64_t fr, int dev,
                       size_t max_part_bytes = SIZE_MAX, bool allow16 = false, int force_bits = 0,
                       bool allow32 = false);
// Whether `a` (strided slices, this thread's method) takes the dyadic kernel.
bool dyadic_applies(const MacArgs& a);
// Whether R outputs take it under this thread's method (any input layout);
// dyadic_shape_ok: under the automatic choice.
bool dyadic_applies_shape(uint64_t R);
bool dyadic_shape_ok(uint64_t R);
// gp: the coefficients of the launch's first block; gp_ld: blocks per output
// group row of the coefficient array; lost: optional per-block masks of inputs
// that count as zero (DyadicArgs::lost).
// gp null: the kernel computes the coefficients from the points xbase + 8b,
// ybase + 8q (both multiples of 8).
hipError_t enqueue_dyadic(const MacArgs& a, const DyadicPlan& p, const uint64_t* gp,
                          uint64_t gp_ld, void* ws, hipStream_t s, const uint32_t* lost = nullptr,
                          uint64_t xbase = 0, uint64_t ybase = 0);
// The plain Cauchy contraction of a call with n_total inputs at points
// (fi + i, fr + r), run over its inputs [first, first + a.N): the dyadic kernel
// where it applies, else the cached coefficient matrix and run_contraction.
// a.coeff is ignored; n_total = 0 means a.N.
int run_cauchy_contraction(MacArgs a, uint64_t fi, uint64_t fr, hipStream_t s, uint64_t first = 0,
                           uint64_t n_total = 0);

// PAR2 GF(2^16) side (engine_gf16.cpp)
int gf16_device_log_exp(const uint16_t** exp_tab, const uint16_t** log_tab);
int gf16_coeffs_device(uint16_t* d_coeff, const uint16_t* d_input_nums, uint64_t N,
                       const uint16_t* d_exps, uint64_t R, hipStream_t s);
int run_gf16_contraction(Gf16Args a, hipStream_t s);

// Switch this thread's kernel timing on or off; returns the previous state
// (launches recorded into a graph are not timed).
bool prof_set(bool on);

// Brackets a kernel launch with HIP events when timing is enabled
// (parhip_prof_enable); no-op otherwise.
class ProfScope {
public:
	ProfScope(const char* name, hipStream_t s);
	~ProfScope();

private:
	hipStream_t s_;
	const char* name_ = nullptr;
	hipEvent_t a_ = nullptr, b_ = nullptr;
	bool on_ = false;
};

// RAII device buffer on a stream (stream-ordered allocation).
struct DevBuf {
	void* p = nullptr;
	hipStream_t s = nullptr;
	DevBuf() = default;
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	int alloc(size_t bytes, hipStream_t st) {
		s = st;
		return dev_alloc(&p, bytes, st);
	}
	~DevBuf() { dev_free(p, s); }
	template <class T>
	T* as() const {
		return (T*)p;
	}
};

}  // namespace parhip
