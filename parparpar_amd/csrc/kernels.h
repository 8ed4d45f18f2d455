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

hipError_t launch_cauchy(uint64_t* out, uint64_t N, uint64_t R, uint64_t fi, uint64_t fr,
                         hipStream_t s);
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

hipError_t launch_gf16_coeffs(uint16_t* coeff, const uint16_t* input_nums, uint64_t N,
                              const uint16_t* exps, uint64_t R, const uint16_t* exp_tab,
                              const uint16_t* inlog_tab, hipStream_t s);
hipError_t launch_gf16_tab(const Gf16Args& a, uint64_t groups, int ng, hipStream_t s);
uint64_t gf16_reduce_scratch_words(uint64_t G, uint64_t R, uint64_t W);
hipError_t launch_gf16_reduce(uint16_t* partial, uint64_t G, uint64_t R, uint64_t W,
                              uint16_t* out_base, uint64_t out_stride, int accumulate,
                              hipStream_t s);

}  // namespace parhip
