// engine_gf16.cpp -- PAR2 GF(2^16) side of the C ABI: coefficient tables on the
// device, the launch plan of the contraction, and the device entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "gf16_host.h"
#include "kernels.h"

namespace parhip {

// ------------------------------------------------------------------ tables
// exp table of generator 2 (65536 entries, exp[65535] = exp[0]) and the PAR2
// input-constant logarithms (32768 entries), uploaded once per device.
struct Gf16DevTables {
	uint16_t* exp_tab = nullptr;
	uint16_t* inlog = nullptr;
	uint16_t* log_tab = nullptr;
};
static std::mutex g_t16_mu;
static std::map<int, Gf16DevTables> g_t16;

int gf16_device_tables(const uint16_t** exp_tab, const uint16_t** inlog) {
	std::lock_guard<std::mutex> lk(g_t16_mu);
	int dev = current_device();
	auto& t = g_t16[dev];
	if (!t.exp_tab) {  // published only once all three are uploaded
		const Gf16HostTables& h = gf16_host_tables();
		Gf16DevTables n;
		hipError_t e = hipMalloc((void**)&n.exp_tab, 65536 * sizeof(uint16_t));
		if (e == hipSuccess) e = hipMalloc((void**)&n.inlog, 32768 * sizeof(uint16_t));
		if (e == hipSuccess) e = hipMalloc((void**)&n.log_tab, 65536 * sizeof(uint16_t));
		if (e == hipSuccess)
			e = hipMemcpy(n.exp_tab, h.exp, 65536 * sizeof(uint16_t), hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemcpy(n.inlog, h.inlog, 32768 * sizeof(uint16_t), hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemcpy(n.log_tab, h.log, 65536 * sizeof(uint16_t), hipMemcpyHostToDevice);
		if (e != hipSuccess) {
			hipFree(n.exp_tab);
			hipFree(n.inlog);
			hipFree(n.log_tab);
			return hip_fail(e, "gf16 device tables");
		}
		t = n;
	}
	*exp_tab = t.exp_tab;
	*inlog = t.inlog;
	return PARHIP_OK;
}

int gf16_device_log_exp(const uint16_t** exp_tab, const uint16_t** log_tab) {
	const uint16_t* il;
	int rc = gf16_device_tables(exp_tab, &il);
	if (rc) return rc;
	std::lock_guard<std::mutex> lk(g_t16_mu);
	*log_tab = g_t16[current_device()].log_tab;
	return PARHIP_OK;
}

// ------------------------------------------------------------------ plan
struct Plan16 {
	int ng;
	uint64_t groups, in_per_group;
};

static Plan16 plan16(uint64_t N, uint64_t R, uint64_t W) {
	Plan16 p;
	p.ng = R > 32 ? 8 : (R > 16 ? 4 : (R > 8 ? 2 : 1));
	const int ib = 4;
	const uint64_t tiles = (W + 1023) / 1024;
	const uint64_t rtiles = (R + 8 * p.ng - 1) / (8 * p.ng);
	const uint64_t target = (uint64_t)device_cus(current_device()) * 4;
	uint64_t groups = 1, base = tiles * rtiles;
	if (base < target) groups = (target + base - 1) / base;
	if (tune_groups_override() > 0) groups = (uint64_t)tune_groups_override();
	groups = std::min<uint64_t>(groups, std::max<uint64_t>(1, N / (4 * ib)));
	groups = std::min<uint64_t>(groups, 65535);
	uint64_t per = (N + groups - 1) / groups;
	per = (per + ib - 1) / ib * ib;
	p.groups = (N + per - 1) / per;
	p.in_per_group = per;
	return p;
}

// GF(2) bit-matrix strategy on MFMA (gf16_mfma.hip): row words of every
// coefficient, then the MFMA contraction; split-K partials folded as for the
// tables.  Output tiles of 8 outputs and 2 column blocks of 32 words per wave,
// two waves per SIMD (measured best for both types at 32768 x 32 KiB, R = 512).
static int run_gf16_mfma(const Gf16Args& m, bool fp4, hipStream_t s) {
	int ot = 4, nb = 2, trt, twpt, tib;
	tune_shape_override(&trt, &twpt, &tib);
	if ((trt || twpt) && gf16_mfma_supported(trt ? trt : ot, twpt ? twpt : nb, fp4)) {
		ot = trt ? trt : ot;
		nb = twpt ? twpt : nb;
	}
	const uint64_t span = 4ull * 32 * nb;
	const uint64_t base = ((m.W + span - 1) / span) * ((m.R + 2 * ot - 1) / (2 * ot));
	const uint64_t target = (uint64_t)device_cus(current_device()) * 2;
	uint64_t groups = base < target ? (target + base - 1) / base : 1;
	if (tune_groups_override() > 0) groups = (uint64_t)tune_groups_override();
	groups = std::min<uint64_t>(groups, std::max<uint64_t>(1, m.N / 16));
	// fp4 counts are fp32 sums of 0/1 products: exact while 16 * inputs < 2^24
	if (fp4) groups = std::max<uint64_t>(groups, (m.N + (1u << 19) - 1) >> 19);
	groups = std::min<uint64_t>(groups, 65535);
	uint64_t per = (m.N + groups - 1) / groups;
	// stages of 4 * ib inputs (ib = 4 unless tuned to 2; 4 inputs per MFMA K
	// step), pairs never straddle groups
	const int gib = fp4 && tib == 2 ? 2 : 4;
	per = (per + 4 * gib - 1) / (4 * gib) * (4 * gib);
	groups = (m.N + per - 1) / per;
	if (fp4 && per >= (1u << 20))
		return fail(PARHIP_ERR_ARG, "gf16 mfma: %llu inputs per group exceed the exact fp32 count",
		            (unsigned long long)per);
	const size_t rows_bytes = (size_t)m.R * ((m.N + 1) / 2) * 16 * sizeof(uint32_t);
	const size_t part_bytes =
	    groups > 1 ? (size_t)gf16_reduce_scratch_words(groups, m.R, m.W) * sizeof(uint16_t) : 0;
	void* ws = nullptr;
	int rc = dev_alloc(&ws, rows_bytes + part_bytes, s);
	if (rc) return rc;
	hipError_t e;
	{
		ProfScope ps("gf16_rows", s);
		e = launch_gf16_rows(m.coeff, m.coeff_ld, m.N, m.R, (uint32_t*)ws, s);
	}
	Gf16MfmaArgs a{};
	a.in_base = m.in_base;
	a.in_stride = m.in_stride;
	a.out_base = m.out_base;
	a.out_stride = m.out_stride;
	a.rows = (const uint32_t*)ws;
	a.N = m.N;
	a.R = m.R;
	a.W = m.W;
	a.in_per_group = per;
	a.partial = groups > 1 ? (uint16_t*)((char*)ws + rows_bytes) : nullptr;
	a.accumulate = m.accumulate;
	if (e == hipSuccess) {
		ProfScope ps("gf16_mfma", s);
		e = launch_gf16_mfma(a, groups, ot, nb, gib, fp4, s);
	}
	if (e == hipSuccess && groups > 1) {
		ProfScope ps("gf16_reduce", s);
		e = launch_gf16_reduce(a.partial, groups, m.R, m.W, m.out_base, m.out_stride, m.accumulate, s);
	}
	dev_free(ws, s);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "k_gf16_mfma");
}

int run_gf16_contraction(Gf16Args a, hipStream_t s) {
	if (a.N == 0 || a.R == 0 || a.W == 0) return PARHIP_OK;
	if (current_method() == PARHIP_METHOD_MFMA || current_method() == PARHIP_METHOD_MFMA_I8)
		return run_gf16_mfma(a, current_method() == PARHIP_METHOD_MFMA, s);
	Plan16 p = plan16(a.N, a.R, a.W);
	a.in_per_group = p.in_per_group;
	a.partial = nullptr;
	void* ws = nullptr;
	if (p.groups > 1) {
		int rc = dev_alloc(&ws, gf16_reduce_scratch_words(p.groups, a.R, a.W) * sizeof(uint16_t), s);
		if (rc) return rc;
		a.partial = (uint16_t*)ws;
	}
	hipError_t e;
	{
		ProfScope ps("gf16_tab", s);
		e = launch_gf16_tab(a, p.groups, p.ng, s);
	}
	if (e != hipSuccess) {
		dev_free(ws, s);
		return hip_fail(e, "k_gf16_tab");
	}
	if (p.groups > 1) {
		ProfScope ps("gf16_reduce", s);
		e = launch_gf16_reduce(a.partial, p.groups, a.R, a.W, a.out_base, a.out_stride,
		                       a.accumulate, s);
		dev_free(ws, s);
		if (e != hipSuccess) return hip_fail(e, "k_gf16_reduce");
	}
	return PARHIP_OK;
}

int gf16_coeffs_device(uint16_t* d_coeff, const uint16_t* d_input_nums, uint64_t N,
                       const uint16_t* d_exps, uint64_t R, hipStream_t s) {
	const uint16_t *et, *il;
	int rc = gf16_device_tables(&et, &il);
	if (rc) return rc;
	hipError_t e = launch_gf16_coeffs(d_coeff, d_input_nums, N, d_exps, R, et, il, s);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "k_gf16_coeffs");
}

}  // namespace parhip

using namespace parhip;

extern "C" {

uint16_t parhip_gf16_coeff(uint16_t inputNum, uint16_t exponent) {
	return gf16_host_coeff(inputNum, exponent);
}

uint16_t parhip_gf16_mul(uint16_t a, uint16_t b) { return gf16_host_mul(a, b); }

int parhip_d_gf16_coeffs(uint16_t* d_coeff, const uint16_t* inputNums, size_t numInputs,
                         const uint16_t* exponents, size_t numOutputs, void* stream) {
	if (numInputs == 0 || numOutputs == 0) return PARHIP_OK;
	if (!d_coeff || !inputNums || !exponents)
		return fail(PARHIP_ERR_ARG, "gf16_coeffs: null pointer");
	hipStream_t s = (hipStream_t)stream;
	DevBuf di, de;
	int rc = di.alloc(numInputs * sizeof(uint16_t), s);
	if (rc) return rc;
	rc = de.alloc(numOutputs * sizeof(uint16_t), s);
	if (rc) return rc;
	hipMemcpyAsync(di.p, inputNums, numInputs * sizeof(uint16_t), hipMemcpyHostToDevice, s);
	hipMemcpyAsync(de.p, exponents, numOutputs * sizeof(uint16_t), hipMemcpyHostToDevice, s);
	rc = gf16_coeffs_device(d_coeff, di.as<uint16_t>(), numInputs, de.as<uint16_t>(), numOutputs, s);
	// pageable-source copies above are staged by the runtime; keep the small
	// index buffers alive until the kernel has consumed them
	hipStreamSynchronize(s);
	return rc;
}

int parhip_d_gf16_muladd(const uint16_t* d_inputs, size_t numInputs, size_t inputStrideWords,
                         uint16_t* d_outputs, size_t numOutputs, size_t outputStrideWords,
                         size_t words, const uint16_t* d_coeff, int accumulate, void* stream) {
	if (numInputs == 0 || numOutputs == 0 || words == 0) return PARHIP_OK;
	if (!d_inputs || !d_outputs || !d_coeff) return fail(PARHIP_ERR_ARG, "gf16_muladd: null pointer");
	if (inputStrideWords < words || outputStrideWords < words)
		return fail(PARHIP_ERR_ARG, "gf16_muladd: stride < words");
	device_cus(current_device());
	Gf16Args a{};
	a.in_base = d_inputs;
	a.in_stride = inputStrideWords;
	a.out_base = d_outputs;
	a.out_stride = outputStrideWords;
	a.coeff = d_coeff;
	a.coeff_ld = numInputs;
	a.N = numInputs;
	a.R = numOutputs;
	a.W = words;
	a.accumulate = accumulate ? 1 : 0;
	return run_gf16_contraction(a, (hipStream_t)stream);
}

}  // extern "C"
