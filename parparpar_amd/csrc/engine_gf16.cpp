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

int run_gf16_contraction(Gf16Args a, hipStream_t s) {
	if (a.N == 0 || a.R == 0 || a.W == 0) return PARHIP_OK;
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
