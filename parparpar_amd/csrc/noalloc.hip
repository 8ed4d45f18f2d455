// noalloc.hip -- the plain Cauchy create on device buffers without any
// allocation per call: parhip_d_gf64_recovery_static.
//
// parhip_d_gf64_recovery takes its scratch memory -- the XOR-block
// coefficients and the split-K partials -- from the coefficient cache and the
// stream-ordered pool.  Callers that must not allocate (work recorded under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc)
// use this entry point instead: both buffers are carved from `scratch`, one
// fixed device array per GPU that stays allocated until the process ends.
// The array is addressed through the device address hipGetSymbolAddress
// returns for the current device (its name in host code is not a device
// address).  Calls that overlap on one GPU share the array and race; calls
// on one stream, or ordered by events, do not.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "kernels.h"

namespace parhip {

// 64 MiB: the coefficients of up to 1 Mi (input block x output group) pairs
// plus the split-K partials, whose row count is reduced to what is left.
static constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) uint64_t scratch[kScratchBytes / sizeof(uint64_t)];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint64_t* g_scratch_dev[kMaxDevices];  // device address of `scratch` per device

// The device address of `scratch` on device `dev`, looked up once per device
// with that device current.
static int scratch_address(int dev, uint64_t** out) {
	if (dev < 0 || dev >= kMaxDevices)
		return fail(PARHIP_ERR_ARG, "recovery_static: device %d out of range", dev);
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_dev[dev]) {
		hipError_t e = hipSetDevice(dev);
		void* p = nullptr;
		if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
		if (!p) return fail(PARHIP_ERR_DEVICE, "recovery_static: no device address for the scratch array");
		g_scratch_dev[dev] = (uint64_t*)p;
	}
	*out = g_scratch_dev[dev];
	return PARHIP_OK;
}

}  // namespace parhip

using namespace parhip;

extern "C" int parhip_d_gf64_recovery_static(const uint64_t* d_inputs, size_t numInputs,
                                             size_t inputStride64, uint64_t* d_recovery,
                                             size_t numRecovery, size_t recoveryStride64,
                                             size_t blockSize64, uint64_t firstInput,
                                             uint64_t firstRecovery, int accumulate, void* stream) {
	if (numInputs == 0 || numRecovery == 0 || blockSize64 == 0) return PARHIP_OK;
	if (!d_inputs || !d_recovery) return fail(PARHIP_ERR_ARG, "recovery_static: null pointer");
	if (inputStride64 < blockSize64 || recoveryStride64 < blockSize64)
		return fail(PARHIP_ERR_ARG, "recovery_static: stride < blockSize64");
	const int dev = current_device();
	const uint64_t nblocks = ((firstInput & 7) + numInputs + 7) / 8;
	const uint64_t Q = ((firstRecovery & 7) + numRecovery + 7) / 8;
	// coefficient words, checked against the array before any multiplication
	// can overflow: nblocks and Q are at most 2^61
	if (Q > 65535 || nblocks > kScratchBytes / 64 || nblocks * Q > kScratchBytes / 64)
		return fail(PARHIP_ERR_RANGE,
		            "recovery_static: %zu inputs x %zu recovery slices need %llu x %llu coefficient "
		            "blocks of 64 B; the scratch array holds %zu MiB",
		            numInputs, numRecovery, (unsigned long long)nblocks, (unsigned long long)Q,
		            kScratchBytes >> 20);
	const size_t coeff_bytes = (size_t)(nblocks * Q * 64 + 255) & ~(size_t)255;
	uint64_t* base = nullptr;
	int rc = scratch_address(dev, &base);
	if (rc) return rc;
	device_cus(dev);
	hipStream_t s = (hipStream_t)stream;
	// every R runs the XOR-block kernel here (parhip_d_gf64_recovery keeps the
	// direct tables below 4 outputs); the bytes are the same
	const DyadicPlan p = plan_dyadic(numInputs, numRecovery, blockSize64, firstInput, firstRecovery,
	                                 dev, kScratchBytes - coeff_bytes);
	MacArgs a{};
	a.in_base = d_inputs;
	a.in_stride = inputStride64;
	a.out_base = d_recovery;
	a.out_stride = recoveryStride64;
	a.N = numInputs;
	a.R = numRecovery;
	a.W = blockSize64;
	a.accumulate = accumulate ? 1 : 0;
	uint64_t* gp = base;
	void* part = (char*)base + coeff_bytes;
	hipError_t e;
	{
		ProfScope ps("dyadic_coeffs", s);
		e = launch_dyadic_coeffs(gp, nblocks, Q, firstInput & ~7ull, firstRecovery & ~7ull, s);
	}
	if (e == hipSuccess) e = enqueue_dyadic(a, p, gp, nblocks, part, s);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "recovery_static");
}
