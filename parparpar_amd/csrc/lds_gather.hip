// lds_gather.hip -- the chip's LDS gather rate for the table kernels' access
// pattern, measured on the device (no counterpart in the reference).
//
// The XOR-block and nibble-table kernels are bound by random 16-byte (and some
// 8-byte) LDS reads: every lane reads one entry of a 16-entry, 256-byte table
// row, a different row per read.  This kernel does only that, with many waves
// per CU and one 3-input XOR per read, so its time per wave-level read is the
// rate those kernels can approach.  Every access is in bounds: offsets are
// 16 * nibble inside rows 0..111 of a 28 KiB table.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/parhip.h"
#include "engine_internal.h"

namespace parhip {

constexpr int kGatherRows = 112;   // 7 pair tables x 16 positions, as k_mac_dyadic's tables
constexpr int kGatherIters = 4096; // steps of 8 reads per thread

template <bool B128>
__global__ __launch_bounds__(256) void k_lds_gather(uint4* __restrict__ out, int iters) {
	__shared__ uint4 tab[kGatherRows * 16];
	for (int i = threadIdx.x; i < kGatherRows * 16; i += 256) tab[i] = make_uint4(i, i * 3u, i * 5u, i * 7u);
	__syncthreads();
	uint32_t x = threadIdx.x * 2654435761u + blockIdx.x * 97u;
	uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
	const char* base = reinterpret_cast<const char*>(tab);
	for (int it = 0; it < iters; it++) {
		x = x * 1664525u + 1013904223u;
		const int row0 = (it * 8) % kGatherRows;  // rows row0 .. row0 + 7 < kGatherRows (112 = 14 x 8)
#pragma unroll
		for (int k = 0; k < 8; k++) {
			const uint32_t ad = ((x >> (4 * k)) & 15u) * 16u;
			if (B128) {
				const uint4 v = *reinterpret_cast<const uint4*>(base + (row0 + k) * 256 + ad);
				a0 = __builtin_amdgcn_bitop3_b32(a0, v.x, v.z, 0x96);
				a1 = __builtin_amdgcn_bitop3_b32(a1, v.y, v.w, 0x96);
			} else {
				const uint2 v = *reinterpret_cast<const uint2*>(base + (row0 + k) * 256 + ad);
				a0 ^= v.x;
				a1 ^= v.y;
			}
		}
	}
	out[blockIdx.x * 256 + threadIdx.x] = make_uint4(a0, a1, a2, a3);
}

}  // namespace parhip

using namespace parhip;

extern "C" int parhip_measure_lds_gather(double* ns_b128, double* ns_b64) {
	if (!ns_b128 || !ns_b64) return fail(PARHIP_ERR_ARG, "measure_lds_gather: null output");
	const int cus = device_cus(current_device());
	if (cus <= 0) return fail(PARHIP_ERR_DEVICE, "measure_lds_gather: no device");
	const int blocks = cus * 8;  // LDS for 5 work-groups per CU; 8 per CU queue behind them
	uint4* out = nullptr;
	hipError_t e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint4));
	if (e != hipSuccess) return hip_fail(e, "measure_lds_gather buffer");
	hipEvent_t a = nullptr, b = nullptr;
	e = hipEventCreate(&a);
	if (e == hipSuccess) e = hipEventCreate(&b);
	double res[2] = {0, 0};
	for (int v = 0; v < 2 && e == hipSuccess; v++) {
		auto k = v == 0 ? k_lds_gather<true> : k_lds_gather<false>;
		hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, nullptr, out, kGatherIters);  // warm-up
		e = hipEventRecord(a, nullptr);
		for (int r = 0; r < 5 && e == hipSuccess; r++) {
			hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, nullptr, out, kGatherIters);
			e = hipGetLastError();
		}
		if (e == hipSuccess) e = hipEventRecord(b, nullptr);
		if (e == hipSuccess) e = hipEventSynchronize(b);
		float ms = 0;
		if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
		// wave-level reads per CU over the 5 launches
		const double reads = 5.0 * blocks * (256 / 64) * (double)kGatherIters * 8 / cus;
		res[v] = ms * 1e6 / reads;
	}
	if (a) (void)hipEventDestroy(a);
	if (b) (void)hipEventDestroy(b);
	(void)hipFree(out);
	if (e != hipSuccess) return hip_fail(e, "measure_lds_gather");
	*ns_b128 = res[0];
	*ns_b64 = res[1];
	return PARHIP_OK;
}
