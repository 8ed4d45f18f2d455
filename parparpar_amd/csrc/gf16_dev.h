// gf16_dev.h -- GF(2^16) element helpers (PAR2 field, polynomial 0x1100B,
// gf16/gf16_global.h:9) for device and host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace parhip {

// c * x mod 0x1100B (c < 2^16)
__host__ __device__ __forceinline__ uint32_t gf16_mulx(uint32_t c) {
	c <<= 1;
	return (c & 0x10000u) ? (c ^ 0x1100Bu) : c;
}

// c * x^s mod 0x1100B
__host__ __device__ __forceinline__ uint32_t gf16_mulx_pow(uint32_t c, int s) {
	for (int k = 0; k < s; k++) c = gf16_mulx(c);
	return c;
}

}  // namespace parhip
