// gf64_host.h -- host-side GF(2^64) helpers of the engine (not a hot path).
#pragma once
#include <cstddef>
#include <cstdint>

namespace parhip {

uint64_t host_gf64_mul(uint64_t a, uint64_t b);
uint64_t host_gf64_inv(uint64_t a);
// In-place Gauss-Jordan with the reference's pivot rule; E receives the
// accumulated row transform (A^-1 on success).  Returns 0, or -1 if singular.
int host_gauss_jordan(uint64_t* A, uint64_t* E, size_t n);

}  // namespace parhip
