// gf16_host.h -- host-side PAR2 GF(2^16) tables and element operations.
#pragma once
#include <cstdint>

namespace parhip {

struct Gf16HostTables {
	uint16_t exp[65536];   // 2^e, e in [0, 65535]; exp[65535] == exp[0]
	uint16_t log[65536];   // log[v] for v != 0
	uint16_t inlog[32768]; // PAR2 input constants: log of the i-th base
};

// Built once, thread-safe.
const Gf16HostTables& gf16_host_tables();
uint16_t gf16_host_mul(uint16_t a, uint16_t b);
// PAR2 coefficient of input `inputNum` for recovery exponent `exponent`
// (gf16/gfmat_coeff.c:84-101): 2^((log_i * exponent) mod 65535).
uint16_t gf16_host_coeff(uint16_t inputNum, uint16_t exponent);

}  // namespace parhip
