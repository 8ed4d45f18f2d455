// gf16_host.cpp -- PAR2 GF(2^16) tables: generator 2 modulo 0x1100B, and the
// PAR2 input constants 2^(log_i), log_i the i-th positive exponent that none of
// 3, 5, 17, 257 divides (the constants of the PAR2 specification; the
// reference derives them the same way in gf16/gfmat_coeff.c:20-32).
#include "gf16_host.h"

#include <mutex>

namespace parhip {

static Gf16HostTables g_tables;
static std::once_flag g_once;

static void build() {
	uint32_t v = 1;
	for (uint32_t e = 0; e < 65535; e++) {
		g_tables.exp[e] = (uint16_t)v;
		g_tables.log[v] = (uint16_t)e;
		v <<= 1;
		if (v & 0x10000u) v ^= 0x1100Bu;
	}
	g_tables.exp[65535] = g_tables.exp[0];
	g_tables.log[0] = 0;
	uint32_t e = 0;
	for (uint32_t i = 0; i < 32768; i++) {
		do {
			e++;
		} while (e % 3 == 0 || e % 5 == 0 || e % 17 == 0 || e % 257 == 0);
		g_tables.inlog[i] = (uint16_t)e;
	}
}

const Gf16HostTables& gf16_host_tables() {
	std::call_once(g_once, build);
	return g_tables;
}

uint16_t gf16_host_mul(uint16_t a, uint16_t b) {
	const Gf16HostTables& t = gf16_host_tables();
	if (!a || !b) return 0;
	return t.exp[((uint32_t)t.log[a] + t.log[b]) % 65535u];
}

uint16_t gf16_host_coeff(uint16_t inputNum, uint16_t exponent) {
	const Gf16HostTables& t = gf16_host_tables();
	uint64_t l = (uint64_t)t.inlog[inputNum & 32767u] * exponent;
	return t.exp[l % 65535u];
}

}  // namespace parhip
