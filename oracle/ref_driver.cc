// ref_driver.cc -- C entry points onto the reference engine compiled in place.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
// gf64/*.c, src/par3_engine.cc and src/par3_topology.cc straight from the
// reference tree into oracle/_ref/libref_par3.so together with this file.
// This file only forwards plain-C calls to the reference's C++ class so that
// tests and the CPU-baseline timer can call it through ctypes.  It contains
// no arithmetic of its own.
#include <cstddef>
#include <cstdint>
#include "par3_engine.h"
#include "../hasher/hasher_blake3.h"

extern "C" {
uint64_t gf64_mul_reference(uint64_t a, uint64_t b);
int gf64_init_dispatch(void);

int ref_init(void) { return gf64_init_dispatch(); }

int ref_method(void) { return (int)gf64_current_method; }

uint64_t ref_gf64_mul(uint64_t a, uint64_t b) { return gf64_mul_reference(a, b); }

uint64_t ref_gf64_inv(uint64_t a) { return gf64_inverse(a); }

void ref_build_cauchy(uint64_t *out, size_t n, size_t r, uint64_t fi, uint64_t fr) {
	GF64Controller::BuildCauchyMatrix(out, n, r, fi, fr);
}

void ref_compute_recovery(const uint64_t *in, size_t n, uint64_t *rec, size_t r, size_t bs,
                          uint64_t fi, uint64_t fr, int threads, int accumulate) {
	GF64Controller::ComputeRecoveryBlocks(in, n, rec, r, bs, fi, fr, threads, accumulate != 0);
}

void ref_compute_recovery_with_coeff(const uint64_t *in, size_t n, uint64_t *rec, size_t r,
                                     size_t bs, const uint64_t *coeff, int threads,
                                     int accumulate) {
	GF64Controller::ComputeRecoveryBlocksWithCoeff(in, n, rec, r, bs, coeff, threads,
	                                               accumulate != 0);
}

int ref_solve_and_reconstruct(uint64_t *A, uint64_t *rhs, size_t n, size_t bs, int threads) {
	return GF64Controller::SolveAndReconstruct(A, rhs, n, bs, threads);
}

/* The reference's BLAKE3 (hasher/hasher_blake3.c), as its packetizer calls it
 * (src/gf64_addon.cc:2928-2933): two updates, 32-byte finalize. */
void ref_blake3_2(const uint8_t *a, size_t alen, const uint8_t *b, size_t blen, uint8_t *out32) {
	blake3_hasher h;
	blake3_hasher_init(&h);
	blake3_hasher_update(&h, a, alen);
	blake3_hasher_update(&h, b, blen);
	blake3_hasher_finalize(&h, out32, 32);
}
}
