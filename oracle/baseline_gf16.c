/*
 * baseline_gf16.c -- CPU baseline for the PAR2 GF(2^16) create, timed by
 * bench.py's cpu_baseline leg beside the GPU contraction.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle library).  The product never
 * links this file.
 *
 * It computes what PAR2ProcCPU computes (gf16/controller_cpu.cpp:463-537 driving
 * Galois16Mul::mul_add_multi_packed, gf16/gf16mul.h:360-390):
 *   out[e][w] = XOR_i coeff(inputNum_i, exponent_e) * in_i[w]
 * over 16-bit little-endian words, coefficients from or_gf16_coeff
 * (gf16/gfmat_coeff.c:84-101).  The method is the split-nibble shuffle form:
 * a product c * v of a 16-bit word v is the XOR of four 16-entry tables indexed
 * by v's nibbles, T_k[n] = c * (n x^{4k}); with the low and high result bytes in
 * separate tables each lookup is one byte shuffle (AVX2 vpshufb, 32 lanes).
 * Words are split into a low-byte and a high-byte plane once per input, so every
 * (input, output) pair costs 8 shuffles and 8 XORs per 32 words; accumulators
 * stay in the split form until the end.  Threads own disjoint output ranges.
 * Words past a multiple of 32 (and hosts without AVX2) use the scalar
 * or_gf16_muladd_region.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

typedef struct {
	const uint16_t *in;
	size_t n, words;
	const uint16_t *input_nums, *exps;
	uint16_t *out;
	size_t e0, e1;  /* this thread's outputs */
	int avx2;
} bl_job;

/* c * x^b for b = 0..15, then T_k[n] low / high bytes for k = 0..3. */
static void bl_tables(uint16_t c, uint8_t lo[4][16], uint8_t hi[4][16]) {
	uint16_t basis[16];
	for (int b = 0; b < 16; b++) basis[b] = or_gf16_mul(c, (uint16_t)(1u << b));
	for (int k = 0; k < 4; k++)
		for (int n = 0; n < 16; n++) {
			uint16_t v = 0;
			for (int j = 0; j < 4; j++)
				if (n & (1 << j)) v ^= basis[4 * k + j];
			lo[k][n] = (uint8_t)v;
			hi[k][n] = (uint8_t)(v >> 8);
		}
}

#if defined(__x86_64__)
/* Split 32-word chunks of one slice into nibble planes: for chunk q,
 * nib[(q*4 + k)*32 .. +32) holds nibble k (0,1 of the low byte, 2,3 of the
 * high byte) of the chunk's words, in the byte order of _mm256_packus_epi16. */
__attribute__((target("avx2"))) static void bl_split(const uint16_t *src, size_t chunks,
                                                      uint8_t *nib) {
	const __m256i m00ff = _mm256_set1_epi16(0x00FF), m0f = _mm256_set1_epi8(0x0F);
	for (size_t q = 0; q < chunks; q++) {
		const __m256i a = _mm256_loadu_si256((const __m256i *)(src + 32 * q));
		const __m256i b = _mm256_loadu_si256((const __m256i *)(src + 32 * q + 16));
		const __m256i lo = _mm256_packus_epi16(_mm256_and_si256(a, m00ff), _mm256_and_si256(b, m00ff));
		const __m256i hi = _mm256_packus_epi16(_mm256_srli_epi16(a, 8), _mm256_srli_epi16(b, 8));
		__m256i *d = (__m256i *)(nib + q * 128);
		_mm256_storeu_si256(d + 0, _mm256_and_si256(lo, m0f));
		_mm256_storeu_si256(d + 1, _mm256_and_si256(_mm256_srli_epi16(lo, 4), m0f));
		_mm256_storeu_si256(d + 2, _mm256_and_si256(hi, m0f));
		_mm256_storeu_si256(d + 3, _mm256_and_si256(_mm256_srli_epi16(hi, 4), m0f));
	}
}

/* acc (lo plane, hi plane per chunk: acc[q*64 .. +32) low bytes, +32 high) ^= c * slice */
__attribute__((target("avx2"))) static void bl_muladd(uint8_t *acc, const uint8_t *nib,
                                                       size_t chunks, uint16_t c) {
	uint8_t lo[4][16], hi[4][16];
	bl_tables(c, lo, hi);
	__m256i tl[4], th[4];
	for (int k = 0; k < 4; k++) {
		tl[k] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)lo[k]));
		th[k] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)hi[k]));
	}
	for (size_t q = 0; q < chunks; q++) {
		const __m256i *s = (const __m256i *)(nib + q * 128);
		const __m256i n0 = _mm256_loadu_si256(s + 0), n1 = _mm256_loadu_si256(s + 1);
		const __m256i n2 = _mm256_loadu_si256(s + 2), n3 = _mm256_loadu_si256(s + 3);
		__m256i *a = (__m256i *)(acc + q * 64);
		__m256i al = _mm256_loadu_si256(a), ah = _mm256_loadu_si256(a + 1);
		al = _mm256_xor_si256(al, _mm256_xor_si256(_mm256_shuffle_epi8(tl[0], n0), _mm256_shuffle_epi8(tl[1], n1)));
		al = _mm256_xor_si256(al, _mm256_xor_si256(_mm256_shuffle_epi8(tl[2], n2), _mm256_shuffle_epi8(tl[3], n3)));
		ah = _mm256_xor_si256(ah, _mm256_xor_si256(_mm256_shuffle_epi8(th[0], n0), _mm256_shuffle_epi8(th[1], n1)));
		ah = _mm256_xor_si256(ah, _mm256_xor_si256(_mm256_shuffle_epi8(th[2], n2), _mm256_shuffle_epi8(th[3], n3)));
		_mm256_storeu_si256(a, al);
		_mm256_storeu_si256(a + 1, ah);
	}
}

/* split accumulators back to words */
__attribute__((target("avx2"))) static void bl_join(const uint8_t *acc, size_t chunks, uint16_t *dst) {
	for (size_t q = 0; q < chunks; q++) {
		const __m256i al = _mm256_loadu_si256((const __m256i *)(acc + q * 64));
		const __m256i ah = _mm256_loadu_si256((const __m256i *)(acc + q * 64 + 32));
		_mm256_storeu_si256((__m256i *)(dst + 32 * q), _mm256_unpacklo_epi8(al, ah));
		_mm256_storeu_si256((__m256i *)(dst + 32 * q + 16), _mm256_unpackhi_epi8(al, ah));
	}
}
#endif

static void *bl_worker(void *arg) {
	bl_job *j = (bl_job *)arg;
	const size_t ne = j->e1 - j->e0;
	if (!ne) return NULL;
	const size_t chunks = j->avx2 ? j->words / 32 : 0, tail0 = chunks * 32;
	uint8_t *acc = chunks ? (uint8_t *)calloc(ne, chunks * 64) : NULL;
	uint8_t *nib = chunks ? (uint8_t *)malloc(chunks * 128) : NULL;
	for (size_t e = j->e0; e < j->e1; e++) memset(j->out + e * j->words, 0, j->words * 2);
	for (size_t i = 0; i < j->n; i++) {
		const uint16_t *src = j->in + i * j->words;
#if defined(__x86_64__)
		if (chunks) bl_split(src, chunks, nib);
#endif
		for (size_t e = j->e0; e < j->e1; e++) {
			const uint16_t c = or_gf16_coeff(j->input_nums[i], j->exps[e]);
#if defined(__x86_64__)
			if (chunks) bl_muladd(acc + (e - j->e0) * chunks * 64, nib, chunks, c);
#endif
			if (tail0 < j->words)
				or_gf16_muladd_region((uint8_t *)(j->out + e * j->words + tail0),
				                      (const uint8_t *)(src + tail0), (j->words - tail0) * 2, c);
		}
	}
#if defined(__x86_64__)
	if (chunks) {
		uint16_t *tmp = (uint16_t *)malloc(tail0 * 2);
		for (size_t e = j->e0; e < j->e1; e++) {
			bl_join(acc + (e - j->e0) * chunks * 64, chunks, tmp);
			memcpy(j->out + e * j->words, tmp, tail0 * 2);
		}
		free(tmp);
	}
#endif
	free(acc);
	free(nib);
	return NULL;
}

/* out (r x words) = PAR2 create of n input slices of `words` 16-bit words,
 * input numbers input_nums[n], exponents exps[r], on `threads` threads.
 * Returns the number of threads used. */
int bl_par2_create(const uint16_t *in, size_t n, size_t words, const uint16_t *input_nums,
                   const uint16_t *exps, size_t r, uint16_t *out, int threads) {
	if (threads < 1) threads = 1;
	if ((size_t)threads > r) threads = (int)(r ? r : 1);
	int avx2 = 0;
#if defined(__x86_64__)
	avx2 = __builtin_cpu_supports("avx2");
#endif
	(void)or_gf16_coeff(0, 0); /* tables built before the threads start */
	bl_job *jobs = (bl_job *)calloc((size_t)threads, sizeof(bl_job));
	pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
	for (int t = 0; t < threads; t++) {
		bl_job *j = &jobs[t];
		j->in = in;
		j->n = n;
		j->words = words;
		j->input_nums = input_nums;
		j->exps = exps;
		j->out = out;
		j->e0 = r * (size_t)t / (size_t)threads;
		j->e1 = r * (size_t)(t + 1) / (size_t)threads;
		j->avx2 = avx2;
		pthread_create(&th[t], NULL, bl_worker, j);
	}
	for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
	free(jobs);
	free(th);
	return threads;
}
