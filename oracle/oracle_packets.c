/*
 * oracle_packets.c -- CPU restatement of PAR3 recovery-packet assembly and
 * its BLAKE3 seal.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * BLAKE3 is restated from its published specification (J. O'Connor,
 * J.-P. Aumasson, S. Neves, Z. Wilcox-O'Hearn, "BLAKE3: one function, fast
 * everywhere", 2020): compression function (section 2.2), chunk chaining
 * (2.3) and the binary tree whose left subtree holds the largest power of two
 * of chunks smaller than the total (2.4), written here as the recursive tree
 * definition rather than the incremental stack of the reference's
 * hasher/hasher_blake3.c, which is the reference's implementation of the same
 * function.  Only the default hash mode is used (key = IV, 32-byte output).
 *
 * Packet layout follows AssembleRecoveryPackets_NAPI (src/gf64_addon.cc:
 * 2890-2938): magic "PAR3\0PKT", 16-byte truncated BLAKE3 over
 * header[24..48] || body, LE64 packet length, 8-byte InputSetID, type
 * "PAR REC\0", body LE64 block index (firstRecoveryIndex + i), LE64 count 1,
 * then the recovery block bytes.
 */
#include <string.h>

#include "oracle.h"

static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const int B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

static uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void b3_g(uint32_t *v, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
	v[a] = v[a] + v[b] + mx;
	v[d] = rotr32(v[d] ^ v[a], 16);
	v[c] = v[c] + v[d];
	v[b] = rotr32(v[b] ^ v[c], 12);
	v[a] = v[a] + v[b] + my;
	v[d] = rotr32(v[d] ^ v[a], 8);
	v[c] = v[c] + v[d];
	v[b] = rotr32(v[b] ^ v[c], 7);
}

/* 7 rounds of column then diagonal G steps, message permuted between rounds;
 * out[0..8) = new chaining value (the first 8 output words). */
void or_blake3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[16]) {
	uint32_t v[16], m[16], t[16];
	for (int i = 0; i < 8; i++) v[i] = cv[i];
	for (int i = 0; i < 4; i++) v[8 + i] = B3_IV[i];
	v[12] = (uint32_t)counter;
	v[13] = (uint32_t)(counter >> 32);
	v[14] = block_len;
	v[15] = flags;
	for (int i = 0; i < 16; i++) m[i] = block[i];
	for (int r = 0; r < 7; r++) {
		b3_g(v, 0, 4, 8, 12, m[0], m[1]);
		b3_g(v, 1, 5, 9, 13, m[2], m[3]);
		b3_g(v, 2, 6, 10, 14, m[4], m[5]);
		b3_g(v, 3, 7, 11, 15, m[6], m[7]);
		b3_g(v, 0, 5, 10, 15, m[8], m[9]);
		b3_g(v, 1, 6, 11, 12, m[10], m[11]);
		b3_g(v, 2, 7, 8, 13, m[12], m[13]);
		b3_g(v, 3, 4, 9, 14, m[14], m[15]);
		for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]];
		for (int i = 0; i < 16; i++) m[i] = t[i];
	}
	for (int i = 0; i < 8; i++) {
		out[i] = v[i] ^ v[i + 8];
		out[i + 8] = v[i + 8] ^ cv[i];
	}
}

static void b3_words(const uint8_t *p, size_t n, uint32_t w[16]) { /* little-endian, zero-padded */
	uint8_t buf[64];
	memset(buf, 0, sizeof buf);
	memcpy(buf, p, n);
	for (int i = 0; i < 16; i++)
		w[i] = (uint32_t)buf[4 * i] | ((uint32_t)buf[4 * i + 1] << 8) |
		       ((uint32_t)buf[4 * i + 2] << 16) | ((uint32_t)buf[4 * i + 3] << 24);
}

/* Chaining value of chunk `index` (len <= 1024 bytes); with root != 0 the
 * last block carries ROOT and out receives the root output words. */
static void b3_chunk(const uint8_t *p, size_t len, uint64_t index, int root, uint32_t out[16]) {
	uint32_t cv[8], w[16], o[16];
	memcpy(cv, B3_IV, sizeof cv);
	const size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
	for (size_t b = 0; b < nblocks; b++) {
		const size_t n = (b + 1) * 64 <= len ? 64 : len - b * 64;
		uint32_t flags = 0;
		if (b == 0) flags |= B3_CHUNK_START;
		if (b == nblocks - 1) flags |= B3_CHUNK_END | (root ? B3_ROOT : 0);
		b3_words(p + b * 64, n, w);
		or_blake3_compress(cv, w, index, (uint32_t)n, flags, o);
		memcpy(cv, o, sizeof cv);
	}
	memcpy(out, o, sizeof o);
}

/* Subtree over chunks [lo, hi) of the message: left holds the largest power
 * of two of chunks below hi - lo. */
static void b3_tree(const uint8_t *msg, size_t len, uint64_t lo, uint64_t hi, int root,
                    uint32_t out[16]) {
	if (hi - lo == 1) {
		const size_t off = (size_t)lo * 1024;
		b3_chunk(msg + off, len - off < 1024 ? len - off : 1024, lo, root, out);
		return;
	}
	uint64_t left = 1;
	while (left * 2 < hi - lo) left *= 2;
	uint32_t l[16], r[16], block[16];
	b3_tree(msg, len, lo, lo + left, 0, l);
	b3_tree(msg, len, lo + left, hi, 0, r);
	memcpy(block, l, 32);
	memcpy(block + 8, r, 32);
	or_blake3_compress(B3_IV, block, 0, 64, B3_PARENT | (root ? B3_ROOT : 0), out);
}

void or_blake3(const uint8_t *msg, size_t len, uint8_t out[32]) {
	const uint64_t chunks = len == 0 ? 1 : (len + 1023) / 1024;
	uint32_t o[16];
	b3_tree(msg, len, 0, chunks, 1, o);
	for (int i = 0; i < 8; i++)
		for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(o[i] >> (8 * b));
}

static void put_le64(uint8_t *p, uint64_t v) {
	for (int b = 0; b < 8; b++) p[b] = (uint8_t)(v >> (8 * b));
}

void or_assemble_recovery_packets(const uint8_t *recovery, const uint8_t inputSetId[8],
                                  uint64_t firstRecoveryIndex, size_t numRecovery,
                                  size_t blockSize, uint8_t *out) {
	static const uint8_t magic[8] = {'P', 'A', 'R', '3', 0, 'P', 'K', 'T'};
	static const uint8_t type[8] = {'P', 'A', 'R', ' ', 'R', 'E', 'C', 0};
	const size_t pkt = 64 + blockSize;
	for (size_t i = 0; i < numRecovery; i++) {
		uint8_t *p = out + i * pkt;
		memcpy(p, magic, 8);
		memset(p + 8, 0, 16);
		put_le64(p + 24, (uint64_t)pkt);
		memcpy(p + 32, inputSetId, 8);
		memcpy(p + 40, type, 8);
		put_le64(p + 48, firstRecoveryIndex + i);
		put_le64(p + 56, 1);
		memcpy(p + 64, recovery + i * blockSize, blockSize);
		uint8_t h[32];
		or_blake3(p + 24, pkt - 24, h);
		memcpy(p + 8, h, 16);
	}
}
