// packets.hip -- PAR3 recovery packets assembled and sealed on the device.
//
// AssembleRecoveryPackets_NAPI (src/gf64_addon.cc:2790-2938) turns R recovery
// blocks into R "PAR REC" packets: 48-byte header (magic "PAR3\0PKT", 16-byte
// BLAKE3 checksum, LE64 packet length, 8-byte InputSetID, type), 16-byte body
// prefix (LE64 block index, LE64 count = 1) and the block bytes; the checksum
// is BLAKE3 over packet[24 .. end) truncated to 16 bytes.  With the recovery
// blocks already in HBM after the contraction, the packets are built and
// hashed here and only the finished packets cross PCIe.
//
// BLAKE3 (published specification; default hash mode) splits the message into
// 1 KiB chunks, each a sequential chain of up to 16 compressions, then merges
// the chunk chaining values in a left-complete binary tree.  Three steps:
//   k_rec_assemble   headers + block copy (grid-stride, coalesced);
//   k_b3_chunks      one thread per (packet, chunk); a single-chunk message
//                    is finished here (ROOT on its last block);
//   k_b3_tree        one work-group per packet: levels of parent compressions,
//                    pairs (2t, 2t+1), an odd last node carried up unchanged,
//                    which is exactly the tree whose left subtree holds the
//                    largest power of two of chunks; ROOT on the last merge.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "kernels.h"

namespace parhip {

__constant__ uint32_t kB3Iv[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
enum : uint32_t { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return __builtin_rotateright32(x, n); }

#define B3_G(a, b, c, d, mx, my)       \
	do {                               \
		v[a] = v[a] + v[b] + (mx);     \
		v[d] = rotr(v[d] ^ v[a], 16);  \
		v[c] = v[c] + v[d];            \
		v[b] = rotr(v[b] ^ v[c], 12);  \
		v[a] = v[a] + v[b] + (my);     \
		v[d] = rotr(v[d] ^ v[a], 8);   \
		v[c] = v[c] + v[d];            \
		v[b] = rotr(v[b] ^ v[c], 7);   \
	} while (0)

// One compression; cv is replaced by the first 8 output words.  The message
// schedule is applied by index (round r uses m[s_r(i)]) instead of permuting
// registers, so the words stay in place.
__device__ __forceinline__ void b3_compress(uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                                            uint32_t block_len, uint32_t flags) {
	// s_r = PERM^r composed: precomputed message index per round and position
	constexpr uint8_t S[7][16] = {
	    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
	    {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8},
	    {3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1},
	    {10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6},
	    {12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4},
	    {9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7},
	    {11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13}};
	uint32_t v[16];
#pragma unroll
	for (int i = 0; i < 8; i++) v[i] = cv[i];
#pragma unroll
	for (int i = 0; i < 4; i++) v[8 + i] = kB3Iv[i];
	v[12] = (uint32_t)counter;
	v[13] = (uint32_t)(counter >> 32);
	v[14] = block_len;
	v[15] = flags;
#pragma unroll
	for (int r = 0; r < 7; r++) {
		B3_G(0, 4, 8, 12, m[S[r][0]], m[S[r][1]]);
		B3_G(1, 5, 9, 13, m[S[r][2]], m[S[r][3]]);
		B3_G(2, 6, 10, 14, m[S[r][4]], m[S[r][5]]);
		B3_G(3, 7, 11, 15, m[S[r][6]], m[S[r][7]]);
		B3_G(0, 5, 10, 15, m[S[r][8]], m[S[r][9]]);
		B3_G(1, 6, 11, 12, m[S[r][10]], m[S[r][11]]);
		B3_G(2, 7, 8, 13, m[S[r][12]], m[S[r][13]]);
		B3_G(3, 4, 9, 14, m[S[r][14]], m[S[r][15]]);
	}
#pragma unroll
	for (int i = 0; i < 8; i++) cv[i] = v[i] ^ v[i + 8];
}

__device__ __forceinline__ void put_le64(uint8_t* p, uint64_t v) {
#pragma unroll
	for (int b = 0; b < 8; b++) p[b] = (uint8_t)(v >> (8 * b));
}

// Headers (checksum zeroed) and block bytes; 8-byte copies when the block size
// keeps every packet 8-byte aligned.
__global__ void k_rec_assemble(const uint8_t* __restrict__ rec, uint64_t isid, uint64_t first,
                               uint64_t num, uint64_t bs, uint8_t* __restrict__ out, int wide) {
	const uint64_t pkt = 64 + bs;
	const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t p = tid; p < num; p += stride) {
		uint8_t* h = out + p * pkt;
		const uint8_t magic[8] = {'P', 'A', 'R', '3', 0, 'P', 'K', 'T'};
		const uint8_t type[8] = {'P', 'A', 'R', ' ', 'R', 'E', 'C', 0};
		for (int b = 0; b < 8; b++) h[b] = magic[b];
		for (int b = 8; b < 24; b++) h[b] = 0;
		put_le64(h + 24, pkt);
		put_le64(h + 32, isid);
		for (int b = 0; b < 8; b++) h[40 + b] = type[b];
		put_le64(h + 48, first + p);
		put_le64(h + 56, 1);
	}
	if (wide) {  // 8-byte aligned buffers and a block size multiple of 8
		const uint64_t wpb = bs / 8, total = num * wpb;
		for (uint64_t t = tid; t < total; t += stride) {
			const uint64_t p = t / wpb, w = t - p * wpb;
			reinterpret_cast<uint64_t*>(out + p * pkt + 64)[w] =
			    reinterpret_cast<const uint64_t*>(rec + p * bs)[w];
		}
	} else {
		const uint64_t total = num * bs;
		for (uint64_t t = tid; t < total; t += stride) {
			const uint64_t p = t / bs, x = t - p * bs;
			out[p * pkt + 64 + x] = rec[p * bs + x];
		}
	}
}

// Message word k (4 bytes at offset o) of packet p: packet bytes [24 + o, +4),
// zero past the message end.  ALIGNED: the block size is a multiple of 4, so
// every word is a 4-byte-aligned load.
template <bool ALIGNED>
__device__ __forceinline__ uint32_t msg_word(const uint8_t* m, uint64_t o, uint64_t len) {
	if (o + 4 <= len) {
		if (ALIGNED) return *reinterpret_cast<const uint32_t*>(m + o);
		return (uint32_t)m[o] | ((uint32_t)m[o + 1] << 8) | ((uint32_t)m[o + 2] << 16) |
		       ((uint32_t)m[o + 3] << 24);
	}
	uint32_t w = 0;
	for (int b = 0; b < 4; b++)
		if (o + b < len) w |= (uint32_t)m[o + b] << (8 * b);
	return w;
}

template <bool ALIGNED>
__global__ __launch_bounds__(256) void k_b3_chunks(uint8_t* __restrict__ out, uint64_t num,
                                                   uint64_t bs, uint64_t chunks,
                                                   uint32_t* __restrict__ cvs) {
	const uint64_t pkt = 64 + bs, len = 40 + bs;
	const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= num * chunks) return;
	const uint64_t p = t / chunks, j = t - p * chunks;
	const uint8_t* m = out + p * pkt + 24;
	const uint64_t c0 = j * 1024;
	const uint64_t clen = (len - c0 < 1024) ? len - c0 : 1024;
	const uint64_t nblocks = (clen + 63) / 64;  // clen >= 1: the message is >= 41 bytes
	uint32_t cv[8];
#pragma unroll
	for (int i = 0; i < 8; i++) cv[i] = kB3Iv[i];
	for (uint64_t b = 0; b < nblocks; b++) {
		const uint64_t o = c0 + b * 64;
		uint32_t w[16];
#pragma unroll
		for (int k = 0; k < 16; k++) w[k] = msg_word<ALIGNED>(m, o + 4 * k, len);
		const uint32_t blen = (uint32_t)((clen - b * 64) < 64 ? clen - b * 64 : 64);
		uint32_t flags = 0;
		if (b == 0) flags |= B3_CHUNK_START;
		if (b == nblocks - 1) flags |= B3_CHUNK_END | (chunks == 1 ? B3_ROOT : 0u);
		b3_compress(cv, w, j, blen, flags);
	}
	if (chunks == 1) {
		uint8_t* h = out + p * pkt + 8;  // 16-byte truncated digest
#pragma unroll
		for (int i = 0; i < 4; i++)
#pragma unroll
			for (int b = 0; b < 4; b++) h[4 * i + b] = (uint8_t)(cv[i] >> (8 * b));
	} else {
#pragma unroll
		for (int i = 0; i < 8; i++) cvs[(p * chunks + j) * 8 + i] = cv[i];
	}
}

// One work-group per packet: merge `chunks` chaining values level by level,
// ping-ponging between a and b (each num x chunks x 8 words).
__global__ __launch_bounds__(256) void k_b3_tree(uint8_t* __restrict__ out, uint64_t bs,
                                                 uint64_t chunks, uint32_t* __restrict__ a,
                                                 uint32_t* __restrict__ b) {
	const uint64_t p = blockIdx.x;
	const uint64_t pkt = 64 + bs;
	uint32_t* src = a + p * chunks * 8;
	uint32_t* dst = b + p * chunks * 8;
	uint64_t n = chunks;
	while (n > 1) {
		const uint64_t half = n / 2;
		const bool root = (n == 2);
		for (uint64_t t = threadIdx.x; t < half; t += blockDim.x) {
			uint32_t m[16], cv[8];
#pragma unroll
			for (int i = 0; i < 16; i++) m[i] = src[2 * t * 8 + i];
#pragma unroll
			for (int i = 0; i < 8; i++) cv[i] = kB3Iv[i];
			b3_compress(cv, m, 0, 64, B3_PARENT | (root ? B3_ROOT : 0u));
			if (root) {
				uint8_t* h = out + p * pkt + 8;
#pragma unroll
				for (int i = 0; i < 4; i++)
#pragma unroll
					for (int k = 0; k < 4; k++) h[4 * i + k] = (uint8_t)(cv[i] >> (8 * k));
			} else {
#pragma unroll
				for (int i = 0; i < 8; i++) dst[t * 8 + i] = cv[i];
			}
		}
		if ((n & 1) && threadIdx.x == 0) {
#pragma unroll
			for (int i = 0; i < 8; i++) dst[half * 8 + i] = src[(n - 1) * 8 + i];
		}
		__syncthreads();  // this level's nodes are complete before the next reads them
		uint32_t* tmp = src;
		src = dst;
		dst = tmp;
		n = half + (n & 1);
	}
}

uint64_t b3_chunks_of(uint64_t bs) { return (40 + bs + 1023) / 1024; }

hipError_t launch_rec_packets(const uint8_t* rec, uint64_t isid, uint64_t first, uint64_t num,
                              uint64_t bs, uint8_t* out, uint32_t* ws, hipStream_t s) {
	if (!num) return hipSuccess;
	const uint64_t chunks = b3_chunks_of(bs);
	const uintptr_t align = reinterpret_cast<uintptr_t>(rec) | reinterpret_cast<uintptr_t>(out);
	const int wide = (bs & 7) == 0 && (align & 7) == 0;
	uint64_t words = wide ? num * bs / 8 : num * bs;
	uint64_t blocks = (std::max<uint64_t>(words, num) + 255) / 256;
	if (blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(k_rec_assemble, dim3((unsigned)blocks), dim3(256), 0, s, rec, isid, first, num,
	                   bs, out, wide);
	const uint64_t threads = num * chunks;
	const uint64_t cblocks = (threads + 255) / 256;
	if (cblocks > 0x7fffffffu) return hipErrorInvalidValue;
	if ((bs & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 3) == 0)
		hipLaunchKernelGGL(k_b3_chunks<true>, dim3((unsigned)cblocks), dim3(256), 0, s, out, num, bs,
		                   chunks, ws);
	else
		hipLaunchKernelGGL(k_b3_chunks<false>, dim3((unsigned)cblocks), dim3(256), 0, s, out, num, bs,
		                   chunks, ws);
	if (chunks > 1) {
		if (num > 0x7fffffffu) return hipErrorInvalidValue;
		hipLaunchKernelGGL(k_b3_tree, dim3((unsigned)num), dim3(256), 0, s, out, bs, chunks, ws,
		                   ws + num * chunks * 8);
	}
	return hipGetLastError();
}

}  // namespace parhip
