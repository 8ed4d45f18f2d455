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
// the chunk chaining values in a left-complete binary tree.
//   k_rec_assemble   headers + block copy (grid-stride, coalesced);
//   k_b3_chunks      a work-group (one wave) = 64 chunks of one packet: the
//                    64 KiB of message are first copied into LDS with coalesced
//                    loads (rows padded to 257 dwords, so the lanes' reads fall
//                    on different banks), then each lane compresses its chunk
//                    from LDS (the four column steps of a round are independent
//                    chains), and the group's 64 values are merged in LDS:
//                    parents (2t, 2t+1) per level, an odd last node carried up.
//                    An aligned run of 64 chunks is its own subtree of the
//                    left-complete tree, so the merge is exact; a packet of one
//                    group (or one chunk) is finished here with ROOT last;
//   k_b3_tree        merges the per-group values the same way, 1024 per
//                    work-group per pass, ROOT on the last merge.
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
// schedule is applied by index (round r uses m[S_r(i)], S_r the published
// permutation applied r times) instead of permuting registers.
__device__ __forceinline__ void b3_compress(uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                                            uint32_t block_len, uint32_t flags) {
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

constexpr int kCvPad = 9;  // LDS row of a chaining value: 8 words + 1 (bank spread)

// Merge n chaining values held in LDS rows (cv, kCvPad words each), a thread
// per parent, levels separated by barriers; the last merge gets ROOT when
// `top` and its first 4 words go to digest.  On return the result is cv[0].
// Every thread of the work-group calls it.
__device__ __forceinline__ void b3_merge_lds(uint32_t* cv, uint32_t* tmp, uint32_t n, bool top,
                                             uint32_t* digest) {
	uint32_t* src = cv;
	uint32_t* dst = tmp;
	while (n > 1) {
		const uint32_t half = n / 2;
		const bool root = top && n == 2;
		for (uint32_t t = threadIdx.x; t < half; t += blockDim.x) {
			uint32_t m[16], h[8];
#pragma unroll
			for (int i = 0; i < 8; i++) {
				m[i] = src[(2 * t) * kCvPad + i];
				m[8 + i] = src[(2 * t + 1) * kCvPad + i];
				h[i] = kB3Iv[i];
			}
			b3_compress(h, m, 0, 64, B3_PARENT | (root ? B3_ROOT : 0u));
			if (root) {
#pragma unroll
				for (int i = 0; i < 4; i++) digest[i] = h[i];
			} else {
#pragma unroll
				for (int i = 0; i < 8; i++) dst[t * kCvPad + i] = h[i];
			}
		}
		if ((n & 1) && threadIdx.x < 8) dst[half * kCvPad + threadIdx.x] = src[(n - 1) * kCvPad + threadIdx.x];
		__syncthreads();
		uint32_t* x = src;
		src = dst;
		dst = x;
		n = half + (n & 1);
	}
	if (src != cv && threadIdx.x < 8) cv[threadIdx.x] = src[threadIdx.x];
	__syncthreads();
}

__device__ __forceinline__ void put_digest(uint8_t* h, const uint32_t* w) {
#pragma unroll
	for (int i = 0; i < 4; i++)
#pragma unroll
		for (int b = 0; b < 4; b++) h[4 * i + b] = (uint8_t)(w[i] >> (8 * b));
}

// Work-group = 64 chunks of one packet (blockIdx.x = packet * groups + group):
// all 256 threads copy the message into LDS, the first wave compresses.
template <bool ALIGNED>
__global__ __launch_bounds__(256) void k_b3_chunks(uint8_t* __restrict__ out, uint64_t bs,
                                                  uint64_t chunks, uint64_t groups,
                                                  uint32_t* __restrict__ gcv) {
	constexpr int kRow = 257;  // dwords per chunk row in LDS
	__shared__ uint32_t msgw[64 * kRow];
	__shared__ uint32_t cv[64 * kCvPad], tmp[32 * kCvPad];
	__shared__ uint32_t digest[4];
	const uint64_t p = blockIdx.x / groups, grp = blockIdx.x - p * groups;
	const uint64_t pkt = 64 + bs, len = 40 + bs;
	const uint8_t* msg = out + p * pkt + 24;
	const uint32_t n = (uint32_t)((chunks - grp * 64) < 64 ? chunks - grp * 64 : 64);
	// coalesced copy of the group's message words (zero past the end): 16 loads
	// in flight per thread; a segment that ends inside the message takes plain
	// aligned loads, the last one the byte-exact path
	const uint64_t gbase = grp * 65536;
	const uint32_t total = n * 256;
	for (uint32_t k0 = 0; k0 < total; k0 += 256 * 16) {
		uint32_t r[16];
		if (ALIGNED && gbase + 4ull * (k0 + 256 * 16) <= len) {
#pragma unroll
			for (int u = 0; u < 16; u++)
				r[u] = *reinterpret_cast<const uint32_t*>(msg + gbase + 4ull * (k0 + u * 256 + threadIdx.x));
		} else {
#pragma unroll
			for (int u = 0; u < 16; u++) r[u] = msg_word<ALIGNED>(msg, gbase + 4ull * (k0 + u * 256 + threadIdx.x), len);
		}
#pragma unroll
		for (int u = 0; u < 16; u++) {
			const uint32_t k = k0 + u * 256 + threadIdx.x;
			if (k < total) msgw[(k >> 8) * kRow + (k & 255)] = r[u];
		}
	}
	__syncthreads();
	const uint32_t t = threadIdx.x;
	if (t < n) {
		const uint64_t j = grp * 64 + t;
		const uint64_t c0 = j * 1024;
		const uint64_t clen = (len - c0 < 1024) ? len - c0 : 1024;
		const uint32_t nblocks = (uint32_t)((clen + 63) / 64);  // >= 1: the message is >= 41 bytes
		uint32_t h[8];
#pragma unroll
		for (int i = 0; i < 8; i++) h[i] = kB3Iv[i];
		for (uint32_t b = 0; b < nblocks; b++) {
			uint32_t w[16];
#pragma unroll
			for (int k = 0; k < 16; k++) w[k] = msgw[t * kRow + 16 * b + k];
			const uint32_t blen = (uint32_t)((clen - b * 64) < 64 ? clen - b * 64 : 64);
			uint32_t flags = 0;
			if (b == 0) flags |= B3_CHUNK_START;
			if (b == nblocks - 1) flags |= B3_CHUNK_END | (chunks == 1 ? B3_ROOT : 0u);
			b3_compress(h, w, j, blen, flags);
		}
		if (chunks == 1) {
			put_digest(out + p * pkt + 8, h);
		} else {
#pragma unroll
			for (int i = 0; i < 8; i++) cv[t * kCvPad + i] = h[i];
		}
	}
	if (chunks == 1) return;  // uniform
	__syncthreads();
	b3_merge_lds(cv, tmp, n, groups == 1, digest);
	if (groups == 1) {
		if (threadIdx.x == 0) put_digest(out + p * pkt + 8, digest);
	} else if (threadIdx.x < 8) {
		gcv[(p * groups + grp) * 8 + threadIdx.x] = cv[threadIdx.x];
	}
}

// Merge a packet's chaining values in aligned runs of up to 1024 (each an
// exact subtree: 1024 values of 2^k chunks each): work-group (packet, run)
// writes the run's value to `outcv`, or the digest when one run is left.
__global__ __launch_bounds__(256) void k_b3_tree(uint8_t* __restrict__ out, uint64_t bs, uint64_t n_in,
                                                 const uint32_t* __restrict__ incv, uint64_t runs,
                                                 uint32_t* __restrict__ outcv) {
	__shared__ uint32_t cv[1024 * kCvPad];
	__shared__ uint32_t tmp[512 * kCvPad];
	__shared__ uint32_t digest[4];
	const uint64_t p = blockIdx.x / runs, run = blockIdx.x - p * runs;
	const uint64_t first = run * 1024;
	const uint32_t n = (uint32_t)((n_in - first) < 1024 ? n_in - first : 1024);
	for (uint32_t t = threadIdx.x; t < n * 8; t += blockDim.x)
		cv[(t / 8) * kCvPad + t % 8] = incv[(p * n_in + first) * 8 + t];
	__syncthreads();
	b3_merge_lds(cv, tmp, n, runs == 1, digest);
	if (runs == 1) {
		if (threadIdx.x == 0) put_digest(out + p * (64 + bs) + 8, digest);
	} else if (threadIdx.x < 8) {
		outcv[(p * runs + run) * 8 + threadIdx.x] = cv[threadIdx.x];
	}
}

uint64_t b3_chunks_of(uint64_t bs) { return (40 + bs + 1023) / 1024; }

hipError_t launch_rec_packets(const uint8_t* rec, uint64_t isid, uint64_t first, uint64_t num,
                              uint64_t bs, uint8_t* out, uint32_t* ws, hipStream_t s) {
	if (!num) return hipSuccess;
	const uint64_t chunks = b3_chunks_of(bs);
	const uint64_t groups = (chunks + 63) / 64;
	const uintptr_t align = reinterpret_cast<uintptr_t>(rec) | reinterpret_cast<uintptr_t>(out);
	const int wide = (bs & 7) == 0 && (align & 7) == 0;
	uint64_t words = wide ? num * bs / 8 : num * bs;
	uint64_t blocks = (std::max<uint64_t>(words, num) + 255) / 256;
	if (blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(k_rec_assemble, dim3((unsigned)blocks), dim3(256), 0, s, rec, isid, first, num,
	                   bs, out, wide);
	const uint64_t cblocks = num * groups;
	if (cblocks > 0x7fffffffu) return hipErrorInvalidValue;
	if ((bs & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 3) == 0)
		hipLaunchKernelGGL(k_b3_chunks<true>, dim3((unsigned)cblocks), dim3(256), 0, s, out, bs, chunks,
		                   groups, ws);
	else
		hipLaunchKernelGGL(k_b3_chunks<false>, dim3((unsigned)cblocks), dim3(256), 0, s, out, bs, chunks,
		                   groups, ws);
	// ws: num x groups values from k_b3_chunks, then ping-pong halves for the passes
	uint32_t* a = ws;
	uint32_t* b = ws + num * groups * 8;
	for (uint64_t n = groups; n > 1;) {
		const uint64_t runs = (n + 1023) / 1024;
		if (num * runs > 0x7fffffffu) return hipErrorInvalidValue;
		hipLaunchKernelGGL(k_b3_tree, dim3((unsigned)(num * runs)), dim3(256), 0, s, out, bs, n, a, runs,
		                   b);
		uint32_t* t = a;
		a = b;
		b = t;
		n = runs == 1 ? 1 : runs;
	}
	return hipGetLastError();
}

}  // namespace parhip
