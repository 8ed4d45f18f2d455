// engine_packets.cpp -- C ABI for PAR3 recovery-packet assembly (packets.hip).
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "kernels.h"

using namespace parhip;

// AssembleRecoveryPackets_NAPI argument rules (src/gf64_addon.cc:2850-2880).
static int packet_args(const void* rec, const uint8_t* isid, int64_t num, int64_t bs,
                       const void* out) {
	if (!rec || !isid || !out) return fail(PARHIP_ERR_ARG, "assemble_recovery_packets: null pointer");
	if (num <= 0) return fail(PARHIP_ERR_RANGE, "numRecovery must be a positive integer");
	if (bs <= 0) return fail(PARHIP_ERR_RANGE, "blockSize must be a positive integer");
	const size_t cap = (size_t)1 << 40;
	if ((size_t)num > cap / (size_t)bs)
		return fail(PARHIP_ERR_RANGE, "numRecovery * blockSize exceeds 1 TiB cap");
	return PARHIP_OK;
}

static int enqueue_packets(const uint8_t* d_rec, const uint8_t* isid, int64_t first, uint64_t num,
                           uint64_t bs, uint8_t* d_out, hipStream_t s) {
	uint64_t id;
	memcpy(&id, isid, 8);  // little-endian host: bytes land in order
	const uint64_t chunks = b3_chunks_of(bs);
	void* ws = nullptr;
	if (chunks > 1) {
		int rc = dev_alloc(&ws, (size_t)(2 * num * chunks * 8 * sizeof(uint32_t)), s);
		if (rc) return rc;
	}
	hipError_t e;
	{
		ProfScope ps("rec_packets", s);
		e = launch_rec_packets(d_rec, id, (uint64_t)first, num, bs, d_out, (uint32_t*)ws, s);
	}
	dev_free(ws, s);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "rec packets");
}

extern "C" {

int parhip_d_assemble_recovery_packets(const void* d_recovery, const uint8_t* inputSetId,
                                       int64_t firstRecoveryIndex, int64_t numRecovery,
                                       int64_t blockSize, void* d_out, void* stream) {
	int rc = packet_args(d_recovery, inputSetId, numRecovery, blockSize, d_out);
	if (rc) return rc;
	return enqueue_packets((const uint8_t*)d_recovery, inputSetId, firstRecoveryIndex,
	                       (uint64_t)numRecovery, (uint64_t)blockSize, (uint8_t*)d_out,
	                       (hipStream_t)stream);
}

int parhip_assemble_recovery_packets(const uint8_t* recovery, const uint8_t* inputSetId,
                                     int64_t firstRecoveryIndex, int64_t numRecovery,
                                     int64_t blockSize, uint8_t* out) {
	int rc = packet_args(recovery, inputSetId, numRecovery, blockSize, out);
	if (rc) return rc;
	const size_t in_bytes = (size_t)numRecovery * (size_t)blockSize;
	const size_t out_bytes = (size_t)numRecovery * (size_t)(64 + blockSize);
	hipStream_t s;
	rc = host_stream(&s);
	if (rc) return rc;
	DevBuf din, dout;
	if ((rc = din.alloc(in_bytes, s)) || (rc = dout.alloc(out_bytes, s))) return rc;
	hipError_t e = hipMemcpyAsync(din.p, recovery, in_bytes, hipMemcpyHostToDevice, s);
	if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
	rc = enqueue_packets(din.as<uint8_t>(), inputSetId, firstRecoveryIndex, (uint64_t)numRecovery,
	                     (uint64_t)blockSize, dout.as<uint8_t>(), s);
	if (rc) {
		hipStreamSynchronize(s);
		return rc;
	}
	e = hipMemcpyAsync(out, dout.p, out_bytes, hipMemcpyDeviceToHost, s);
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "packets D2H");
}

}  // extern "C"
