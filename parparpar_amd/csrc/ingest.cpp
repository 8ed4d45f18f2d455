// ingest.cpp -- host -> HBM streaming of input slices, overlapped with compute.
//
// The reference's host entries take caller-owned host buffers (JS Buffers,
// aligned_alloc'd reads or mmaps: src/gf64_create_streaming.cc:93-240) and
// the chunked create path re-enters with accumulate=true per chunk
// (lib/par3gen.js:2306-2316).  Here a host call streams its inputs through a
// ring of pinned staging slots:
//
//   fill slot k (host threads: memcpy from the caller's buffer or pread from
//   the file) -> hipMemcpyAsync slot k -> device slot k on a copy stream ->
//   the contraction of that chunk on the compute stream, XOR-accumulating
//   into the device outputs.
//
// While the host fills slot k+1, the DMA engine moves slot k and the GPU
// reduces slot k-1.  Events order reuse: a pinned slot is refilled only after
// its copy finished, a device slot is overwritten only after the contraction
// that read it finished.  Linearity makes the chunked result identical to a
// single contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "ingest.h"

namespace parhip {

// ------------------------------------------------------------------ fill pool
// Persistent host threads for the fill step (one job at a time; concurrent
// callers queue on job_mu_).  The calling thread works on the job too.
class FillPool {
public:
	static FillPool& get() {
		static FillPool p;
		return p;
	}
	void run(size_t nparts, const std::function<void(size_t)>& fn) {
		if (nparts == 0) return;
		std::lock_guard<std::mutex> jl(job_mu_);
		std::unique_lock<std::mutex> lk(mu_);
		job_ = &fn;
		nparts_ = nparts;
		next_ = 0;
		remaining_ = nparts;
		cv_.notify_all();
		while (next_ < nparts_) {
			const size_t i = next_++;
			lk.unlock();
			fn(i);
			lk.lock();
			--remaining_;
		}
		done_cv_.wait(lk, [&] { return remaining_ == 0; });
		job_ = nullptr;
	}
	size_t threads() const { return th_.size() + 1; }

private:
	FillPool() {
		unsigned hw = std::thread::hardware_concurrency();
		unsigned n = hw > 2 ? std::min(hw / 2, 8u) - 1 : 0;
		for (unsigned t = 0; t < n; t++) th_.emplace_back([this] { loop(); });
	}
	~FillPool() {
		{
			std::lock_guard<std::mutex> lk(mu_);
			stop_ = true;
		}
		cv_.notify_all();
		for (auto& t : th_) t.join();
	}
	void loop() {
		std::unique_lock<std::mutex> lk(mu_);
		for (;;) {
			cv_.wait(lk, [&] { return stop_ || (job_ && next_ < nparts_); });
			if (stop_) return;
			const size_t i = next_++;
			const std::function<void(size_t)>* fn = job_;
			lk.unlock();
			(*fn)(i);
			lk.lock();
			if (--remaining_ == 0) done_cv_.notify_all();
		}
	}
	std::mutex job_mu_, mu_;
	std::condition_variable cv_, done_cv_;
	const std::function<void(size_t)>* job_ = nullptr;
	size_t nparts_ = 0, next_ = 0, remaining_ = 0;
	bool stop_ = false;
	std::vector<std::thread> th_;
};

void parallel_parts(size_t nparts, const std::function<void(size_t)>& fn) {
	FillPool::get().run(nparts, fn);
}

// ------------------------------------------------------------------ pinned ring
static constexpr int kSlots = 3;
static constexpr size_t kDefaultChunk = 64ull << 20;
static thread_local size_t t_chunk_bytes = 0;  // 0 = kDefaultChunk

struct Ring {
	int dev = -1;
	size_t bytes = 0;
	void* host[kSlots] = {};
	void* dbuf[kSlots] = {};
	hipEvent_t copied[kSlots] = {}, consumed[kSlots] = {};
	hipStream_t copy = nullptr;

	~Ring() { release(); }
	void release() {
		if (dev < 0) return;
		hipSetDevice(dev);
		if (copy) hipStreamSynchronize(copy);
		for (int k = 0; k < kSlots; k++) {
			if (copied[k]) hipEventDestroy(copied[k]);
			if (consumed[k]) hipEventDestroy(consumed[k]);
			if (host[k]) hipHostFree(host[k]);
			if (dbuf[k]) hipFree(dbuf[k]);
			copied[k] = consumed[k] = nullptr;
			host[k] = dbuf[k] = nullptr;
		}
		if (copy) hipStreamDestroy(copy);
		copy = nullptr;
		dev = -1;
		bytes = 0;
	}
	int ensure(int d, size_t need) {
		if (dev == d && bytes >= need) return PARHIP_OK;
		release();
		dev = d;
		bytes = need;
		hipError_t e = hipStreamCreateWithFlags(&copy, hipStreamNonBlocking);
		for (int k = 0; k < kSlots && e == hipSuccess; k++) {
			e = hipHostMalloc(&host[k], need, hipHostMallocDefault);
			if (e == hipSuccess) e = hipMalloc(&dbuf[k], need);
			if (e == hipSuccess) e = hipEventCreateWithFlags(&copied[k], hipEventDisableTiming);
			if (e == hipSuccess) e = hipEventCreateWithFlags(&consumed[k], hipEventDisableTiming);
		}
		if (e != hipSuccess) {
			release();
			return hip_fail(e, "staging ring");
		}
		return PARHIP_OK;
	}
};
static thread_local Ring t_ring;

void ingest_release() { t_ring.release(); }

size_t ingest_chunk_bytes() { return t_chunk_bytes ? t_chunk_bytes : kDefaultChunk; }

int stream_slices(uint64_t n, size_t slice_bytes, const FillFn& fill, const ConsumeFn& consume,
                  hipStream_t compute) {
	if (n == 0 || slice_bytes == 0) return PARHIP_OK;
	const size_t chunk = ingest_chunk_bytes();
	const uint64_t per = std::max<uint64_t>(1, chunk / slice_bytes);
	RC_(t_ring.ensure(current_device(), (size_t)std::min<uint64_t>(per, n) * slice_bytes));
	Ring& ring = t_ring;
	bool used[kSlots] = {};
	int rc = PARHIP_OK;
	uint64_t c = 0;
	for (uint64_t first = 0; first < n && rc == PARHIP_OK; first += per, c++) {
		const uint64_t count = std::min<uint64_t>(per, n - first);
		const size_t bytes = (size_t)count * slice_bytes;
		const int k = (int)(c % kSlots);
		hipError_t e = hipSuccess;
		if (used[k]) e = hipEventSynchronize(ring.copied[k]);  // pinned slot free
		if (e != hipSuccess) {
			rc = hip_fail(e, "hipEventSynchronize");
			break;
		}
		rc = fill(ring.host[k], first, count);
		if (rc) break;
		if (used[k]) e = hipStreamWaitEvent(ring.copy, ring.consumed[k], 0);  // device slot free
		if (e == hipSuccess)
			e = hipMemcpyAsync(ring.dbuf[k], ring.host[k], bytes, hipMemcpyHostToDevice, ring.copy);
		if (e == hipSuccess) e = hipEventRecord(ring.copied[k], ring.copy);
		if (e == hipSuccess) e = hipStreamWaitEvent(compute, ring.copied[k], 0);
		if (e != hipSuccess) {
			rc = hip_fail(e, "staging copy");
			break;
		}
		rc = consume(ring.dbuf[k], first, count, compute);
		if (rc) break;
		e = hipEventRecord(ring.consumed[k], compute);
		if (e != hipSuccess) {
			rc = hip_fail(e, "hipEventRecord");
			break;
		}
		used[k] = true;
	}
	if (rc) {  // drain what is in flight before the caller's buffers go away
		hipStreamSynchronize(ring.copy);
		hipStreamSynchronize(compute);
	}
	return rc;
}

}  // namespace parhip

extern "C" int parhip_set_stream_chunk(size_t bytes) {
	if (bytes != 0 && bytes < 4096)
		return parhip::fail(PARHIP_ERR_ARG, "set_stream_chunk: %zu bytes is below 4096", bytes);
	parhip::t_chunk_bytes = bytes;
	return PARHIP_OK;
}
