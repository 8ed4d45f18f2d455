// controller_hip.cpp -- the PAR2 engine (controller_hip.h) and its C ABI.
#include "controller_hip.h"

#include <emmintrin.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>

#include "../../include/parhip.h"
#include "../../include/parhip_par2.h"
#include "engine_internal.h"
#include "gf16_host.h"
#include "kernels.h"

namespace parhip {


// Checksum word of a staged slice of n words, in m = ceil(n/8) chunks of 8
// (the last zero-padded): word w = 8j + k carries the weight x^(m(7-k) + m-1-j)
// in GF(2^16), every word a different power of x below x^(8m).  It is the
// positional checksum of the reference's staging (gf16/gf16_checksum_generic.h:
// per block, multiply the running checksum by x and XOR the block in) with the
// 8 lanes of a block folded into one word.  Linear over GF(2^16), so the muladd
// carries it (cksum(sum_i c_i in_i) = sum_i c_i cksum(in_i)), and words that land
// in the wrong place change it.  Computed with SSE2 over 8 lanes, the chunks
// split into 4 runs stepped independently (a step is x per chunk) and joined
// with one multiply per lane.
static inline __m128i ck_step(__m128i a, __m128i v) {
	const __m128i carry = _mm_and_si128(_mm_srai_epi16(a, 15), _mm_set1_epi16(0x100B));
	return _mm_xor_si128(_mm_xor_si128(_mm_slli_epi16(a, 1), carry), v);
}

// Chunk range [c0, c1) of run r out of `runs` over n words: the full chunks
// split evenly; the last run also takes the zero-padded partial chunk.
struct CkRun {
	size_t c0, c1;
	bool tail;
};
static CkRun ck_run(size_t n, size_t runs, size_t r) {
	const size_t full = n / 8;
	CkRun k{full * r / runs, full * (r + 1) / runs, r + 1 == runs && (n % 8) != 0};
	return k;
}

// Join run values: run r is followed by x^(chunks after it); then fold the 8
// lanes with x^(m(7-k)).
static uint16_t ck_join(const __m128i* v, size_t runs, size_t n) {
	const Gf16HostTables& t = gf16_host_tables();
	const size_t m = (n + 7) / 8;
	uint16_t lane[8] = {0, 0, 0, 0, 0, 0, 0, 0}, run[8];
	for (size_t r = 0; r < runs; r++) {
		_mm_storeu_si128(reinterpret_cast<__m128i*>(run), v[r]);
		const CkRun k = ck_run(n, runs, r);
		const uint64_t after = m - (k.c1 + (k.tail ? 1 : 0));
		const uint16_t f = t.exp[after % 65535u];
		for (int q = 0; q < 8; q++) lane[q] ^= gf16_host_mul(run[q], f);
	}
	uint16_t ck = 0;
	for (int q = 0; q < 8; q++) ck ^= gf16_host_mul(lane[q], t.exp[((uint64_t)m * (7 - q)) % 65535u]);
	return ck;
}

// The checksum of n words with four runs stepped together on this thread
// (independent dependency chains).
static uint16_t slice_checksum(const uint16_t* p, size_t n) {
	constexpr size_t R = 4;
	CkRun k[R];
	for (size_t r = 0; r < R; r++) k[r] = ck_run(n, R, r);
	__m128i a[R];
	for (size_t r = 0; r < R; r++) a[r] = _mm_setzero_si128();
	const size_t L = k[0].c1 - k[0].c0;  // runs 0..2 are L chunks, run 3 L or more
	for (size_t j = 0; j < L; j++)
		for (size_t r = 0; r < R; r++)
			a[r] = ck_step(a[r], _mm_loadu_si128(reinterpret_cast<const __m128i*>(p + 8 * (k[r].c0 + j))));
	for (size_t r = 0; r < R; r++) {  // the chunks of runs longer than L (at most one) and the tail
		for (size_t j = k[r].c0 + L; j < k[r].c1; j++)
			a[r] = ck_step(a[r], _mm_loadu_si128(reinterpret_cast<const __m128i*>(p + 8 * j)));
		if (k[r].tail) {
			uint16_t tail[8] = {0, 0, 0, 0, 0, 0, 0, 0};
			std::memcpy(tail, p + 8 * k[r].c1, (n - 8 * k[r].c1) * 2);
			a[r] = ck_step(a[r], _mm_loadu_si128(reinterpret_cast<const __m128i*>(tail)));
		}
	}
	return ck_join(a, R, n);
}

// Stage one slice: copy `size` bytes of src into the slot (n words, zero-padded
// past the source) and return the checksum of the n words -- one pass that
// reads the caller's bytes once and writes the pinned slot with
// non-temporal stores (no read of the destination lines, no second read for
// the checksum).  Same chunk runs as slice_checksum, so the same value; `dst`
// is 16-byte aligned.
static uint16_t stage_copy_checksum(uint16_t* dst, const void* src, size_t size, size_t n) {
	constexpr size_t R = 4;
	const unsigned char* s8 = static_cast<const unsigned char*>(src);
	CkRun k[R];
	for (size_t r = 0; r < R; r++) k[r] = ck_run(n, R, r);
	__m128i a[R];
	for (size_t r = 0; r < R; r++) a[r] = _mm_setzero_si128();
	// chunk c (words 8c..8c+7): bytes [16c, 16c + 16) of the source, zeros past it
	auto chunk = [&](size_t c) -> __m128i {
		const size_t b0 = 16 * c;
		if (b0 + 16 <= size) return _mm_loadu_si128(reinterpret_cast<const __m128i*>(s8 + b0));
		alignas(16) unsigned char t[16] = {0};
		if (b0 < size) std::memcpy(t, s8 + b0, size - b0);
		return _mm_load_si128(reinterpret_cast<const __m128i*>(t));
	};
	const size_t L = k[0].c1 - k[0].c0;
	for (size_t j = 0; j < L; j++)
		for (size_t r = 0; r < R; r++) {
			const size_t c = k[r].c0 + j;
			const __m128i v = chunk(c);
			_mm_stream_si128(reinterpret_cast<__m128i*>(dst + 8 * c), v);
			a[r] = ck_step(a[r], v);
		}
	for (size_t r = 0; r < R; r++) {
		for (size_t c = k[r].c0 + L; c < k[r].c1; c++) {
			const __m128i v = chunk(c);
			_mm_stream_si128(reinterpret_cast<__m128i*>(dst + 8 * c), v);
			a[r] = ck_step(a[r], v);
		}
		if (k[r].tail) {  // the partial last chunk: its words below n, zeros above
			const size_t c = k[r].c1;
			alignas(16) uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
			const __m128i v = chunk(c);
			_mm_store_si128(reinterpret_cast<__m128i*>(t), v);
			for (size_t w = n - 8 * c; w < 8; w++) t[w] = 0;
			std::memcpy(dst + 8 * c, t, 16);
			a[r] = ck_step(a[r], _mm_load_si128(reinterpret_cast<const __m128i*>(t)));
		}
	}
	_mm_sfence();  // the streamed stores are visible before the slot is reported staged
	return ck_join(a, R, n);
}

// ---------------------------------------------------------------- engine
static constexpr unsigned kPollers = 2;  // copy workers polling for slots at once

Par2Engine::Par2Engine(int device, unsigned stagingAreas, unsigned nworkers) : dev(device) {
	areas.resize(stagingAreas ? stagingAreas : 2);
	pend.reset(new std::atomic<unsigned>[areas.size()]);
	slotq.reset(new std::atomic<uint64_t>[areas.size()]);
	for (size_t a = 0; a < areas.size(); a++) pend[a] = 0, slotq[a] = 0;
	hipError_t e = hipSetDevice(dev);
	if (e == hipSuccess) e = hipStreamCreateWithFlags(&strm, hipStreamNonBlocking);
	if (e == hipSuccess) e = hipStreamCreateWithFlags(&cstrm, hipStreamNonBlocking);
	if (e != hipSuccess) setErr(std::string("stream: ") + hipGetErrorString(e));
	if (!nworkers) {
		const unsigned hw = std::thread::hardware_concurrency();
		nworkers = std::max(1u, std::min(8u, hw ? hw : 1u));
	}
	for (unsigned w = 0; w < nworkers; w++) workers.emplace_back(&Par2Engine::worker, this, w);
	devThread = std::thread(&Par2Engine::deviceLoop, this);
}

Par2Engine::~Par2Engine() { deinit(); }

void Par2Engine::setErr(const std::string& what) {
	std::lock_guard<std::mutex> lk(emu);
	if (err.empty()) err = what;
}

bool Par2Engine::ok() {
	std::lock_guard<std::mutex> lk(emu);
	return err.empty();
}

std::string Par2Engine::lastError() {
	std::lock_guard<std::mutex> lk(emu);
	return err;
}

bool Par2Engine::init(unsigned batch) {
	std::unique_lock<std::mutex> lk(mu);
	drain(lk);
	autoBatch = !batch;
	inputBatchSize = batch ? batch : batchFor(sliceSize);
	if (autoMin) minInBatchSize = inputBatchSize;
	currentArea = fillArea = fillSlot = 0;
	statBatchesStarted = 0;
	processingAdd = false;
	return allocStaging() && allocOutputs();
}

// Inputs per batch when the caller leaves it to the engine: about 64 MiB of
// slices (a contraction over fewer inputs rereads the recovery slices more
// often per input byte), between 16 and 512 inputs.
unsigned Par2Engine::batchFor(size_t slice) {
	const size_t n = (64u << 20) / std::max<size_t>(slice, 1);
	unsigned b = 16;
	while (b < 512 && 2 * (size_t)b <= n) b *= 2;
	return b;
}

// Everything staged and processed, no worker task open.
void Par2Engine::drain(std::unique_lock<std::mutex>& lk) {
	cv.wait(lk, [&] { return activeCount == 0 && tasksOpen == 0; });
}

void Par2Engine::freeStaging() {
	hipSetDevice(dev);
	if (cstrm) hipStreamSynchronize(cstrm);
	if (strm) hipStreamSynchronize(strm);
	for (auto& st : areas) {
		if (st.hsrc) hipHostFree(st.hsrc);
		if (st.hcoef) hipHostFree(st.hcoef);
		if (st.hnum) hipHostFree(st.hnum);
		if (st.dsrc) hipFree(st.dsrc);
		if (st.dcoef) hipFree(st.dcoef);
		if (st.dnum) hipFree(st.dnum);
		if (st.copied) hipEventDestroy(st.copied);
		st = Area();
	}
}

bool Par2Engine::allocStaging() {
	freeStaging();
	if (!allocWords || !inputBatchSize) return true;
	hipSetDevice(dev);
	const size_t nOut = std::max<size_t>(outputExponents.size(), 1);
	for (auto& st : areas) {
		const size_t sb = (size_t)inputBatchSize * strideWords * sizeof(uint16_t);
		const size_t cb = nOut * inputBatchSize * sizeof(uint16_t);
		const size_t nb = (size_t)inputBatchSize * sizeof(uint16_t);
		hipError_t e;
		if ((e = hipHostMalloc((void**)&st.hsrc, sb, 0)) != hipSuccess ||
		    (e = hipHostMalloc((void**)&st.hcoef, cb, 0)) != hipSuccess ||
		    (e = hipHostMalloc((void**)&st.hnum, nb, 0)) != hipSuccess ||
		    (e = hipMalloc((void**)&st.dsrc, sb)) != hipSuccess ||
		    (e = hipMalloc((void**)&st.dcoef, cb)) != hipSuccess ||
		    (e = hipMalloc((void**)&st.dnum, nb)) != hipSuccess ||
		    (e = hipEventCreateWithFlags(&st.copied, hipEventDisableTiming)) != hipSuccess) {
			setErr(std::string("staging allocation: ") + hipGetErrorString(e));
			return false;
		}
		std::memset(st.hsrc, 0, sb);
		st.slots.assign(inputBatchSize, Task{});
	}
	return true;
}

bool Par2Engine::allocOutputs() {
	hipSetDevice(dev);
	if (dout) {
		hipStreamSynchronize(strm);
		hipFree(dout);
		dout = nullptr;
	}
	for (auto* h : hout) hipHostFree(h);
	hout.clear();
	if (!allocWords || outputExponents.empty()) return true;
	hipError_t e = hipMalloc((void**)&dout, outputExponents.size() * strideWords * sizeof(uint16_t));
	for (size_t w = 0; e == hipSuccess && w <= workers.size(); w++) {
		uint16_t* h = nullptr;
		e = hipHostMalloc((void**)&h, strideWords * sizeof(uint16_t), 0);
		if (e == hipSuccess) hout.push_back(h);
	}
	if (e != hipSuccess) {
		setErr(std::string("output allocation: ") + hipGetErrorString(e));
		return false;
	}
	allocOutputsCount = outputExponents.size();
	return true;
}

void Par2Engine::setSliceSize(size_t size) {
	std::unique_lock<std::mutex> lk(mu);
	drain(lk);
	sliceSize = currentSliceSize = size;
	if (autoBatch) {
		inputBatchSize = batchFor(size);
		if (autoMin) minInBatchSize = inputBatchSize;
	}
	words = (size + 1) / 2 + 1;  // payload words (odd tail byte padded) + checksum word
	allocWords = words;
	strideWords = (words + 7) & ~(size_t)7;  // 16-byte aligned rows (streamed stores)
	allocStaging();
	allocOutputs();
}

bool Par2Engine::setCurrentSliceSize(size_t size) {
	const size_t w = (size + 1) / 2 + 1;
	{
		std::unique_lock<std::mutex> lk(mu);
		drain(lk);
		if (w <= allocWords) {
			currentSliceSize = size;
			words = w;
			return true;
		}
	}
	setSliceSize(size);
	return ok();
}

bool Par2Engine::setRecoverySlices(unsigned numSlices, const uint16_t* exponents) {
	std::unique_lock<std::mutex> lk(mu);
	drain(lk);
	hipSetDevice(dev);
	outputExponents.assign(numSlices, 1);  // no exponents: custom coefficients (controller_cpu.cpp:176)
	if (exponents) std::copy(exponents, exponents + numSlices, outputExponents.begin());
	processingAdd = false;
	bool okay = true;
	if (numSlices > allocOutputsCount || !dout) okay = allocOutputs() && allocStaging();
	// the exponents on the device, for the batches' coefficient kernel
	if (dexps) hipFree(dexps);
	dexps = nullptr;
	if (numSlices) {
		hipError_t e = hipMalloc((void**)&dexps, numSlices * sizeof(uint16_t));
		if (e == hipSuccess)
			e = hipMemcpy(dexps, outputExponents.data(), numSlices * sizeof(uint16_t), hipMemcpyHostToDevice);
		if (e != hipSuccess) {
			setErr(std::string("exponents: ") + hipGetErrorString(e));
			okay = false;
		}
	}
	lk.unlock();
	return okay && ok();
}

void Par2Engine::setBatchCallbacks(Par2BatchFn started, Par2BatchFn finished, void* user) {
	std::lock_guard<std::mutex> lk(mu);
	cbStarted = started;
	cbFinished = finished;
	cbUser = user;
}

Par2AddState Par2Engine::canAdd() {
	std::lock_guard<std::mutex> lk(mu);
	if (areas[currentArea].active) return PAR2_ADD_FULL;
	return activeCount + 1 < areas.size() ? PAR2_ADD_OK : PAR2_ADD_OK_BUSY;
}

void Par2Engine::waitForAdd() {
	std::unique_lock<std::mutex> lk(mu);
	cv.wait(lk, [&] { return !areas[currentArea].active; });
}

void Par2Engine::setCoeffs(Area& st, unsigned idx, uint16_t inputNum) {
	// gf16_host_coeff with the modulo 65535 folded by hand: the product of two
	// values below 2^16 fits 32 bits, and x mod 65535 = (x & 0xFFFF) + (x >> 16)
	// up to one subtraction
	const Gf16HostTables& t = gf16_host_tables();
	const uint32_t li = t.inlog[inputNum & 32767u];
	for (size_t e = 0; e < outputExponents.size(); e++) {
		uint32_t x = li * outputExponents[e];
		x = (x & 0xFFFFu) + (x >> 16);
		if (x >= 65535u) x -= 65535u;
		st.hcoef[e * inputBatchSize + idx] = t.exp[x];
	}
}

void Par2Engine::setCoeffs(Area& st, unsigned idx, const uint16_t* coeffs) {
	for (size_t e = 0; e < outputExponents.size(); e++)
		st.hcoef[e * inputBatchSize + idx] = coeffs[e];
}

// Close area a (mu held): no more slots; it goes to the device, in close
// order, once all of its slots are staged.  Rule of the reference backends
// (gf16/controller_ocl.cpp:441-456): an area closes when full, on flush, or
// early when nothing is on the device and it holds minInBatchSize inputs.
void Par2Engine::closeArea(unsigned a) {
	Area& st = areas[a];
	st.closed = true;
	st.active = true;
	st.overwrite = !processingAdd;
	processingAdd = true;
	st.seq = closeSeq++;
	activeCount++;
	statBatchesStarted++;
	currentArea = (a + 1) % (unsigned)areas.size();
}

// Claim the next slot of the current area (mu held via lk), waiting while that
// area is still on the device.
bool Par2Engine::claimSlot(bool flushNow, unsigned* area, unsigned* slot, bool* closed,
                           unsigned* closedN, std::unique_lock<std::mutex>& lk) {
	cv.wait(lk, [&] { return !areas[currentArea].active; });
	if (!areas[currentArea].hsrc) {
		setErr("add: no staging memory (slice size not set)");
		return false;
	}
	const unsigned a = currentArea;
	Area& st = areas[a];
	*area = a;
	*slot = st.count++;
	*closed = st.count >= inputBatchSize || flushNow ||
	          (activeCount == 0 && areas.size() > 1 && st.count >= minInBatchSize);
	*closedN = st.count;
	if (*closed) closeArea(a);
	return true;
}

// Queue a read-back task (slots are published per area instead).
void Par2Engine::push(const Task& t) {
	{
		std::lock_guard<std::mutex> lk(qmu);
		queue.push_back(t);
		queued++;
		if (sleepers.load()) qcv.notify_one();
	}
}

// Publish slot `slot` of area a to the workers (mu held); slots are claimed
// and published in order under mu, so the published count becomes slot + 1.
void Par2Engine::publish(unsigned a, unsigned slot, const Task& t) {
	areas[a].slots[slot] = t;
	pend[a]++;
	tasksOpen++;
	stageAvail++;  // before the slot appears: never below the slots available
	slotq[a].fetch_add(1ull << 32);
}

// After a slot is published: wake a worker only if every worker sleeps.  An
// awake worker, busy or polling, looks for slots before it sleeps, and a
// worker that takes a slot with more published behind it and finds no other
// worker polling wakes one (helpWanted), so the adding thread makes no system
// call while the workers keep up.  The notify is under qmu, so a worker
// checking its wait condition cannot miss the new work; the counter is read
// after the slot was published and a worker goes to sleep in the opposite
// order (sleepers, then the check), so one of the two sees the other.
void Par2Engine::wake() {
	if (sleepers.load() == workers.size()) {
		std::lock_guard<std::mutex> lk(qmu);
		qcv.notify_one();
	}
}

void Par2Engine::helpWanted() {
	if (stageAvail.load() && !spinners.load() && sleepers.load()) {
		std::lock_guard<std::mutex> lk(qmu);
		qcv.notify_one();
	}
}

bool Par2Engine::addInput(const void* buffer, size_t size, uint16_t inputNum, bool flushNow,
                          Par2DoneFn done, void* user) {
	unsigned a, slot, n;
	bool closed;
	Par2BatchFn started;
	void* cu;
	{
		std::unique_lock<std::mutex> lk(mu);
		if (!claimSlot(flushNow, &a, &slot, &closed, &n, lk)) return false;
		areas[a].hnum[slot] = inputNum;  // coefficients: on the device at submission
		publish(a, slot, Task{0, a, slot, inputNum, buffer, size, done, nullptr, user, 0, nullptr});
		started = cbStarted;
		cu = cbUser;
	}
	wake();
	if (closed && started) started(cu, a, n);
	return true;
}

bool Par2Engine::addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flushNow,
                          Par2DoneFn done, void* user) {
	unsigned a, slot, n;
	bool closed;
	Par2BatchFn started;
	void* cu;
	{
		std::unique_lock<std::mutex> lk(mu);
		if (!claimSlot(flushNow, &a, &slot, &closed, &n, lk)) return false;
		setCoeffs(areas[a], slot, coeffs);  // the caller's array may change after the call
		areas[a].custom = true;
		publish(a, slot, Task{2, a, slot, 0, buffer, size, done, nullptr, user, 0, nullptr});
		started = cbStarted;
		cu = cbUser;
	}
	wake();
	if (closed && started) started(cu, a, n);
	return true;
}

// Benchmarking (gf16/controller.h:303-305): an input over whatever its slot
// holds (fillInput), no copy.
void Par2Engine::dummyInput(uint16_t inputNum, bool flushNow) {
	unsigned a, slot, n;
	bool closed;
	Par2BatchFn started;
	void* cu;
	{
		std::unique_lock<std::mutex> lk(mu);
		if (!claimSlot(flushNow, &a, &slot, &closed, &n, lk)) return;
		areas[a].hnum[slot] = inputNum;
		// published like a staged slot (slots are taken in order), nothing to copy
		publish(a, slot, Task{3, a, slot, inputNum, nullptr, 0, nullptr, nullptr, nullptr, 0, nullptr});
		started = cbStarted;
		cu = cbUser;
	}
	wake();
	if (closed && started) started(cu, a, n);
}

// Benchmarking (gf16/controller.h:303-305): fill the next staging slot with
// `buffer` (the current slice size), as PAR2ProcCPU::fillInput does
// (gf16/controller_cpu.cpp:357-369); true once every slot of every staging
// area holds it, after which dummyInput() adds inputs over that data.
bool Par2Engine::fillInput(const void* buffer) {
	std::unique_lock<std::mutex> lk(mu);
	cv.wait(lk, [&] { return !areas[fillArea].active; });
	Area& st = areas[fillArea];
	if (!st.hsrc) {
		setErr("fillInput: no staging memory");
		return true;
	}
	uint16_t* slot = st.hsrc + (size_t)fillSlot * strideWords;
	const size_t payloadWords = words - 1;
	std::memcpy(slot, buffer, currentSliceSize);
	std::memset(reinterpret_cast<char*>(slot) + currentSliceSize, 0, words * 2 - currentSliceSize);
	slot[payloadWords] = slice_checksum(slot, payloadWords);
	if (++fillSlot == inputBatchSize) {
		fillSlot = 0;
		if (++fillArea == areas.size()) {
			fillArea = 0;
			return true;  // all filled
		}
	}
	return false;
}

void Par2Engine::flush() {
	unsigned a, n;
	Par2BatchFn started;
	void* cu;
	{
		std::lock_guard<std::mutex> lk(mu);
		a = currentArea;
		n = areas[a].count;
		if (!n || areas[a].closed) return;
		closeArea(a);
		submitReady();
		started = cbStarted;
		cu = cbUser;
	}
	if (started) started(cu, a, n);
}

// Hand every closed area whose slots are all staged to the device thread, in
// close order (mu held).
void Par2Engine::submitReady() {
	for (bool again = true; again;) {
		again = false;
		for (unsigned a = 0; a < areas.size(); a++) {
			Area& st = areas[a];
			if (st.closed && !st.submitted && pend[a].load() == 0 && st.seq == submitSeq) {
				st.submitted = true;
				submitSeq++;
				devPush(DevEvent{0, a, st.count});
				again = true;
			}
		}
	}
}

void Par2Engine::devPush(const DevEvent& ev) {
	{
		std::lock_guard<std::mutex> lk(dmu);
		devq.push_back(ev);
	}
	dcv.notify_one();
}

// The device thread: submits areas in the order they were handed over and
// frees the areas whose batches finished.
void Par2Engine::deviceLoop() {
	hipSetDevice(dev);
	for (;;) {
		DevEvent ev;
		{
			std::unique_lock<std::mutex> lk(dmu);
			dcv.wait(lk, [&] { return !devq.empty(); });
			ev = devq.front();
			devq.pop_front();
		}
		if (ev.kind == 2) return;
		if (ev.kind == 0) submit(ev.area);
		else finishArea(ev.area, ev.n);
	}
}

// An area's batch finished on the device: the area takes slots again.
void Par2Engine::finishArea(unsigned a, unsigned n) {
	Par2BatchFn finished;
	void* cu;
	{
		std::lock_guard<std::mutex> lk(mu);
		Area& st = areas[a];
		st.count = 0;
		st.closed = st.active = st.submitted = st.custom = false;
		slotq[a].store(0);  // every slot of the area was taken and staged
		activeCount--;
		finished = cbFinished;
		cu = cbUser;
	}
	cv.notify_all();
	if (finished) finished(cu, a, n);
}

// One batch to the device (device thread; the area is closed and staged, so
// nothing else writes it until finishArea): slots and coefficients, the
// contraction (the pass's first batch overwrites, later ones XOR), then a host
// function that reports the batch finished.
void Par2Engine::submit(unsigned a) {
	Area& st = areas[a];
	const unsigned n = st.count;
	hipSetDevice(dev);
	// host-to-device copies on the copy stream, so that they overlap the
	// previous batch's contraction; the compute stream waits for them.  The
	// area's device buffers are written again only after this batch finished.
	hipError_t e = hipMemcpyAsync(st.dsrc, st.hsrc, (size_t)n * strideWords * 2, hipMemcpyHostToDevice,
	                              cstrm);
	// coefficients: from the input numbers on the device (PAR2's rule,
	// gf16/gfmat_coeff.c:84-101), or, where the caller gave some, the host
	// columns with the input-number slots filled in here
	uint64_t coeff_ld = n;
	if (!st.custom) {
		if (e == hipSuccess)
			e = hipMemcpyAsync(st.dnum, st.hnum, (size_t)n * 2, hipMemcpyHostToDevice, cstrm);
	} else {
		for (unsigned k = 0; k < n; k++)
			if (st.slots[k].kind != 2) setCoeffs(st, k, st.hnum[k]);
		coeff_ld = inputBatchSize;
		if (e == hipSuccess)
			e = hipMemcpyAsync(st.dcoef, st.hcoef, outputExponents.size() * inputBatchSize * 2,
			                   hipMemcpyHostToDevice, cstrm);
	}
	if (e == hipSuccess) e = hipEventRecord(st.copied, cstrm);
	if (e == hipSuccess) e = hipStreamWaitEvent(strm, st.copied, 0);
	if (e == hipSuccess && !st.custom &&
	    gf16_coeffs_device(st.dcoef, st.dnum, n, dexps, outputExponents.size(), strm) != PARHIP_OK)
		setErr("gf16 coefficient kernel launch failed");
	if (e == hipSuccess) {
		Gf16Args g{};
		g.in_base = st.dsrc;
		g.in_stride = strideWords;
		g.out_base = dout;
		g.out_stride = strideWords;
		g.coeff = st.dcoef;
		g.coeff_ld = coeff_ld;
		g.N = n;
		g.R = outputExponents.size();
		g.W = words;
		g.accumulate = st.overwrite ? 0 : 1;
		if (run_gf16_contraction(g, strm) != PARHIP_OK) setErr("gf16 contraction launch failed");
	} else {
		setErr(std::string("staging copy: ") + hipGetErrorString(e));
	}
	auto* ctx = new DoneCtx{this, a, n};
	e = hipLaunchHostFunc(strm, &Par2Engine::onBatchDone, ctx);
	if (e != hipSuccess) {
		// no host function: wait here (no lock held) and free the area
		delete ctx;
		setErr(std::string("hipLaunchHostFunc: ") + hipGetErrorString(e));
		hipStreamSynchronize(strm);
		finishArea(a, n);
	}
}

// Host function on the stream after a batch.  Runs on the runtime's callback
// thread, makes no HIP call and takes only the device queue's lock, which no
// thread holds across a call.
void Par2Engine::onBatchDone(void* p) {
	auto* ctx = static_cast<DoneCtx*>(p);
	ctx->e->devPush(DevEvent{1, ctx->area, ctx->n});
	delete ctx;
}

void Par2Engine::stageSlot(const Task& t) {
	Area& st = areas[t.area];
	uint16_t* slot = st.hsrc + (size_t)t.slot * strideWords;
	const size_t payloadWords = words - 1;
	size_t size = t.size;
	if (size > payloadWords * 2) size = payloadWords * 2;
	slot[payloadWords] = stage_copy_checksum(slot, t.buf, size, payloadWords);
}

// Take the next published slot of any area, starting at area `start`.
bool Par2Engine::takeSlot(unsigned start, Task* t) {
	const unsigned na = (unsigned)areas.size();
	for (unsigned k = 0; k < na; k++) {
		const unsigned a = (start + k) % na;
		uint64_t q = slotq[a].load();
		while ((uint32_t)q < (uint32_t)(q >> 32)) {
			if (slotq[a].compare_exchange_weak(q, q + 1)) {
				*t = areas[a].slots[(uint32_t)q];
				stageAvail--;
				return true;
			}
		}
	}
	return false;
}

void Par2Engine::worker(unsigned w) {
	hipSetDevice(dev);
	unsigned start = w;
	for (;;) {
		Task t;
		bool got = stageAvail.load() && takeSlot(start++, &t);
		if (got) helpWanted();
		if (!got && queued.load()) {
			std::lock_guard<std::mutex> lk(qmu);
			if (!queue.empty()) {
				t = queue.front();
				queue.pop_front();
				queued--;
				got = true;
			}
		}
		if (!got) {
			// poll for a while before blocking (a 32 KiB slot is staged in ~4 us;
			// batches arrive every few hundred us when the device is the limit)
			// (at most kPollers poll at once: polling threads compete with the
			// adding thread and the device thread for the cores)
			bool idle = spinners.fetch_add(1) >= kPollers;
			const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(300);
			for (unsigned i = 1; !idle && !stageAvail.load() && !queued.load() && !stopping.load(); i++) {
				_mm_pause();
				if (!(i & 63) && std::chrono::steady_clock::now() >= until) {
					idle = true;
					break;
				}
			}
			spinners--;
			// work appeared (another worker may take it first): look again; on
			// stopping, fall through to the exit check below
			if (!idle && !stopping.load()) continue;
			std::unique_lock<std::mutex> lk(qmu);
			if (!stageAvail.load() && queue.empty() && !stopping) {
				sleepers++;
				qcv.wait(lk, [&] { return stopping || stageAvail.load() || !queue.empty(); });
				sleepers--;
			}
			if (stopping && !stageAvail.load() && queue.empty()) return;
			continue;
		}
		if (t.kind != 1) {
			if (t.kind != 3) stageSlot(t);
			if (t.done) t.done(t.user);  // the caller's buffer may be reused
			// the last staged slot of an area takes the lock: the area may now
			// go to the device (if it is closed; else closing it submits it)
			if (pend[t.area].fetch_sub(1) == 1) {
				std::lock_guard<std::mutex> lk(mu);
				submitReady();
			}
		} else {
			{
				std::unique_lock<std::mutex> lk(mu);
				cv.wait(lk, [&] { return activeCount == 0; });
			}
			const bool valid = readOutput(t.index, t.out, hout.size() > w ? hout[w] : nullptr);
			if (t.valid) t.valid(t.user, valid ? 1 : 0);
		}
		if (tasksOpen.fetch_sub(1) == 1) {  // drain() may be waiting
			std::lock_guard<std::mutex> lk(mu);
			cv.notify_all();
		}
	}
}

void Par2Engine::endInput() {
	flush();
	std::unique_lock<std::mutex> lk(mu);
	cv.wait(lk, [&] { return activeCount == 0; });
}

// One recovery slice back through a pinned buffer, checksum verified there.
bool Par2Engine::readOutput(unsigned index, void* output, uint16_t* pinned) {
	bool added;
	{
		std::lock_guard<std::mutex> lk(mu);
		added = processingAdd;
	}
	if (index >= outputExponents.size()) return false;
	if (!added) {  // nothing added in this pass: all-zero recovery
		std::memset(output, 0, currentSliceSize);
		return true;
	}
	if (!pinned) {
		setErr("output: no read-back buffer");
		return false;
	}
	hipError_t e = hipMemcpy(pinned, dout + (size_t)index * strideWords, words * 2, hipMemcpyDeviceToHost);
	if (e != hipSuccess) {
		setErr(std::string("output copy: ") + hipGetErrorString(e));
		return false;
	}
	const size_t n = words - 1;
	std::memcpy(output, pinned, currentSliceSize);
	return slice_checksum(pinned, n) == pinned[n];
}

bool Par2Engine::getOutput(unsigned index, void* output) {
	{
		std::unique_lock<std::mutex> lk(mu);
		cv.wait(lk, [&] { return activeCount == 0; });
	}
	std::lock_guard<std::mutex> lk(hout_mu);
	return readOutput(index, output, hout.empty() ? nullptr : hout.back());
}

void Par2Engine::getOutputAsync(unsigned index, void* output, Par2ValidFn fn, void* user) {
	tasksOpen++;
	push(Task{1, 0, 0, 0, nullptr, 0, nullptr, fn, user, index, output});
}

void Par2Engine::discardOutput() {
	std::lock_guard<std::mutex> lk(mu);
	processingAdd = false;
}

uint16_t* Par2Engine::deviceOutput(unsigned index, size_t* strideBytes) const {
	if (strideBytes) *strideBytes = strideWords * sizeof(uint16_t);
	if (!dout || index >= outputExponents.size()) return nullptr;
	return dout + (size_t)index * strideWords;
}

void Par2Engine::freeProcessingMem() {
	std::unique_lock<std::mutex> lk(mu);
	drain(lk);
	hipSetDevice(dev);
	if (strm) hipStreamSynchronize(strm);
	if (dout) hipFree(dout);
	dout = nullptr;
	for (auto* h : hout) hipHostFree(h);
	hout.clear();
	allocOutputsCount = 0;
}

void Par2Engine::deinit() {
	{
		std::unique_lock<std::mutex> lk(mu);
		drain(lk);
	}
	{
		std::lock_guard<std::mutex> lk(qmu);
		stopping = true;
	}
	qcv.notify_all();
	for (auto& t : workers) t.join();
	workers.clear();
	if (devThread.joinable()) {
		devPush(DevEvent{2, 0, 0});
		devThread.join();
	}
	freeStaging();
	freeProcessingMem();
	if (dexps) hipFree(dexps);
	dexps = nullptr;
	if (strm) {
		hipStreamSynchronize(strm);
		hipStreamDestroy(strm);
	}
	if (cstrm) {
		hipStreamSynchronize(cstrm);
		hipStreamDestroy(cstrm);
	}
	strm = cstrm = nullptr;
}

}  // namespace parhip

using namespace parhip;

// ============================================================ C ABI
struct parhip_par2 {
	Par2Engine* be;
};

// Blocks a synchronous call until a worker reports its completion.
struct Par2Wait {
	std::mutex m;
	std::condition_variable c;
	bool done = false;
	int valid = 0;
	static void on_done(void* u) {
		auto* w = static_cast<Par2Wait*>(u);
		std::lock_guard<std::mutex> lk(w->m);
		w->done = true;
		w->c.notify_all();
	}
	static void on_valid(void* u, int v) {
		auto* w = static_cast<Par2Wait*>(u);
		std::lock_guard<std::mutex> lk(w->m);
		w->valid = v;
		w->done = true;
		w->c.notify_all();
	}
	void wait() {
		std::unique_lock<std::mutex> lk(m);
		c.wait(lk, [&] { return done; });
	}
};

extern "C" {

parhip_par2* parhip_par2_create_areas(int device, size_t sliceSize, unsigned inputBatchSize,
                                      unsigned stagingAreas) {
	if (stagingAreas > 64) {
		fail(PARHIP_ERR_RANGE, "par2_create: %u staging areas (at most 64)", stagingAreas);
		return nullptr;
	}
	auto* h = new (std::nothrow) parhip_par2;
	if (!h) return nullptr;
	h->be = new (std::nothrow) Par2Engine(device, stagingAreas ? stagingAreas : 2, 0);
	if (!h->be) {
		delete h;
		return nullptr;
	}
	if (sliceSize) h->be->setSliceSize(sliceSize);
	if (!h->be->init(inputBatchSize) || !h->be->ok()) {
		fail(PARHIP_ERR_DEVICE, "par2_create: %s", h->be->lastError().c_str());
		delete h->be;
		delete h;
		return nullptr;
	}
	return h;
}

parhip_par2* parhip_par2_create(int device, size_t sliceSize, unsigned inputBatchSize) {
	return parhip_par2_create_areas(device, sliceSize, inputBatchSize, 2);
}

void parhip_par2_destroy(parhip_par2* h) {
	if (!h) return;
	delete h->be;
	delete h;
}

static int be_status(parhip_par2* h) {
	if (!h->be->ok()) return fail(PARHIP_ERR_DEVICE, "par2: %s", h->be->lastError().c_str());
	return PARHIP_OK;
}

int parhip_par2_set_slice_size(parhip_par2* h, size_t sliceSize) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->setSliceSize(sliceSize);
	return be_status(h);
}

int parhip_par2_set_recovery_slices(parhip_par2* h, unsigned numSlices, const uint16_t* exponents) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	if (!h->be->setRecoverySlices(numSlices, exponents)) return be_status(h) ? be_status(h) : PARHIP_ERR_NOMEM;
	return PARHIP_OK;
}

int parhip_par2_set_current_slice_size(parhip_par2* h, size_t size) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->setCurrentSliceSize(size);
	return be_status(h);
}

int parhip_par2_can_add(parhip_par2* h) { return h ? (int)h->be->canAdd() : PARHIP_ERR_ARG; }

int parhip_par2_wait_for_add(parhip_par2* h) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->waitForAdd();
	return be_status(h);
}

int parhip_par2_add_input(parhip_par2* h, const void* buffer, size_t size, uint16_t inputNum,
                          int flush) {
	if (!h || (!buffer && size)) return fail(PARHIP_ERR_ARG, "par2_add_input: null pointer");
	Par2Wait w;
	if (h->be->addInput(buffer, size, inputNum, flush != 0, &Par2Wait::on_done, &w)) w.wait();
	return be_status(h);
}

int parhip_par2_add_input_coeffs(parhip_par2* h, const void* buffer, size_t size,
                                 const uint16_t* coeffs, int flush) {
	if (!h || (!buffer && size) || !coeffs)
		return fail(PARHIP_ERR_ARG, "par2_add_input_coeffs: null pointer");
	Par2Wait w;
	if (h->be->addInput(buffer, size, coeffs, flush != 0, &Par2Wait::on_done, &w)) w.wait();
	return be_status(h);
}

int parhip_par2_add_input_async(parhip_par2* h, const void* buffer, size_t size, uint16_t inputNum,
                                int flush, parhip_par2_done_fn done, void* user) {
	if (!h || (!buffer && size)) return fail(PARHIP_ERR_ARG, "par2_add_input_async: null pointer");
	h->be->addInput(buffer, size, inputNum, flush != 0, done, user);
	return be_status(h);
}

int parhip_par2_add_input_coeffs_async(parhip_par2* h, const void* buffer, size_t size,
                                       const uint16_t* coeffs, int flush, parhip_par2_done_fn done,
                                       void* user) {
	if (!h || (!buffer && size) || !coeffs)
		return fail(PARHIP_ERR_ARG, "par2_add_input_coeffs_async: null pointer");
	h->be->addInput(buffer, size, coeffs, flush != 0, done, user);
	return be_status(h);
}

int parhip_par2_set_batch_callbacks(parhip_par2* h, parhip_par2_batch_fn started,
                                    parhip_par2_batch_fn finished, void* user) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->setBatchCallbacks(started, finished, user);
	return PARHIP_OK;
}

int parhip_par2_dummy_input(parhip_par2* h, uint16_t inputNum, int flush) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->dummyInput(inputNum, flush != 0);
	return be_status(h);
}

int parhip_par2_fill_input(parhip_par2* h, const void* buffer) {
	if (!h || !buffer) return fail(PARHIP_ERR_ARG, "par2_fill_input: null pointer");
	const bool all = h->be->fillInput(buffer);
	const int rc = be_status(h);
	return rc ? rc : (all ? 1 : 0);
}

int parhip_par2_flush(parhip_par2* h) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->flush();
	return be_status(h);
}

int parhip_par2_end_input(parhip_par2* h) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->endInput();
	return be_status(h);
}

int parhip_par2_get_output(parhip_par2* h, unsigned index, void* output) {
	if (!h || !output) return fail(PARHIP_ERR_ARG, "par2_get_output: null pointer");
	if (index >= h->be->numRecovery())
		return fail(PARHIP_ERR_ARG, "par2_get_output: index %u out of range", index);
	bool valid = h->be->getOutput(index, output);
	int rc = be_status(h);
	return rc ? rc : (valid ? 1 : 0);
}

int parhip_par2_get_output_async(parhip_par2* h, unsigned index, void* output,
                                 parhip_par2_valid_fn fn, void* user) {
	if (!h || !output) return fail(PARHIP_ERR_ARG, "par2_get_output_async: null pointer");
	if (index >= h->be->numRecovery())
		return fail(PARHIP_ERR_ARG, "par2_get_output_async: index %u out of range", index);
	h->be->getOutputAsync(index, output, fn, user);
	return be_status(h);
}

int parhip_par2_discard_output(parhip_par2* h) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->discardOutput();
	return PARHIP_OK;
}

int parhip_par2_set_min_input_batch_size(parhip_par2* h, unsigned size) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->setMinInputBatchSize(size);
	return PARHIP_OK;
}

int parhip_par2_input_batch_size(parhip_par2* h) {
	return h ? (int)h->be->getInputBatchSize() : PARHIP_ERR_ARG;
}

int parhip_par2_batches_started(parhip_par2* h) {
	return h ? (int)h->be->getBatchesStarted() : PARHIP_ERR_ARG;
}

int parhip_par2_staging_areas(parhip_par2* h) {
	return h ? (int)h->be->stagingAreas() : PARHIP_ERR_ARG;
}

int parhip_par2_free_processing_mem(parhip_par2* h) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->freeProcessingMem();
	return be_status(h);
}

int parhip_par2_output_device_ptr(parhip_par2* h, unsigned index, void** d_row, size_t* strideBytes) {
	if (!h || !d_row) return fail(PARHIP_ERR_ARG, "par2_output_device_ptr: null pointer");
	*d_row = h->be->deviceOutput(index, strideBytes);
	if (!*d_row) return fail(PARHIP_ERR_ARG, "par2_output_device_ptr: no recovery slice %u", index);
	return PARHIP_OK;
}

}  // extern "C"
