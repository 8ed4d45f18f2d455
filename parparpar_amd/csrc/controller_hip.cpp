// controller_hip.cpp -- PAR2ProcHIP (see controller_hip.h) and its C ABI.
#include "controller_hip.h"

#include <emmintrin.h>

#include <algorithm>
#include <cstring>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "gf16_host.h"
#include "ingest.h"
#include "kernels.h"

namespace parhip {

int run_gf16_contraction(Gf16Args a, hipStream_t s);

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

// One run's Horner value (lanes), x per chunk.
static __m128i ck_run_value(const uint16_t* p, size_t n, const CkRun& k) {
	__m128i a = _mm_setzero_si128();
	for (size_t j = k.c0; j < k.c1; j++)
		a = ck_step(a, _mm_loadu_si128(reinterpret_cast<const __m128i*>(p + 8 * j)));
	if (k.tail) {
		uint16_t tail[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		std::memcpy(tail, p + 8 * k.c1, (n - 8 * k.c1) * 2);
		a = ck_step(a, _mm_loadu_si128(reinterpret_cast<const __m128i*>(tail)));
	}
	return a;
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

// Slices of this many bytes and more are copied and checksummed on the host
// copy threads.
static constexpr size_t kParallelBytes = 256u << 10;

static void copy_parallel(void* dst, const void* src, size_t bytes) {
	const size_t parts = 8;
	const size_t step = ((bytes + parts - 1) / parts + 15) & ~(size_t)15;
	parallel_parts(parts, [&](size_t q) {
		const size_t b0 = std::min(bytes, q * step), b1 = std::min(bytes, b0 + step);
		if (b1 > b0) std::memcpy(static_cast<char*>(dst) + b0, static_cast<const char*>(src) + b0, b1 - b0);
	});
}

static uint16_t checksum_parallel(const uint16_t* p, size_t n) {
	constexpr size_t parts = 8;
	__m128i v[parts];
	parallel_parts(parts, [&](size_t r) { v[r] = ck_run_value(p, n, ck_run(n, parts, r)); });
	return ck_join(v, parts, n);
}

// Stage a large slice: copy and checksum in parallel on the host copy threads
// (the same bytes and checksum as memcpy + slice_checksum).
static uint16_t stage_large(uint16_t* slot, const void* src, size_t size, size_t n,
                            size_t slot_bytes) {
	copy_parallel(slot, src, size);
	std::memset(reinterpret_cast<char*>(slot) + size, 0, slot_bytes - size);
	return checksum_parallel(slot, n);
}

static std::future<void> ready_future() {
	std::promise<void> p;
	p.set_value();
	return p.get_future();
}

PAR2ProcHIP::PAR2ProcHIP(int device, unsigned stagingAreas) : dev(device) {
	staging.resize(stagingAreas ? stagingAreas : 2);
	hipError_t e = hipSetDevice(dev);
	if (e == hipSuccess) e = hipStreamCreateWithFlags(&strm, hipStreamNonBlocking);
	if (e != hipSuccess) setErr("stream", e);
}

PAR2ProcHIP::~PAR2ProcHIP() { deinit(); }

void PAR2ProcHIP::setErr(const char* what, hipError_t e) {
	err = std::string(what) + ": " + hipGetErrorString(e);
}

bool PAR2ProcHIP::init(unsigned batch) {
	inputBatchSize = batch ? batch : 64;
	minInBatchSize = inputBatchSize;
	currentStagingArea = currentStagingInputs = 0;
	statBatchesStarted = 0;
	processingAdd = false;
	return allocStaging() && allocOutputs();
}

void PAR2ProcHIP::freeStaging() {
	hipSetDevice(dev);
	for (auto& st : staging) {
		if (st.done) hipEventSynchronize(st.done);
		if (st.hsrc) hipHostFree(st.hsrc);
		if (st.hcoef) hipHostFree(st.hcoef);
		if (st.dsrc) hipFree(st.dsrc);
		if (st.dcoef) hipFree(st.dcoef);
		if (st.done) hipEventDestroy(st.done);
		st = Staging();
	}
}

bool PAR2ProcHIP::allocStaging() {
	freeStaging();
	if (!allocWords || !inputBatchSize) return true;
	hipSetDevice(dev);
	const size_t nOut = std::max<size_t>(outputExponents.size(), 1);
	for (auto& st : staging) {
		const size_t sb = (size_t)inputBatchSize * strideWords * sizeof(uint16_t);
		const size_t cb = nOut * inputBatchSize * sizeof(uint16_t);
		hipError_t e;
		if ((e = hipHostMalloc((void**)&st.hsrc, sb, 0)) != hipSuccess ||
		    (e = hipHostMalloc((void**)&st.hcoef, cb, 0)) != hipSuccess ||
		    (e = hipMalloc((void**)&st.dsrc, sb)) != hipSuccess ||
		    (e = hipMalloc((void**)&st.dcoef, cb)) != hipSuccess ||
		    (e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming)) != hipSuccess) {
			setErr("staging allocation", e);
			return false;
		}
		std::memset(st.hsrc, 0, sb);
	}
	return true;
}

bool PAR2ProcHIP::allocOutputs() {
	hipSetDevice(dev);
	if (dout) {
		hipStreamSynchronize(strm);
		hipFree(dout);
		dout = nullptr;
	}
	if (hout) {
		hipHostFree(hout);
		hout = nullptr;
	}
	if (!allocWords || outputExponents.empty()) return true;
	hipError_t e = hipMalloc((void**)&dout, outputExponents.size() * strideWords * sizeof(uint16_t));
	if (e == hipSuccess) e = hipHostMalloc((void**)&hout, strideWords * sizeof(uint16_t), 0);
	if (e != hipSuccess) {
		setErr("output allocation", e);
		return false;
	}
	allocOutputsCount = outputExponents.size();
	return true;
}

void PAR2ProcHIP::setSliceSize(size_t size) {
	sliceSize = currentSliceSize = size;
	words = (size + 1) / 2 + 1;  // payload words (odd tail byte padded) + checksum word
	allocWords = words;
	strideWords = (words + 3) & ~(size_t)3;  // 8-byte aligned rows
	allocStaging();
	allocOutputs();
}

bool PAR2ProcHIP::setCurrentSliceSize(size_t size) {
	const size_t w = (size + 1) / 2 + 1;
	if (w > allocWords) {
		setSliceSize(size);
		return ok();
	}
	currentSliceSize = size;
	words = w;
	return true;
}

bool PAR2ProcHIP::setRecoverySlices(unsigned numSlices, const uint16_t* exponents) {
	hipSetDevice(dev);
	hipStreamSynchronize(strm);
	outputExponents.assign(numSlices, 1);  // no exponents: custom coefficients (controller_cpu.cpp:176)
	if (exponents) std::copy(exponents, exponents + numSlices, outputExponents.begin());
	processingAdd = false;
	bool okay = true;
	if (numSlices > allocOutputsCount || !dout) okay = allocOutputs() && allocStaging();
	return okay && ok();
}

PAR2ProcBackendAddResult PAR2ProcHIP::canAdd() const {
	const Staging& st = staging[currentStagingArea];
	if (st.active && hipEventQuery(st.done) == hipErrorNotReady) return PROC_ADD_FULL;
	for (const auto& s : staging)
		if (s.active && hipEventQuery(s.done) == hipErrorNotReady) return PROC_ADD_OK_BUSY;
	return PROC_ADD_OK;
}

void PAR2ProcHIP::waitForAdd() {
	Staging& st = staging[currentStagingArea];
	if (st.active) {
		hipEventSynchronize(st.done);
		st.active = false;
	}
}

void PAR2ProcHIP::setCoeffs(Staging& st, unsigned idx, uint16_t inputNum) {
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

void PAR2ProcHIP::setCoeffs(Staging& st, unsigned idx, const uint16_t* coeffs) {
	for (size_t e = 0; e < outputExponents.size(); e++)
		st.hcoef[e * inputBatchSize + idx] = coeffs[e];
}

template <class T>
std::future<void> PAR2ProcHIP::addInputT(const void* buffer, size_t size, T numOrCoeffs,
                                         bool flushNow) {
	waitForAdd();  // the area we are about to fill must not be in flight
	Staging& st = staging[currentStagingArea];
	uint16_t* slot = st.hsrc + (size_t)currentStagingInputs * strideWords;
	const size_t payloadWords = words - 1;
	if (size > payloadWords * 2) size = payloadWords * 2;
	if (size >= kParallelBytes) {
		slot[payloadWords] = stage_large(slot, buffer, size, payloadWords, words * 2);
	} else {
		std::memcpy(slot, buffer, size);
		std::memset(reinterpret_cast<char*>(slot) + size, 0, (words * 2) - size);
		slot[payloadWords] = slice_checksum(slot, payloadWords);
	}
	setCoeffs(st, currentStagingInputs, numOrCoeffs);
	currentStagingInputs++;
	if (currentStagingInputs >= inputBatchSize || flushNow) flush();
	return ready_future();  // the caller's buffer has been copied: reusable now
}

std::future<void> PAR2ProcHIP::addInput(const void* buffer, size_t size, uint16_t inputNum,
                                        bool flushNow) {
	return addInputT(buffer, size, inputNum, flushNow);
}

std::future<void> PAR2ProcHIP::addInput(const void* buffer, size_t size, const uint16_t* coeffs,
                                        bool flushNow) {
	return addInputT(buffer, size, coeffs, flushNow);
}

void PAR2ProcHIP::dummyInput(uint16_t inputNum, bool flushNow) {
	waitForAdd();
	setCoeffs(staging[currentStagingArea], currentStagingInputs, inputNum);
	currentStagingInputs++;
	if (currentStagingInputs >= inputBatchSize || flushNow) flush();
}

// Benchmarking (gf16/controller.h:303-305): fill the next staging slot with
// `buffer` (the current slice size), as PAR2ProcCPU::fillInput does
// (gf16/controller_cpu.cpp:357-369); true once every slot of every staging
// area holds it, after which dummyInput() adds inputs over that data.
bool PAR2ProcHIP::fillInput(const void* buffer) {
	Staging& st = staging[currentStagingArea];
	if (!st.hsrc) {
		err = "fillInput: no staging memory";
		return true;
	}
	if (st.active) {
		hipEventSynchronize(st.done);
		st.active = false;
	}
	uint16_t* slot = st.hsrc + (size_t)currentStagingInputs * strideWords;
	const size_t payloadWords = words - 1;
	std::memcpy(slot, buffer, currentSliceSize);
	std::memset(reinterpret_cast<char*>(slot) + currentSliceSize, 0, words * 2 - currentSliceSize);
	slot[payloadWords] = slice_checksum(slot, payloadWords);
	if (++currentStagingInputs == inputBatchSize) {
		currentStagingInputs = 0;
		if (++currentStagingArea == staging.size()) {
			currentStagingArea = 0;
			return true;  // all filled
		}
	}
	return false;
}

void PAR2ProcHIP::flush() {
	if (!currentStagingInputs) return;
	run_kernel(currentStagingArea, currentStagingInputs);
	currentStagingArea = (currentStagingArea + 1) % staging.size();
	currentStagingInputs = 0;
}

void PAR2ProcHIP::run_kernel(unsigned area, unsigned numInputs) {
	Staging& st = staging[area];
	hipSetDevice(dev);
	hipError_t e = hipMemcpyAsync(st.dsrc, st.hsrc, (size_t)numInputs * strideWords * 2,
	                              hipMemcpyHostToDevice, strm);
	if (e == hipSuccess)
		e = hipMemcpyAsync(st.dcoef, st.hcoef, outputExponents.size() * inputBatchSize * 2,
		                   hipMemcpyHostToDevice, strm);
	if (e != hipSuccess) {
		setErr("staging copy", e);
		return;
	}
	Gf16Args a{};
	a.in_base = st.dsrc;
	a.in_stride = strideWords;
	a.out_base = dout;
	a.out_stride = strideWords;
	a.coeff = st.dcoef;
	a.coeff_ld = inputBatchSize;
	a.N = numInputs;
	a.R = outputExponents.size();
	a.W = words;
	a.accumulate = processingAdd ? 1 : 0;
	if (run_gf16_contraction(a, strm) != PARHIP_OK) err = "gf16 contraction launch failed";
	processingAdd = true;
	hipEventRecord(st.done, strm);
	st.active = true;
	statBatchesStarted++;
}

std::future<void> PAR2ProcHIP::endInput() {
	flush();
	hipStreamSynchronize(strm);
	for (auto& st : staging) st.active = false;
	return ready_future();
}

std::future<bool> PAR2ProcHIP::getOutput(unsigned index, void* output) {
	std::promise<bool> pr;
	hipSetDevice(dev);
	hipStreamSynchronize(strm);
	if (index >= outputExponents.size()) {
		pr.set_value(false);
		return pr.get_future();
	}
	if (!processingAdd) {  // nothing added in this pass: all-zero recovery
		std::memset(output, 0, currentSliceSize);
		pr.set_value(true);
		return pr.get_future();
	}
	// through a pinned buffer (DMA at full rate), checksum verified there
	hipError_t e = hipMemcpy(hout, dout + (size_t)index * strideWords, words * 2,
	                         hipMemcpyDeviceToHost);
	if (e != hipSuccess) {
		setErr("output copy", e);
		pr.set_value(false);
		return pr.get_future();
	}
	const size_t n = words - 1;
	bool valid;
	if (currentSliceSize >= kParallelBytes) {
		copy_parallel(output, hout, currentSliceSize);
		valid = checksum_parallel(hout, n) == hout[n];
	} else {
		std::memcpy(output, hout, currentSliceSize);
		valid = slice_checksum(hout, n) == hout[n];
	}
	pr.set_value(valid);
	return pr.get_future();
}

void PAR2ProcHIP::freeProcessingMem() {
	hipSetDevice(dev);
	hipStreamSynchronize(strm);
	if (dout) hipFree(dout);
	dout = nullptr;
	if (hout) hipHostFree(hout);
	hout = nullptr;
	allocOutputsCount = 0;
}

void PAR2ProcHIP::deinit() {
	if (strm) hipStreamSynchronize(strm);
	freeStaging();
	freeProcessingMem();
	if (strm) hipStreamDestroy(strm);
	strm = nullptr;
}

}  // namespace parhip

using namespace parhip;

// ============================================================ C ABI
struct parhip_par2 {
	PAR2ProcHIP* be;
};

extern "C" {

parhip_par2* parhip_par2_create(int device, size_t sliceSize, unsigned inputBatchSize) {
	auto* h = new (std::nothrow) parhip_par2;
	if (!h) return nullptr;
	h->be = new (std::nothrow) PAR2ProcHIP(device, 2);
	if (!h->be) {
		delete h;
		return nullptr;
	}
	h->be->setSliceSize(sliceSize);
	if (!h->be->init(inputBatchSize) || !h->be->ok()) {
		fail(PARHIP_ERR_DEVICE, "par2_create: %s", h->be->lastError().c_str());
		delete h->be;
		delete h;
		return nullptr;
	}
	return h;
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

int parhip_par2_add_input(parhip_par2* h, const void* buffer, size_t size, uint16_t inputNum,
                          int flush) {
	if (!h || (!buffer && size)) return fail(PARHIP_ERR_ARG, "par2_add_input: null pointer");
	h->be->addInput(buffer, size, inputNum, flush != 0).get();
	return be_status(h);
}

int parhip_par2_add_input_coeffs(parhip_par2* h, const void* buffer, size_t size,
                                 const uint16_t* coeffs, int flush) {
	if (!h || (!buffer && size) || !coeffs)
		return fail(PARHIP_ERR_ARG, "par2_add_input_coeffs: null pointer");
	h->be->addInput(buffer, size, coeffs, flush != 0).get();
	return be_status(h);
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
	h->be->endInput().get();
	return be_status(h);
}

int parhip_par2_get_output(parhip_par2* h, unsigned index, void* output) {
	if (!h || !output) return fail(PARHIP_ERR_ARG, "par2_get_output: null pointer");
	if ((int)index >= h->be->getNumRecoverySlices())
		return fail(PARHIP_ERR_ARG, "par2_get_output: index %u out of range", index);
	bool valid = h->be->getOutput(index, output).get();
	int rc = be_status(h);
	return rc ? rc : (valid ? 1 : 0);
}

int parhip_par2_discard_output(parhip_par2* h) {
	if (!h) return fail(PARHIP_ERR_ARG, "par2: null handle");
	h->be->discardOutput();
	return PARHIP_OK;
}

int parhip_par2_input_batch_size(parhip_par2* h) {
	return h ? (int)h->be->getInputBatchSize() : PARHIP_ERR_ARG;
}

}  // extern "C"
