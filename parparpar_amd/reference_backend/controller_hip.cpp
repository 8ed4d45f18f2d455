// controller_hip.cpp -- PAR2ProcHIP, the reference controller's backend over
// the MI355X PAR2 engine (see controller_hip.h).
#include "controller_hip.h"

#include <algorithm>
#include <stdexcept>
#include <string>

#ifdef USE_LIBUV
struct hip_add_req {  // an add whose buffer is still being copied
	PAR2ProcHIP* parent;
	PAR2ProcPlainCb cb;
};
struct hip_out_req {  // a recovery slice being copied out
	PAR2ProcHIP* parent;
	PAR2ProcOutputCb cb;
	bool valid;
};
#endif
typedef struct PAR2ProcBackendBaseComputeReq<PAR2ProcHIP> hip_compute_req;

static std::runtime_error engine_error(const char* what) {
	return std::runtime_error(std::string("PAR2ProcHIP: ") + what + ": " + parhip_last_error());
}

PAR2ProcHIP::PAR2ProcHIP(IF_LIBUV(uv_loop_t* _loop, ) int _deviceId, int stagingAreas)
    : IPAR2ProcBackend(IF_LIBUV(_loop)), eng(nullptr), deviceId(_deviceId), sliceSize(0), minSent(0),
      staging(stagingAreas > 0 ? stagingAreas : 2) {
	processingAdd = false;
	currentStagingArea = currentStagingInputs = 0;
	inputBatchSize = minInBatchSize = 0;
	statBatchesStarted = 0;
}

PAR2ProcHIP::~PAR2ProcHIP() {
	deinit();
	_deinit();  // (libuv: deinit() returns early once the loop was detached)
}

bool PAR2ProcHIP::init(unsigned targetInputBatch) {
	if (eng) _deinit();
	eng = parhip_par2_create_areas(deviceId, sliceSize, targetInputBatch, (unsigned)staging.size());
	if (!eng) return false;
	if (parhip_par2_set_batch_callbacks(eng, &PAR2ProcHIP::onStarted, &PAR2ProcHIP::onFinished, this))
		return false;
	refreshBatchSize();
	// every area starts free (the base class's staging promise starts unkept),
	// as the reference backends reset them (gf16/controller_ocl.cpp:241-246)
	IF_LIBUV(stagingActiveCount = 0);
	IF_NOT_LIBUV(stagingActiveCount.store(0));
	for (auto& area : staging) area.setIsActive(false);
	currentStagingArea = currentStagingInputs = 0;
	processingAdd = false;
	return true;
}

void PAR2ProcHIP::refreshBatchSize() {
	const int b = parhip_par2_input_batch_size(eng);
	if (b > 0) inputBatchSize = minInBatchSize = minSent = (unsigned)b;
}

void PAR2ProcHIP::_deinit() {
	if (eng) parhip_par2_destroy(eng);  // waits for every batch and copy in flight
	eng = nullptr;
}

void PAR2ProcHIP::setSliceSize(size_t size) {
	sliceSize = size;
	if (!eng) return;  // the usual order: the controller's init, then init()
	if (parhip_par2_set_slice_size(eng, size)) throw engine_error("setSliceSize");
	refreshBatchSize();
}

bool PAR2ProcHIP::setCurrentSliceSize(size_t size) {
	return eng && parhip_par2_set_current_slice_size(eng, size) == 0;
}

bool PAR2ProcHIP::setRecoverySlices(unsigned numSlices, const uint16_t* exponents) {
	if (!eng) return false;
	outputExponents.assign(numSlices, 0);
	if (exponents) std::copy(exponents, exponents + numSlices, outputExponents.begin());
	processingAdd = false;
	return parhip_par2_set_recovery_slices(eng, numSlices, exponents) == 0;
}

// The same decision as the reference backends (gf16/controller_ocl.cpp:372-375).
PAR2ProcBackendAddResult PAR2ProcHIP::canAdd() const {
	if (staging[currentStagingArea].getIsActive()) return PROC_ADD_FULL;
	return stagingActiveCount_get() < staging.size() - 1 ? PROC_ADD_OK : PROC_ADD_OK_BUSY;
}

void PAR2ProcHIP::syncBase() {
	// discardOutput() clears processingAdd in the base class only; the engine
	// starts its pass over (the next batch overwrites) when told here.  While
	// processingAdd is false no batch of this pass has started, so this is
	// also harmless when the engine never had one.
	if (!processingAdd) parhip_par2_discard_output(eng);
	if (minInBatchSize != minSent) {
		parhip_par2_set_min_input_batch_size(eng, minInBatchSize);
		minSent = minInBatchSize;
	}
}

// A batch of numInputs inputs was closed in `area` (adding thread).
void PAR2ProcHIP::onStarted(void* p, unsigned area, unsigned numInputs) {
	auto* self = static_cast<PAR2ProcHIP*>(p);
	(void)numInputs;
	self->stagingActiveCount_inc();
#ifdef USE_LIBUV
	self->staging[area].setIsActive(true);
#else
	// the engine freed the area before it reports the finish: wait for that
	// report, so that the area's promise is replaced only after it was kept
	self->staging[area].promFuture.wait();
	{
		std::lock_guard<std::mutex> lk(self->areaMu);
		self->staging[area].setIsActive(true);
	}
#endif
	self->statBatchesStarted++;
	self->processingAdd = true;
	self->currentStagingInputs = 0;
	self->currentStagingArea = (area + 1) % (unsigned)self->staging.size();
}

// The batch in `area` finished on the device; the area is free (device thread).
void PAR2ProcHIP::onFinished(void* p, unsigned area, unsigned numInputs) {
	auto* self = static_cast<PAR2ProcHIP*>(p);
#ifdef USE_LIBUV
	auto* req = new hip_compute_req;
	req->numInputs = (uint16_t)numInputs;
	req->procIdx = area;
	req->parent = self;
	self->_queueProc.notify(req);
#else
	(void)numInputs;
	self->stagingActiveCount_dec();
	std::lock_guard<std::mutex> lk(self->areaMu);
	self->staging[area].setIsActive(false);
#endif
}

void PAR2ProcHIP::run_kernel(unsigned, unsigned) {
	// never called: the engine submits each staged batch itself
}

#ifdef USE_LIBUV
// Loop thread: an area is free; adds may continue (gf16/controller_ocl.cpp:595-604).
void PAR2ProcHIP::_notifyProc(void* _req) {
	auto* req = static_cast<hip_compute_req*>(_req);
	stagingActiveCount_dec();
	staging[req->procIdx].setIsActive(false);
	if (progressCb) progressCb(req->numInputs);
	delete req;
}

// Loop thread: an added buffer may be reused (gf16/controller_ocl.cpp:361-369).
void PAR2ProcHIP::_notifySent(void* _req) {
	auto* req = static_cast<hip_add_req*>(_req);
	pendingInCallbacks--;
	if (req->cb) req->cb();
	delete req;
	// this may arrive after the last batch's finish: then the controller
	// learns here that everything is done
	if (endSignalled && isEmpty() && progressCb) progressCb(0);
}

// Loop thread: a recovery slice was copied out (gf16/controller_ocl.cpp:550-557).
void PAR2ProcHIP::_notifyRecv(void* _req) {
	auto* req = static_cast<hip_out_req*>(_req);
	pendingOutCallbacks--;
	req->cb(req->valid);
	delete req;
	if (pendingOutCallbacks < 1 && deinitCallback) deinit(deinitCallback);
}

// Copy-worker threads: hand the completion to the loop thread.
void PAR2ProcHIP::onAdded(void* p) {
	auto* req = static_cast<hip_add_req*>(p);
	req->parent->_queueSent.notify(req);
}

void PAR2ProcHIP::onOutput(void* p, int valid) {
	auto* req = static_cast<hip_out_req*>(p);
	req->valid = valid != 0;
	req->parent->_queueRecv.notify(req);
}

void PAR2ProcHIP::addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush,
                           const PAR2ProcPlainCb& cb) {
	syncBase();
	auto* req = new hip_add_req{this, cb};
	pendingInCallbacks++;
	if (parhip_par2_add_input_async(eng, buffer, size, inputNum, flush ? 1 : 0, onAdded, req)) {
		pendingInCallbacks--;
		delete req;
		throw engine_error("addInput");
	}
}

void PAR2ProcHIP::addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush,
                           const PAR2ProcPlainCb& cb) {
	syncBase();
	auto* req = new hip_add_req{this, cb};
	pendingInCallbacks++;
	if (parhip_par2_add_input_coeffs_async(eng, buffer, size, coeffs, flush ? 1 : 0, onAdded, req)) {
		pendingInCallbacks--;
		delete req;
		throw engine_error("addInput");
	}
}

void PAR2ProcHIP::getOutput(unsigned index, void* output, const PAR2ProcOutputCb& cb) {
	auto* req = new hip_out_req{this, cb, false};
	pendingOutCallbacks++;
	if (parhip_par2_get_output_async(eng, index, output, onOutput, req)) {
		pendingOutCallbacks--;
		delete req;
		throw engine_error("getOutput");
	}
}

#else  // futures

// Copy-worker threads: keep the promise the caller holds the future of.
void PAR2ProcHIP::onAdded(void* p) {
	auto* prom = static_cast<std::promise<void>*>(p);
	prom->set_value();
	delete prom;
}

void PAR2ProcHIP::onOutput(void* p, int valid) {
	auto* prom = static_cast<std::promise<bool>*>(p);
	prom->set_value(valid != 0);
	delete prom;
}

std::future<void> PAR2ProcHIP::addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush) {
	syncBase();
	auto* prom = new std::promise<void>();
	auto fut = prom->get_future();
	if (parhip_par2_add_input_async(eng, buffer, size, inputNum, flush ? 1 : 0, onAdded, prom)) {
		prom->set_exception(std::make_exception_ptr(engine_error("addInput")));
		delete prom;
	}
	return fut;
}

std::future<void> PAR2ProcHIP::addInput(const void* buffer, size_t size, const uint16_t* coeffs,
                                        bool flush) {
	syncBase();
	auto* prom = new std::promise<void>();
	auto fut = prom->get_future();
	if (parhip_par2_add_input_coeffs_async(eng, buffer, size, coeffs, flush ? 1 : 0, onAdded, prom)) {
		prom->set_exception(std::make_exception_ptr(engine_error("addInput")));
		delete prom;
	}
	return fut;
}

std::future<bool> PAR2ProcHIP::getOutput(unsigned index, void* output) {
	auto* prom = new std::promise<bool>();
	auto fut = prom->get_future();
	if (parhip_par2_get_output_async(eng, index, output, onOutput, prom)) {
		prom->set_exception(std::make_exception_ptr(engine_error("getOutput")));
		delete prom;
	}
	return fut;
}

// Every area's current batch finished, then processing_finished (the base
// class's helper, gf16/controller.h:118-130); the controller flushed first.
std::future<void> PAR2ProcHIP::endInput() {
	return IPAR2ProcBackend::_endInput(staging);
}

void PAR2ProcHIP::waitForAdd() {
	IPAR2ProcBackend::_waitForAdd(staging[currentStagingArea]);
}
#endif

void PAR2ProcHIP::dummyInput(uint16_t inputNum, bool flush) {
	syncBase();
	if (parhip_par2_dummy_input(eng, inputNum, flush ? 1 : 0)) throw engine_error("dummyInput");
}

bool PAR2ProcHIP::fillInput(const void* buffer) {
	const int rc = parhip_par2_fill_input(eng, buffer);
	if (rc < 0) throw engine_error("fillInput");
	return rc == 1;
}

void PAR2ProcHIP::flush() {
	syncBase();
	if (parhip_par2_flush(eng)) throw engine_error("flush");
}

void PAR2ProcHIP::processing_finished() {
	IF_LIBUV(endSignalled = false);
}

void PAR2ProcHIP::freeProcessingMem() {
	if (eng) parhip_par2_free_processing_mem(eng);
}
