// controller_hip.h -- the PAR2 GF(2^16) recovery engine on one MI355X, behind
// the parhip_par2_* C ABI (include/parhip_par2.h).
//
// It implements the staging model of the reference's PAR2 backends
// (IPAR2ProcBackend, gf16/controller.h:77-217; PAR2ProcOCL,
// gf16/controller_ocl.cpp:382-412): inputs go into a slot of the current
// staging area, an area goes to the device as one batch when it is full or
// flushed, the first batch of a pass overwrites the recovery slices and later
// batches XOR into them (gf16/controller_ocl.cpp:626-634), and each recovery
// slice comes back with its checksum verified (cksumValid).
//
// Staging runs on a pool of copy workers: addInput only claims a slot, notes
// its input number and publishes it; a worker takes the next published slot
// (a compare-and-swap, no lock), copies the caller's bytes into pinned memory,
// appends the slot's positional checksum word and reports that the caller's
// buffer may be reused.  Once every slot of a closed area is staged, the area
// is handed, in the order the areas were closed, to the device thread, which
// enqueues host-to-device copies, the batch's coefficients computed on the
// device from the input numbers, and the contraction on the engine's stream,
// then a host function that reports when the batch has finished.  The host
// function only queues that report for the device thread, which frees the
// area: no thread holds the staging lock across a HIP call, and the runtime's
// callback thread never waits for that lock.  Reads of recovery slices run on
// the workers, each into its own pinned buffer.
//
// The reference-side backend class (parparpar_amd/reference_backend/) derives
// from IPAR2ProcBackend and drives this engine through the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace parhip {

// Values as PROC_ADD_* in gf16/controller.h:41-46.
enum Par2AddState { PAR2_ADD_OK = 0, PAR2_ADD_OK_BUSY = 1, PAR2_ADD_FULL = 2 };

typedef void (*Par2DoneFn)(void* user);
typedef void (*Par2ValidFn)(void* user, int valid);
typedef void (*Par2BatchFn)(void* user, unsigned area, unsigned numInputs);

class Par2Engine {
public:
	Par2Engine(int device, unsigned stagingAreas, unsigned workers);
	~Par2Engine();
	Par2Engine(const Par2Engine&) = delete;
	Par2Engine& operator=(const Par2Engine&) = delete;

	// inputBatchSize = inputs per batch (0: chosen from the slice size,
	// about 64 MiB per batch, 16 to 512 inputs)
	bool init(unsigned inputBatchSize);
	void setSliceSize(size_t size);
	bool setCurrentSliceSize(size_t size);
	// exponents null: custom coefficients per input (gf16/controller_cpu.cpp:176)
	bool setRecoverySlices(unsigned numSlices, const uint16_t* exponents);
	unsigned numRecovery() const { return (unsigned)outputExponents.size(); }
	Par2AddState canAdd();
	void waitForAdd();
	// Claim a slot and publish its copy to the workers; done(user) fires on a
	// worker once `buffer` may be reused (done may be null).  Blocks only while
	// the current staging area is still on the device.
	bool addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush, Par2DoneFn done,
	              void* user);
	bool addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush,
	              Par2DoneFn done, void* user);
	void dummyInput(uint16_t inputNum, bool flush);
	bool fillInput(const void* buffer);
	void flush();
	// Flush, then wait until every batch has been staged and processed.
	void endInput();
	// Recovery slice `index` into output (current slice size bytes); returns
	// cksumValid.  The asynchronous form calls fn(user, valid) from a worker.
	bool getOutput(unsigned index, void* output);
	void getOutputAsync(unsigned index, void* output, Par2ValidFn fn, void* user);
	void discardOutput();
	// started(user, area, n): a batch of n inputs was closed (on the calling
	// thread, inside the call that closed it); finished(user, area, n): the
	// batch finished on the device and its area is free (device thread).
	void setBatchCallbacks(Par2BatchFn started, Par2BatchFn finished, void* user);
	void setMinInputBatchSize(unsigned size) {
		autoMin = !size;
		minInBatchSize = size ? size : inputBatchSize;
	}
	unsigned getInputBatchSize() const { return inputBatchSize; }
	unsigned getBatchesStarted() const { return statBatchesStarted; }
	unsigned stagingAreas() const { return (unsigned)areas.size(); }
	// Device address and row stride of recovery slice `index` (checksum word last).
	uint16_t* deviceOutput(unsigned index, size_t* strideBytes) const;
	void freeProcessingMem();
	void deinit();
	bool ok();
	std::string lastError();

private:
	struct Task {
		// 0 stage, coefficients from inputNum; 2 stage, caller's coefficients;
		// 3 dummy slot (data from fillInput); 1 read a recovery slice back
		int kind;
		unsigned area, slot;
		uint16_t inputNum;
		const void* buf;
		size_t size;
		Par2DoneFn done;
		Par2ValidFn valid;
		void* user;
		unsigned index;
		void* out;
	};
	struct Area {
		uint16_t* hsrc = nullptr;   // pinned: inputBatchSize x strideWords
		uint16_t* hcoef = nullptr;  // pinned: numOutputs x inputBatchSize (custom coefficients)
		uint16_t* hnum = nullptr;   // pinned: input number of each slot
		uint16_t* dsrc = nullptr;
		uint16_t* dcoef = nullptr;
		uint16_t* dnum = nullptr;
		hipEvent_t copied = nullptr;  // the batch's host-to-device copies are done
		std::vector<Task> slots;    // what each claimed slot stages (kind 3: no data)
		bool custom = false;        // a slot has caller-given coefficients
		unsigned count = 0;         // slots claimed (slots still being staged: pend[area])
		bool closed = false;        // no more slots; goes to the device once staged
		bool active = false;        // closed and not yet finished on the device
		bool overwrite = false;     // first batch of a pass
		bool submitted = false;     // on the stream
		uint64_t seq = 0;           // close order
	};
	struct DoneCtx {
		Par2Engine* e;
		unsigned area, n;
	};
	// work for the device thread: 0 submit an area, 1 an area finished, 2 stop
	struct DevEvent {
		int kind;
		unsigned area, n;
	};

	bool claimSlot(bool flush, unsigned* area, unsigned* slot, bool* closed, unsigned* closedN,
	               std::unique_lock<std::mutex>& lk);
	void closeArea(unsigned a);
	void submitReady();  // with mu held: hands ready areas to the device thread
	void submit(unsigned a);  // device thread, no lock held
	void finishArea(unsigned a, unsigned n);
	void deviceLoop();
	void devPush(const DevEvent& ev);
	static void onBatchDone(void* ctx);
	void worker(unsigned w);
	void stageSlot(const Task& t);
	bool readOutput(unsigned index, void* output, uint16_t* pinned);
	void setCoeffs(Area& st, unsigned idx, uint16_t inputNum);
	void setCoeffs(Area& st, unsigned idx, const uint16_t* coeffs);
	bool allocStaging();
	void freeStaging();
	bool allocOutputs();
	void drain(std::unique_lock<std::mutex>& lk);
	void push(const Task& t);
	void publish(unsigned a, unsigned slot, const Task& t);  // mu held
	void wake();
	void helpWanted();  // a worker took a slot: wake another if slots wait
	bool takeSlot(unsigned start, Task* t);
	void setErr(const std::string& what);

	int dev;
	hipStream_t strm = nullptr;   // coefficients, contraction, host functions
	hipStream_t cstrm = nullptr;  // host-to-device copies of the staging areas
	std::vector<Area> areas;
	// per area: slots claimed but not yet staged; and one word holding the
	// slots published to the workers (high half) and taken by a worker (low
	// half), taken with a compare-and-swap on the whole word so that a worker
	// holding a stale value from an earlier batch of the area cannot succeed
	std::unique_ptr<std::atomic<unsigned>[]> pend;
	std::unique_ptr<std::atomic<uint64_t>[]> slotq;
	std::atomic<unsigned> stageAvail{0};    // published slots not yet taken (all areas)
	uint16_t* dexps = nullptr;              // device copy of the exponents
	std::vector<uint16_t> outputExponents;
	uint16_t* dout = nullptr;               // numOutputs x strideWords
	std::vector<uint16_t*> hout;            // pinned read-back buffer per worker, + one for getOutput
	std::mutex hout_mu;                     // the synchronous getOutput buffer
	size_t sliceSize = 0, currentSliceSize = 0;
	size_t allocWords = 0;                  // per-slice words allocated (payload + checksum)
	size_t words = 0;                       // current payload words + 1 checksum word
	size_t strideWords = 0;
	size_t allocOutputsCount = 0;
	unsigned inputBatchSize = 64, minInBatchSize = 64;
	bool autoBatch = false, autoMin = true;  // batch size from the slice size; minimum = batch size
	static unsigned batchFor(size_t slice);
	unsigned currentArea = 0;
	unsigned fillArea = 0, fillSlot = 0;    // fillInput's position
	unsigned statBatchesStarted = 0;
	unsigned activeCount = 0;
	uint64_t closeSeq = 0, submitSeq = 0;
	bool processingAdd = false;             // a batch of this pass has been closed
	Par2BatchFn cbStarted = nullptr, cbFinished = nullptr;
	void* cbUser = nullptr;

	std::mutex mu;                          // staging state
	std::condition_variable cv;             // area freed / slot staged / task done
	std::atomic<unsigned> tasksOpen{0};     // queued or running worker tasks
	std::mutex emu;
	std::string err;

	// workers
	std::vector<std::thread> workers;
	std::mutex qmu;
	std::condition_variable qcv;
	std::deque<Task> queue;
	std::atomic<unsigned> queued{0};        // tasks in `queue` (read while spinning)
	std::atomic<unsigned> sleepers{0};      // workers blocked on qcv
	std::atomic<unsigned> spinners{0};      // workers polling for work
	std::atomic<bool> stopping{false};

	// device thread (lock order: mu before dmu; dmu is never held across a call)
	std::thread devThread;
	std::mutex dmu;
	std::condition_variable dcv;
	std::deque<DevEvent> devq;
};

}  // namespace parhip
