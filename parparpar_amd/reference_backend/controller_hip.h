// controller_hip.h -- PAR2ProcHIP: a PAR2 backend of the reference's
// controller (class IPAR2ProcBackend, gf16/controller.h:77-217) over the
// MI355X PAR2 engine's C ABI (include/parhip_par2.h, libparhip.so).
//
// The class is written against the reference's headers: a maintainer adds
// controller_hip.h and controller_hip.cpp to the reference's gf16/ directory,
// beside controller_ocl.{h,cpp}, lists controller_hip.cpp in the addon's
// sources next to the OpenCL backend (binding.gyp:78) and links libparhip.so.
// It builds in both flavours the reference builds its backends in: futures
// (the default) and USE_LIBUV (completions delivered on the loop thread
// through the base class's ThreadNotifyQueue hooks).
//
// The engine does what the reference's GPU backend does on its transfer
// thread and device queue (PAR2ProcOCL, gf16/controller_ocl.cpp:382-540): an
// add claims a slot of the current staging area and returns; a copy worker
// copies the slice into pinned memory with its checksum word and reports the
// caller's buffer free; full or flushed areas go to the device in order.
// This class keeps the base class's bookkeeping in step with the engine.  The
// engine reports a batch start on the adding thread, inside the call that
// closed the area (stagingActiveCount up, the area active, statBatchesStarted,
// processingAdd), and a batch finish on its device thread (count down, the
// area free; in the libuv flavour through _queueProc to the loop thread,
// followed by progressCb as the reference backends do).
#ifndef __GF16_CONTROLLER_HIP
#define __GF16_CONTROLLER_HIP

#include "controller.h"
#include "parhip.h"
#include "parhip_par2.h"
#include <mutex>

class PAR2ProcHIP : public IPAR2ProcBackend {
	parhip_par2* eng;
	int deviceId;
	size_t sliceSize;
	unsigned minSent;  // minInBatchSize as last passed to the engine
	// the engine's staging areas as the base class sees them; an area is
	// active from the call that closed it until its batch finished
	std::vector<IPAR2ProcStaging> staging;
#ifndef USE_LIBUV
	std::mutex areaMu;  // an area's promise between the adding and the device thread
#endif

	static void onStarted(void* self, unsigned area, unsigned numInputs);
	static void onFinished(void* self, unsigned area, unsigned numInputs);
	static void onAdded(void* req);               // an added buffer may be reused
	static void onOutput(void* req, int valid);   // a recovery slice was copied out
	// state the base class changes without a virtual call (discardOutput,
	// setMinInputBatchSize), passed on before the engine takes an input
	void syncBase();
	void refreshBatchSize();

	// batches are submitted by the engine (its device thread) once staged
	void run_kernel(unsigned stagingArea, unsigned numInputs) override;
	void _deinit() override;
#ifdef USE_LIBUV
	void _notifySent(void* req) override;
	void _notifyRecv(void* req) override;
	void _notifyProc(void* req) override;
#endif

	PAR2ProcHIP(const PAR2ProcHIP&);
	PAR2ProcHIP& operator=(const PAR2ProcHIP&);

public:
	// deviceId: HIP device ordinal; stagingAreas as the reference backends'
	// constructors take it (gf16/controller_ocl.h:202)
	explicit PAR2ProcHIP(IF_LIBUV(uv_loop_t* _loop, ) int deviceId = 0, int stagingAreas = 2);
	~PAR2ProcHIP();
	// After the controller's init (which sets the slice size): creates the
	// engine.  targetInputBatch 0 lets the engine choose the batch size from
	// the slice size.  False on failure (parhip_last_error() says why).
	bool init(unsigned targetInputBatch = 0);

	void setSliceSize(size_t size) override;
	bool setCurrentSliceSize(size_t size) override;
	bool setRecoverySlices(unsigned numSlices, const uint16_t* exponents = nullptr) override;
	PAR2ProcBackendAddResult canAdd() const override;
	FUTURE_RETURN_T addInput(const void* buffer, size_t size, uint16_t inputNum,
	                         bool flush IF_LIBUV(, const PAR2ProcPlainCb& cb)) override;
	FUTURE_RETURN_T addInput(const void* buffer, size_t size, const uint16_t* coeffs,
	                         bool flush IF_LIBUV(, const PAR2ProcPlainCb& cb)) override;
	void dummyInput(uint16_t inputNum, bool flush = false) override;
	bool fillInput(const void* buffer) override;
	void flush() override;
#ifndef USE_LIBUV
	FUTURE_RETURN_T endInput() override;
	void waitForAdd() override;
#endif
	FUTURE_RETURN_BOOL_T getOutput(unsigned index, void* output IF_LIBUV(, const PAR2ProcOutputCb& cb)) override;
	void processing_finished() override;
	void freeProcessingMem() override;
};

#endif  // __GF16_CONTROLLER_HIP
