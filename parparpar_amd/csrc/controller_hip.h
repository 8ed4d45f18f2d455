// controller_hip.h -- PAR2 GF(2^16) recovery backend on one MI355X.
//
// PAR2ProcHIP follows the backend contract of the reference's PAR2 controller
// (IPAR2ProcBackend, gf16/controller.h:77-217, non-libuv flavour with futures):
// the same member names, argument meaning and add/flush/end/get sequencing, so
// that PAR2Proc (gf16/controller.cpp) can drive it beside PAR2ProcCPU and
// PAR2ProcOCL.  Inside, inputs are staged in pinned host memory with a linear
// positional checksum word appended, copied to HBM a batch at a time, and
// multiplied into the device-resident recovery slices by the GF(2^16) kernels;
// the first batch of a pass overwrites the outputs and later batches XOR into
// them (controller_ocl.cpp:626-634).  getOutput() copies one recovery slice back and
// reports whether its checksum word still matches (the reference's cksumValid).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <future>
#include <string>
#include <vector>

namespace parhip {

// Values as in gf16/controller.h:41-46.
enum PAR2ProcBackendAddResult { PROC_ADD_OK, PROC_ADD_OK_BUSY, PROC_ADD_FULL, PROC_ADD_ALL_FULL };

class PAR2ProcHIP {
public:
	explicit PAR2ProcHIP(int device = 0, unsigned stagingAreas = 2);
	~PAR2ProcHIP();
	PAR2ProcHIP(const PAR2ProcHIP&) = delete;
	PAR2ProcHIP& operator=(const PAR2ProcHIP&) = delete;

	// inputBatchSize = inputs multiplied per kernel launch (0 = default)
	bool init(unsigned inputBatchSize = 0);
	void setSliceSize(size_t size);
	bool setCurrentSliceSize(size_t size);
	bool setRecoverySlices(unsigned numSlices, const uint16_t* exponents = nullptr);
	int getNumRecoverySlices() const { return (int)outputExponents.size(); }
	PAR2ProcBackendAddResult canAdd() const;
	std::future<void> addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush);
	std::future<void> addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush);
	void dummyInput(uint16_t inputNum, bool flush = false);
	bool fillInput(const void* buffer);
	void flush();
	std::future<void> endInput();
	std::future<bool> getOutput(unsigned index, void* output);
	void discardOutput() { processingAdd = false; }
	void waitForAdd();
	void freeProcessingMem();
	void deinit();
	unsigned getInputBatchSize() const { return inputBatchSize; }
	unsigned getBatchesStarted() const { return statBatchesStarted; }
	void setMinInputBatchSize(unsigned size) { minInBatchSize = size ? size : inputBatchSize; }
	bool ok() const { return err.empty(); }
	const std::string& lastError() const { return err; }
	hipStream_t stream() const { return strm; }

private:
	struct Staging {
		uint16_t* hsrc = nullptr;   // pinned: inputBatchSize x strideWords
		uint16_t* hcoef = nullptr;  // pinned: numOutputs x inputBatchSize
		uint16_t* dsrc = nullptr;
		uint16_t* dcoef = nullptr;
		hipEvent_t done = nullptr;
		bool active = false;
	};
	template <class T>
	std::future<void> addInputT(const void* buffer, size_t size, T numOrCoeffs, bool flush);
	void setCoeffs(Staging& st, unsigned idx, uint16_t inputNum);
	void setCoeffs(Staging& st, unsigned idx, const uint16_t* coeffs);
	void run_kernel(unsigned area, unsigned numInputs);
	bool allocStaging();
	void freeStaging();
	bool allocOutputs();
	void setErr(const char* what, hipError_t e);

	int dev;
	hipStream_t strm = nullptr;
	std::vector<Staging> staging;
	std::vector<uint16_t> outputExponents;
	uint16_t* dout = nullptr;     // numOutputs x strideWords
	uint16_t* hout = nullptr;     // pinned: one output slice on its way back (getOutput)
	size_t sliceSize = 0, currentSliceSize = 0;
	size_t allocWords = 0;        // per-slice words allocated (payload + checksum)
	size_t words = 0;             // current payload words + 1 checksum word
	size_t strideWords = 0;
	size_t allocOutputsCount = 0;
	unsigned inputBatchSize = 64, minInBatchSize = 64;
	unsigned currentStagingArea = 0, currentStagingInputs = 0;
	unsigned statBatchesStarted = 0;
	bool processingAdd = false;
	std::string err;
};

}  // namespace parhip
