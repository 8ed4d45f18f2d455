// par2_backend_driver.cc -- TEST INFRASTRUCTURE ONLY (never shipped): the
// reference's PAR2 controller (PAR2Proc, gf16/controller.cpp, compiled from
// the reference tree into _ref/) driving PAR2ProcHIP backends the way the
// reference's controller test drives its backends
// (test/gf16/test-ctrl.cpp:140-215): wait for room, add every input, end,
// read every recovery slice with its cksumValid, deinit.  Built twice, for the
// futures flavour and for USE_LIBUV (a uv loop, adds resumed from the
// controller's progress callback).
//
//   par2_backend_driver <case file> <result file>
//
// case file (little-endian): u64 sliceSize, u64 lastSliceSize, u32 numInputs,
//   u32 numOutputs, u32 inputBatch (0: the engine's choice), u32 stagingAreas,
//   u32 flags, u16 inputNums[numInputs], u16 exponents[numOutputs], then
//   numInputs x sliceSize bytes of input slices (the last one is added with
//   lastSliceSize bytes).
// flags: 1 = each slice split over two backends (the even-sized first part on
//   one, the rest on the other); 2 = a first pass over the first half of the
//   inputs that is discarded (discardOutput) before the full pass; 4 = every
//   odd input added with explicit coefficients (gfmat_coeff) instead of its
//   number.
// result file: numOutputs x sliceSize bytes of recovery, then numOutputs bytes
//   of cksumValid.  Exit status 0 on success.
#include "controller.h"
#include "controller_hip.h"
#include "gfmat_coeff.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <vector>

struct Case {
	uint64_t sliceSize, lastSliceSize;
	uint32_t numInputs, numOutputs, inputBatch, stagingAreas, flags;
	std::vector<uint16_t> inputNums, exps;
	std::vector<uint8_t> data;
	std::vector<std::vector<uint16_t>> coeffs;  // per input (flag 4)
};

static bool read_case(const char* path, Case& c) {
	FILE* f = fopen(path, "rb");
	if (!f) return false;
	bool ok = fread(&c.sliceSize, 8, 1, f) == 1 && fread(&c.lastSliceSize, 8, 1, f) == 1 &&
	          fread(&c.numInputs, 4, 1, f) == 1 && fread(&c.numOutputs, 4, 1, f) == 1 &&
	          fread(&c.inputBatch, 4, 1, f) == 1 && fread(&c.stagingAreas, 4, 1, f) == 1 &&
	          fread(&c.flags, 4, 1, f) == 1;
	if (ok) {
		c.inputNums.resize(c.numInputs);
		c.exps.resize(c.numOutputs);
		c.data.resize((size_t)c.numInputs * c.sliceSize);
		ok = fread(c.inputNums.data(), 2, c.numInputs, f) == c.numInputs &&
		     fread(c.exps.data(), 2, c.numOutputs, f) == c.numOutputs &&
		     fread(c.data.data(), 1, c.data.size(), f) == c.data.size();
	}
	fclose(f);
	return ok;
}

static const void* slice(const Case& c, unsigned i) { return c.data.data() + (size_t)i * c.sliceSize; }
static size_t slice_len(const Case& c, unsigned i) {
	return i == c.numInputs - 1 ? c.lastSliceSize : c.sliceSize;
}

struct Run {
	Case c;
	PAR2Proc par2;
	std::vector<PAR2ProcHIP*> backends;
	std::vector<uint8_t> out, valid;
	int failures = 0;
	~Run() {
		for (auto* b : backends) delete b;
	}
};

static bool setup(Run& r IF_LIBUV(, uv_loop_t* loop, const PAR2ProcCompleteCb& progress)) {
	const Case& c = r.c;
	std::vector<struct PAR2ProcBackendAlloc> alloc;
	const unsigned nb = (c.flags & 1) && c.sliceSize >= 4 ? 2 : 1;
	size_t first = c.sliceSize;
	if (nb == 2) {
		first = c.sliceSize / 2;
		first += first & 1;  // backends start on 16-bit word boundaries
	}
	for (unsigned b = 0; b < nb; b++) {
		r.backends.push_back(new PAR2ProcHIP(IF_LIBUV(loop, ) 0, (int)c.stagingAreas));
		alloc.push_back({r.backends.back(), b ? first : 0, b ? c.sliceSize - first : first});
	}
	if (!r.par2.init(c.sliceSize, alloc IF_LIBUV(, progress))) return false;
	for (auto* b : r.backends)
		if (!b->init(c.inputBatch)) {
			fprintf(stderr, "backend init: %s\n", parhip_last_error());
			return false;
		}
	if (!r.par2.setRecoverySlices(c.numOutputs, c.exps.data())) return false;
	r.out.assign((size_t)c.numOutputs * c.sliceSize, 0);
	r.valid.assign(c.numOutputs, 0);
	return true;
}

static bool write_result(const Run& r, const char* path) {
	FILE* f = fopen(path, "wb");
	if (!f) return false;
	bool ok = fwrite(r.out.data(), 1, r.out.size(), f) == r.out.size() &&
	          fwrite(r.valid.data(), 1, r.valid.size(), f) == r.valid.size();
	return fclose(f) == 0 && ok;
}

#ifndef USE_LIBUV
static void add_pass(Run& r, unsigned n) {
	for (unsigned i = 0; i < n; i++) {
		r.par2.waitForAdd();
		std::future<void> f;
		if ((r.c.flags & 4) && (i & 1))
			f = r.par2.addInput(slice(r.c, i), slice_len(r.c, i), r.c.coeffs[i].data(), false);
		else
			f = r.par2.addInput(slice(r.c, i), slice_len(r.c, i), r.c.inputNums[i], false);
		f.get();  // the buffer may be reused (the test keeps it anyway)
	}
	r.par2.endInput().get();
}

static int run(Run& r) {
	if (!setup(r)) return 2;
	if (r.c.flags & 2) {
		add_pass(r, r.c.numInputs / 2);
		r.par2.discardOutput();
	}
	add_pass(r, r.c.numInputs);
	for (unsigned o = 0; o < r.c.numOutputs; o++)
		r.valid[o] = r.par2.getOutput(o, r.out.data() + (size_t)o * r.c.sliceSize).get() ? 1 : 0;
	r.par2.deinit();
	return 0;
}
#else
// The loop thread drives everything: adds until the controller is full, then
// more from the progress callback; after endInput's callback, the outputs.
struct UvState {
	Run* r;
	uv_timer_t timer;  // starts the second pass outside endInput's callback
	unsigned pass = 0, next = 0, n = 0, outsDone = 0;
	bool ending = false, finished = false;
};

static void uv_add_more(UvState& s);

static void uv_outputs(UvState& s) {
	Run& r = *s.r;
	for (unsigned o = 0; o < r.c.numOutputs; o++) {
		r.par2.getOutput(o, r.out.data() + (size_t)o * r.c.sliceSize, [&s, o](bool valid) {
			s.r->valid[o] = valid ? 1 : 0;
			if (++s.outsDone == s.r->c.numOutputs)
				s.r->par2.deinit([&s]() {
					s.finished = true;
					uv_close(reinterpret_cast<uv_handle_t*>(&s.timer), nullptr);
				});
		});
	}
}

static void uv_end_pass(UvState& s) {
	s.ending = true;
	s.r->par2.endInput([&s]() {
		s.ending = false;
		if (s.pass == 0 && (s.r->c.flags & 2)) {
			// the controller clears its finish callback after calling it, so
			// the next pass (which ends with a new one) starts from the loop
			uv_timer_start(&s.timer, [](uv_timer_t* t) {
				auto& st = *static_cast<UvState*>(t->data);
				st.r->par2.discardOutput();
				st.pass = 1;
				st.next = 0;
				st.n = st.r->c.numInputs;
				uv_add_more(st);
			}, 0, 0);
		} else {
			uv_outputs(s);
		}
	});
}

static void uv_add_more(UvState& s) {
	Run& r = *s.r;
	while (s.next < s.n && !s.ending) {
		const unsigned i = s.next;
		bool added;
		if ((r.c.flags & 4) && (i & 1))
			added = r.par2.addInput(slice(r.c, i), slice_len(r.c, i), r.c.coeffs[i].data(), false, nullptr);
		else
			added = r.par2.addInput(slice(r.c, i), slice_len(r.c, i), r.c.inputNums[i], false, nullptr);
		if (!added) return;  // full: resumed from the progress callback
		if (++s.next == s.n) uv_end_pass(s);
	}
}

static int run(Run& r) {
	uv_loop_t loop;
	uv_loop_init(&loop);
	UvState s;
	s.r = &r;
	s.n = (r.c.flags & 2) ? r.c.numInputs / 2 : r.c.numInputs;
	uv_timer_init(&loop, &s.timer);
	s.timer.data = &s;
	if (!setup(r, &loop, [&s](unsigned) { uv_add_more(s); })) return 2;
	if (s.n)
		uv_add_more(s);
	else
		uv_end_pass(s);
	uv_run(&loop, UV_RUN_DEFAULT);
	uv_loop_close(&loop);
	return s.finished ? 0 : 3;
}
#endif

int main(int argc, char** argv) {
	if (argc != 3) {
		fprintf(stderr, "usage: %s <case file> <result file>\n", argv[0]);
		return 1;
	}
	auto r = std::unique_ptr<Run>(new Run);
	if (!read_case(argv[1], r->c) || r->c.numInputs == 0 || r->c.numOutputs == 0) {
		fprintf(stderr, "bad case file %s\n", argv[1]);
		return 1;
	}
	gfmat_init();
	if (r->c.flags & 4) {
		r->c.coeffs.resize(r->c.numInputs);
		for (unsigned i = 0; i < r->c.numInputs; i++)
			for (unsigned o = 0; o < r->c.numOutputs; o++)
				r->c.coeffs[i].push_back(gfmat_coeff(r->c.inputNums[i], r->c.exps[o]));
	}
	int rc = run(*r);
	if (rc) {
		fprintf(stderr, "run failed (%d)\n", rc);
		return rc;
	}
	if (!write_result(*r, argv[2])) {
		fprintf(stderr, "cannot write %s\n", argv[2]);
		return 1;
	}
	return 0;
}
