// repair.cpp -- PAR3 repair on the device and a continuous-batching repair
// server.
//
// One repair (value semantics of lib/par3gen.js:3840-4012 followed by
// SolveAndReconstruct, src/par3_engine.cc:980-1048): k input slices are lost;
// with recovery rows eq (ids firstRecovery + rows[eq]),
//   RHS[eq] = REC[rows[eq]] ^ XOR_{known j} c(eq, j) * D[j]
//   X       = A^-1 RHS,   A[eq][m] = c(eq, missing[m])
// and X[m] equals the lost slice missing[m].  On the device: the RHS is one
// contraction over [known slices, used recovery slices] with coefficients
// [Cauchy | identity], A^-1 comes from the closed form of a Cauchy inverse
// (repair_kernels.hip), and X is a second contraction.
//
// The closed form holds only for a true Cauchy matrix.  When a missing input's
// point equals a used recovery point (overlapping ranges, which the reference
// accepts: test/par3-fenger-padded-engine.js:110-112), the reference's
// coefficient for that pair is 1 (src/par3_engine.cc:392-394) and A is no
// longer Cauchy; then A is inverted on the host by Gauss-Jordan with the
// reference's pivot rule (src/par3_engine.cc:989-1045), and a singular A gives
// PARHIP_ERR_SINGULAR (-1), the reference's result.
//
// The server keeps archives resident in HBM, admits requests in arrival
// order while the number in flight and their device memory stay within the
// configured limits, gives every admitted request its own stream (their work
// overlaps on the device), and records host-observed latency from submit to
// completion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/parhip.h"
#include "../../include/parhip_repair.h"
#include "engine_internal.h"
#include "gf64_host.h"
#include "kernels.h"

namespace parhip {

struct RepairArchive {
	const uint64_t* data;
	uint64_t N, W, data_stride;
	const uint64_t* rec;
	uint64_t R, rec_stride;
	uint64_t fi, fr;
};

// Host-side state of one enqueued repair; must outlive the device work.
struct RepairPlan {
	std::vector<uint64_t> host_idx;  // [xs_known | xs_missing | ys]
	std::vector<const uint64_t*> in_ptrs;
	std::vector<uint64_t> host_inv;  // A^-1 from the host (colliding points only)
	void* ws = nullptr;
	size_t ws_bytes = 0;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t repair_ws_bytes(uint64_t N, uint64_t W, uint64_t k) {
	size_t b = 0;
	b += align256(N * 8);      // indices: xs_known (N-k) + xs_missing (k)
	b += align256(k * 8);      // ys
	b += align256(N * 8);      // input pointers
	b += align256(k * N * 8);  // RHS coefficients [Cauchy | I]
	b += align256(k * 8) * 2;  // alpha, beta
	b += align256(k * k * 8);  // A^-1
	b += align256(k * W * 8);  // RHS blocks
	return b;
}

// Device bytes a repair needs beyond its output: workspace + the split-K
// partials of its two contractions.
static size_t repair_bytes(uint64_t N, uint64_t W, uint64_t k) {
	const int dev = current_device();
	return repair_ws_bytes(N, W, k) + plan_workspace_bytes(plan_contraction(N, k, W, dev), k, W) +
	       plan_workspace_bytes(plan_contraction(k, k, W, dev), k, W);
}

static int repair_enqueue(const RepairArchive& ar, const uint64_t* missing, uint64_t k,
                          const uint64_t* rows, uint64_t* d_out, uint64_t out_stride,
                          hipStream_t s, RepairPlan& plan) {
	const uint64_t N = ar.N, W = ar.W;
	if (k == 0) return PARHIP_OK;
	if (k > N || k > ar.R) return fail(PARHIP_ERR_ARG, "repair: k=%llu exceeds N or R", (unsigned long long)k);
	std::vector<char> lost(N, 0);
	for (uint64_t m = 0; m < k; m++) {
		if (missing[m] >= N || lost[missing[m]])
			return fail(PARHIP_ERR_ARG, "repair: bad or repeated missing index");
		lost[missing[m]] = 1;
	}
	std::vector<char> used(ar.R, 0);
	for (uint64_t e = 0; e < k; e++) {
		const uint64_t r = rows ? rows[e] : e;
		if (r >= ar.R || used[r]) return fail(PARHIP_ERR_ARG, "repair: bad or repeated recovery row");
		used[r] = 1;
	}
	const uint64_t nk = N - k;
	plan.host_idx.resize(N + k);
	plan.in_ptrs.resize(N);
	uint64_t* xs_known = plan.host_idx.data();
	uint64_t* xs_miss = xs_known + nk;
	uint64_t* ys = xs_known + N;
	uint64_t c = 0;
	for (uint64_t j = 0; j < N; j++)
		if (!lost[j]) {
			xs_known[c] = ar.fi + j;
			plan.in_ptrs[c] = ar.data + j * ar.data_stride;
			c++;
		}
	for (uint64_t m = 0; m < k; m++) xs_miss[m] = ar.fi + missing[m];
	for (uint64_t e = 0; e < k; e++) {
		const uint64_t r = rows ? rows[e] : e;
		ys[e] = ar.fr + r;
		plan.in_ptrs[nk + e] = ar.rec + r * ar.rec_stride;
	}
	// A point shared by a missing input and a used recovery row makes A a
	// Cauchy matrix with some entries replaced by 1: invert it on the host.
	bool collide = false;
	{
		std::vector<uint64_t> ys_sorted(ys, ys + k);
		std::sort(ys_sorted.begin(), ys_sorted.end());
		for (uint64_t m = 0; m < k && !collide; m++)
			collide = std::binary_search(ys_sorted.begin(), ys_sorted.end(), xs_miss[m]);
	}
	if (collide) {
		std::vector<uint64_t> A(k * k);
		for (uint64_t eq = 0; eq < k; eq++)
			for (uint64_t m = 0; m < k; m++) {
				const uint64_t d = ys[eq] ^ xs_miss[m];
				A[eq * k + m] = d ? host_gf64_inv(d) : 1u;
			}
		plan.host_inv.resize(k * k);
		if (host_gauss_jordan(A.data(), plan.host_inv.data(), k))
			return fail(PARHIP_ERR_SINGULAR, "repair: the %llu x %llu system is singular",
			            (unsigned long long)k, (unsigned long long)k);
	}
	// workspace carve-up
	plan.ws_bytes = repair_ws_bytes(N, W, k);
	int rc = dev_alloc(&plan.ws, plan.ws_bytes, s);
	if (rc) return rc;
	char* p = (char*)plan.ws;
	uint64_t* d_idx = (uint64_t*)p;          p += align256(N * 8);
	uint64_t* d_ys = (uint64_t*)p;           p += align256(k * 8);
	const uint64_t** d_ptrs = (const uint64_t**)p; p += align256(N * 8);
	uint64_t* d_coef = (uint64_t*)p;         p += align256(k * N * 8);
	uint64_t* d_alpha = (uint64_t*)p;        p += align256(k * 8);
	uint64_t* d_beta = (uint64_t*)p;         p += align256(k * 8);
	uint64_t* d_inv = (uint64_t*)p;          p += align256(k * k * 8);
	uint64_t* d_rhs = (uint64_t*)p;
	hipError_t e = hipMemcpyAsync(d_idx, plan.host_idx.data(), N * 8, hipMemcpyHostToDevice, s);
	if (e == hipSuccess) e = hipMemcpyAsync(d_ys, ys, k * 8, hipMemcpyHostToDevice, s);
	if (e == hipSuccess)
		e = hipMemcpyAsync(d_ptrs, plan.in_ptrs.data(), N * 8, hipMemcpyHostToDevice, s);
	if (e == hipSuccess) e = launch_repair_coeffs(d_coef, N, d_idx, nk, d_ys, k, s);
	if (e == hipSuccess) {
		if (collide)
			e = hipMemcpyAsync(d_inv, plan.host_inv.data(), k * k * 8, hipMemcpyHostToDevice, s);
		else
			e = launch_cauchy_inverse(d_ys, d_idx + nk, k, d_alpha, d_beta, d_inv, s);
	}
	if (e != hipSuccess) return hip_fail(e, "repair setup");
	// RHS = [C_known | I] * [known slices ; recovery rows]
	MacArgs a{};
	a.in_ptrs = d_ptrs;
	a.out_base = d_rhs;
	a.out_stride = W;
	a.coeff = d_coef;
	a.coeff_ld = N;
	a.N = N;
	a.R = k;
	a.W = W;
	a.accumulate = 0;
	rc = run_contraction(a, s);
	if (rc) return rc;
	// X = A^-1 * RHS
	MacArgs b{};
	b.in_base = d_rhs;
	b.in_stride = W;
	b.out_base = d_out;
	b.out_stride = out_stride;
	b.coeff = d_inv;
	b.coeff_ld = k;
	b.N = k;
	b.R = k;
	b.W = W;
	b.accumulate = 0;
	return run_contraction(b, s);
}

}  // namespace parhip

using namespace parhip;

// ======================================================================== server
struct RepairReq {
	int64_t id;
	int archive;
	std::vector<uint64_t> missing;
	uint64_t* out = nullptr;
	size_t out_bytes = 0, charge = 0;
	int stream = -1;
	hipEvent_t done = nullptr;
	RepairPlan plan;
	std::chrono::steady_clock::time_point t_submit, t_done;
	int state = 0;  // 0 pending, 1 in flight, 2 done, 3 failed
	int rc = 0;
};

struct parhip_repair_server {
	int dev = 0;
	size_t budget = 0, used = 0, peak = 0;
	unsigned max_inflight = 4;
	std::vector<hipStream_t> streams;
	std::vector<bool> stream_busy;
	std::vector<RepairArchive> archives;
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<RepairReq*> pending;
	std::vector<RepairReq*> inflight;
	std::map<int64_t, RepairReq*> reqs;
	std::vector<double> latencies_ms;
	unsigned peak_inflight = 0;
	int64_t next_id = 0;
	bool stop = false;
	std::thread sched;
};

static void finish_req(parhip_repair_server* S, RepairReq* r, int state) {
	r->t_done = std::chrono::steady_clock::now();
	r->state = state;
	// the workspace goes back to the pool on the request's stream; only the
	// output stays charged until release
	if (r->plan.ws) dev_free(r->plan.ws, S->streams[r->stream]);
	r->plan.ws = nullptr;
	S->used -= (r->charge - r->out_bytes);
	S->stream_busy[r->stream] = false;
	if (state == 2)
		S->latencies_ms.push_back(
		    std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
}

static void scheduler(parhip_repair_server* S) {
	parhip_set_device(S->dev);  // this thread's engine calls target the server's device
	std::unique_lock<std::mutex> lk(S->mu);
	while (true) {
		// completions
		for (size_t i = 0; i < S->inflight.size();) {
			RepairReq* r = S->inflight[i];
			if (hipEventQuery(r->done) == hipSuccess) {
				finish_req(S, r, r->rc ? 3 : 2);
				S->inflight.erase(S->inflight.begin() + i);
				S->cv_done.notify_all();
			} else {
				i++;
			}
		}
		// admissions, in arrival order
		while (!S->pending.empty() && S->inflight.size() < S->max_inflight) {
			RepairReq* r = S->pending.front();
			if (S->used + r->charge > S->budget && !S->inflight.empty()) break;  // wait for memory
			int si = -1;
			for (size_t k = 0; k < S->streams.size(); k++)
				if (!S->stream_busy[k]) {
					si = (int)k;
					break;
				}
			if (si < 0) break;
			S->pending.pop_front();
			r->stream = si;
			S->stream_busy[si] = true;
			S->used += r->charge;
			S->peak = std::max(S->peak, S->used);
			hipStream_t st = S->streams[si];
			const RepairArchive& ar = S->archives[r->archive];
			int rc = dev_alloc((void**)&r->out, r->out_bytes, st);
			if (!rc)
				rc = repair_enqueue(ar, r->missing.data(), r->missing.size(), nullptr, r->out, ar.W,
				                    st, r->plan);
			r->rc = rc;
			hipEventRecord(r->done, st);
			r->state = 1;
			S->inflight.push_back(r);
			S->peak_inflight = std::max<unsigned>(S->peak_inflight, (unsigned)S->inflight.size());
		}
		if (S->stop && S->pending.empty() && S->inflight.empty()) break;
		if (S->inflight.empty() && S->pending.empty())
			S->cv_work.wait(lk);
		else
			S->cv_work.wait_for(lk, std::chrono::microseconds(20));
	}
}

extern "C" {

int parhip_d_gf64_repair(const uint64_t* d_data, size_t numInputs, size_t dataStride64,
                         const uint64_t* d_recovery, size_t numRecovery, size_t recoveryStride64,
                         size_t blockSize64, uint64_t firstInput, uint64_t firstRecovery,
                         const uint64_t* missing, size_t k, const uint64_t* recoveryRows,
                         uint64_t* d_out, size_t outStride64, void* stream) {
	if (k == 0 || blockSize64 == 0) return PARHIP_OK;
	if (!d_data || !d_recovery || !missing || !d_out)
		return fail(PARHIP_ERR_ARG, "repair: null pointer");
	if (dataStride64 < blockSize64 || recoveryStride64 < blockSize64 || outStride64 < blockSize64)
		return fail(PARHIP_ERR_ARG, "repair: stride < blockSize64");
	device_cus(current_device());
	RepairArchive ar{d_data, numInputs, blockSize64, dataStride64, d_recovery, numRecovery,
	                 recoveryStride64, firstInput, firstRecovery};
	hipStream_t s = (hipStream_t)stream;
	RepairPlan plan;
	int rc = repair_enqueue(ar, missing, k, recoveryRows, d_out, outStride64, s, plan);
	// the host index arrays feed async copies: finish before they go away
	hipError_t e = hipStreamSynchronize(s);
	if (plan.ws) dev_free(plan.ws, s);
	if (rc) return rc;
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "repair");
}

size_t parhip_gf64_repair_workspace_bytes(size_t numInputs, size_t blockSize64, size_t k) {
	return repair_bytes(numInputs, blockSize64, k);
}

parhip_repair_server* parhip_repair_server_create(size_t hbmBudgetBytes, unsigned maxInflight,
                                                  unsigned numStreams) {
	auto* S = new (std::nothrow) parhip_repair_server;
	if (!S) return nullptr;
	S->dev = current_device();
	S->budget = hbmBudgetBytes;
	S->max_inflight = maxInflight ? maxInflight : 4;
	if (!numStreams) numStreams = S->max_inflight;
	hipSetDevice(S->dev);
	device_cus(S->dev);
	for (unsigned i = 0; i < numStreams; i++) {
		hipStream_t st;
		if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
			fail(PARHIP_ERR_DEVICE, "repair_server: stream creation failed");
			for (auto x : S->streams) hipStreamDestroy(x);
			delete S;
			return nullptr;
		}
		S->streams.push_back(st);
		S->stream_busy.push_back(false);
	}
	S->max_inflight = std::min<unsigned>(S->max_inflight, numStreams);
	S->sched = std::thread(scheduler, S);
	return S;
}

int parhip_repair_server_add_archive(parhip_repair_server* S, const uint64_t* d_data,
                                     size_t numInputs, size_t dataStride64,
                                     const uint64_t* d_recovery, size_t numRecovery,
                                     size_t recoveryStride64, size_t blockSize64,
                                     uint64_t firstInput, uint64_t firstRecovery) {
	if (!S || !d_data || !d_recovery || !numInputs || !numRecovery || !blockSize64)
		return fail(PARHIP_ERR_ARG, "repair_server_add_archive: bad argument");
	std::lock_guard<std::mutex> lk(S->mu);
	S->archives.push_back(RepairArchive{d_data, numInputs, blockSize64, dataStride64, d_recovery,
	                                    numRecovery, recoveryStride64, firstInput, firstRecovery});
	return (int)S->archives.size() - 1;
}

int64_t parhip_repair_submit(parhip_repair_server* S, int archive, const uint64_t* missing,
                             size_t k) {
	if (!S || !missing || k == 0) return fail(PARHIP_ERR_ARG, "repair_submit: bad argument");
	std::lock_guard<std::mutex> lk(S->mu);
	if (archive < 0 || (size_t)archive >= S->archives.size())
		return fail(PARHIP_ERR_ARG, "repair_submit: unknown archive %d", archive);
	const RepairArchive& ar = S->archives[archive];
	if (k > ar.R || k > ar.N) return fail(PARHIP_ERR_ARG, "repair_submit: k exceeds R");
	auto* r = new RepairReq;
	r->id = S->next_id++;
	r->archive = archive;
	r->missing.assign(missing, missing + k);
	r->out_bytes = k * ar.W * sizeof(uint64_t);
	r->charge = r->out_bytes + repair_bytes(ar.N, ar.W, k);
	if (r->charge > S->budget) {
		delete r;
		return fail(PARHIP_ERR_RANGE, "repair_submit: request needs more than the HBM budget");
	}
	hipEventCreateWithFlags(&r->done, hipEventDisableTiming);
	r->t_submit = std::chrono::steady_clock::now();
	S->reqs[r->id] = r;
	S->pending.push_back(r);
	S->cv_work.notify_all();
	return r->id;
}

int parhip_repair_wait(parhip_repair_server* S, int64_t id, double* latency_ms) {
	if (!S) return fail(PARHIP_ERR_ARG, "repair_wait: null server");
	std::unique_lock<std::mutex> lk(S->mu);
	auto it = S->reqs.find(id);
	if (it == S->reqs.end()) return fail(PARHIP_ERR_ARG, "repair_wait: unknown request");
	RepairReq* r = it->second;
	S->cv_done.wait(lk, [&] { return r->state >= 2; });
	if (latency_ms)
		*latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
	if (r->state == 3) return r->rc ? r->rc : PARHIP_ERR_DEVICE;
	return PARHIP_OK;
}

int parhip_repair_result(parhip_repair_server* S, int64_t id, uint64_t** d_out, size_t* k) {
	if (!S || !d_out) return fail(PARHIP_ERR_ARG, "repair_result: bad argument");
	std::lock_guard<std::mutex> lk(S->mu);
	auto it = S->reqs.find(id);
	if (it == S->reqs.end() || it->second->state != 2)
		return fail(PARHIP_ERR_ARG, "repair_result: request not complete");
	*d_out = it->second->out;
	if (k) *k = it->second->missing.size();
	return PARHIP_OK;
}

int parhip_repair_release(parhip_repair_server* S, int64_t id) {
	if (!S) return fail(PARHIP_ERR_ARG, "repair_release: null server");
	std::lock_guard<std::mutex> lk(S->mu);
	auto it = S->reqs.find(id);
	if (it == S->reqs.end() || it->second->state < 2)
		return fail(PARHIP_ERR_ARG, "repair_release: request not complete");
	RepairReq* r = it->second;
	if (r->out) dev_free(r->out, S->streams[r->stream]);
	S->used -= r->out_bytes;
	hipEventDestroy(r->done);
	delete r;
	S->reqs.erase(it);
	S->cv_work.notify_all();
	return PARHIP_OK;
}

int parhip_repair_server_stats(parhip_repair_server* S, uint64_t* completed, double* p50_ms,
                               double* p99_ms, unsigned* peak_inflight, uint64_t* peak_bytes) {
	if (!S) return fail(PARHIP_ERR_ARG, "repair_stats: null server");
	std::lock_guard<std::mutex> lk(S->mu);
	std::vector<double> v = S->latencies_ms;
	std::sort(v.begin(), v.end());
	auto pct = [&](double q) {
		if (v.empty()) return 0.0;
		size_t i = (size_t)(q * (v.size() - 1) + 0.5);
		return v[std::min(i, v.size() - 1)];
	};
	if (completed) *completed = v.size();
	if (p50_ms) *p50_ms = pct(0.50);
	if (p99_ms) *p99_ms = pct(0.99);
	if (peak_inflight) *peak_inflight = S->peak_inflight;
	if (peak_bytes) *peak_bytes = S->peak;
	return PARHIP_OK;
}

void parhip_repair_server_destroy(parhip_repair_server* S) {
	if (!S) return;
	{
		std::lock_guard<std::mutex> lk(S->mu);
		S->stop = true;
		S->cv_work.notify_all();
	}
	S->sched.join();
	hipSetDevice(S->dev);
	for (auto& kv : S->reqs) {
		RepairReq* r = kv.second;
		if (r->out) dev_free(r->out, S->streams[r->stream >= 0 ? r->stream : 0]);
		if (r->done) hipEventDestroy(r->done);
		delete r;
	}
	for (auto st : S->streams) {
		hipStreamSynchronize(st);
		hipStreamDestroy(st);
	}
	delete S;
}

}  // extern "C"
