// repair.cpp -- PAR3 repair on the device and a repair server.
//
// One repair (value semantics of lib/par3gen.js:3840-4012 followed by
// SolveAndReconstruct, src/par3_engine.cc:980-1048): k input slices are lost;
// with recovery rows eq (ids firstRecovery + rows[eq]),
//   RHS[eq] = REC[rows[eq]] ^ XOR_{known j} c(eq, j) * D[j]
//   X       = A^-1 RHS,   A[eq][m] = c(eq, missing[m])
// and X[m] equals the lost slice missing[m].  On the device the RHS is the
// create's XOR-block contraction (gf64_dyadic_dev.h) over the whole archive
// with the lost inputs masked to zero, XORed into a copy of the recovery rows,
// when the rows used are consecutive (their points one range); otherwise one
// contraction over [known slices, used recovery slices] with an explicit
// coefficient matrix [Cauchy | identity].  A^-1 comes from the closed form of
// a Cauchy inverse (repair_kernels.hip), and X is a second contraction.
//
// The closed form holds only for a true Cauchy matrix.  When a missing input's
// point equals a used recovery point (overlapping ranges, which the reference
// accepts: test/par3-fenger-padded-engine.js:110-112), the reference's
// coefficient for that pair is 1 (src/par3_engine.cc:392-394) and A is no
// longer Cauchy; then A is inverted on the host by Gauss-Jordan with the
// reference's pivot rule (src/par3_engine.cc:989-1045), and a singular A gives
// PARHIP_ERR_SINGULAR (-1), the reference's result.
//
// The server keeps archives resident in HBM, admits requests in arrival order
// while the number in flight and the device memory in use (workspace, outputs
// not yet released, cached inverses) stay within the configured limits, gives
// every admitted request its own stream (their work overlaps on the device),
// and records host-observed latency from submit to completion.  Inverted
// matrices stay resident in an LRU keyed by (archive, missing set), charged to
// the same budget and evicted, least recently used first, when a request needs
// the memory -- the analogue of the reference's coefficient cache
// (src/par3_engine.cc:451-482) for the repair matrices.
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
	std::vector<uint64_t> host_idx;  // table route: [xs_known | xs_missing | ys]; else [xs_missing | ys]
	std::vector<const uint64_t*> in_ptrs;
	std::vector<uint32_t> host_mask;  // XOR-block route: lost inputs per block of 8 or 16 points
	std::vector<uint64_t> host_inv;   // A^-1 from the host (colliding points only)
	void* ws = nullptr;
	size_t ws_bytes = 0;
};

// Where A^-1 lives when the caller keeps it (the server's inverse cache): d
// holds k x k words; ready = it already holds A^-1 for this repair.
struct RepairInv {
	uint64_t* d = nullptr;
	bool ready = false;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The RHS takes the XOR-block kernel when the recovery rows used are
// consecutive and the method allows the kernel (4 or more rows: below that the
// direct tables cost less, as for the create).  auto_method: as under the
// automatic choice (the server's scheduler), else this thread's method.
static bool repair_dyadic(const uint64_t* rows, uint64_t k, bool auto_method = false) {
	if (!(auto_method ? dyadic_shape_ok(k) : dyadic_applies_shape(k))) return false;
	if (rows)
		for (uint64_t e = 1; e < k; e++)
			if (rows[e] != rows[0] + e) return false;
	return true;
}

static size_t repair_ws_bytes(uint64_t N, uint64_t W, uint64_t k, const DyadicPlan* dp) {
	size_t b = 0;
	if (dp) {
		b += align256(dp->nblocks * 4);             // lost-input masks
		b += align256(dp->nblocks * dp->Q * 8 * (1ull << dp->bits));  // XOR-block coefficients
		b += align256(dp->part_bytes);              // split-K partials of the RHS
		b += align256(k * 8) * 2;                   // xs_missing, ys
	} else {
		b += align256(N * 8);      // indices: xs_known (N-k) + xs_missing (k)
		b += align256(k * 8);      // ys
		b += align256(N * 8);      // input pointers
		b += align256(k * N * 8);  // RHS coefficients [Cauchy | I]
	}
	b += align256(k * 8) * 2;  // alpha, beta
	b += align256(k * k * 8);  // A^-1
	b += align256(k * W * 8);  // RHS blocks
	return b;
}

// Device bytes a repair of the first k recovery rows needs beyond its output:
// workspace + the split-K partials of its contractions.
// force_bits: an XOR-block size whatever the rule picks (0: the rule's).
static size_t repair_bytes(uint64_t N, uint64_t W, uint64_t k, uint64_t fi, uint64_t fr,
                           bool auto_method, int force_bits = 0) {
	const int dev = current_device();
	const size_t solve = plan_workspace_bytes(plan_contraction(k, k, W, dev), k, W);
	if (repair_dyadic(nullptr, k, auto_method)) {
		const DyadicPlan dp = plan_dyadic(N, k, W, fi, fr, dev, SIZE_MAX, true, force_bits, true);
		return repair_ws_bytes(N, W, k, &dp) + solve;
	}
	return repair_ws_bytes(N, W, k, nullptr) +
	       plan_workspace_bytes(plan_contraction(N, k, W, dev), k, W) + solve;
}

static int repair_enqueue(const RepairArchive& ar, const uint64_t* missing, uint64_t k,
                          const uint64_t* rows, uint64_t* d_out, uint64_t out_stride,
                          hipStream_t s, RepairPlan& plan, RepairInv inv = RepairInv()) {
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
	const bool dy = repair_dyadic(rows, k);
	const uint64_t nk = N - k;
	// points: the known inputs (table route only), the missing inputs, the rows used
	plan.host_idx.resize(dy ? 2 * k : N + k);
	uint64_t* xs_known = plan.host_idx.data();
	uint64_t* xs_miss = dy ? xs_known : xs_known + nk;
	uint64_t* ys = xs_miss + k;
	if (!dy) {
		plan.in_ptrs.resize(N);
		uint64_t c = 0;
		for (uint64_t j = 0; j < N; j++)
			if (!lost[j]) {
				xs_known[c] = ar.fi + j;
				plan.in_ptrs[c] = ar.data + j * ar.data_stride;
				c++;
			}
	}
	for (uint64_t m = 0; m < k; m++) xs_miss[m] = ar.fi + missing[m];
	for (uint64_t e = 0; e < k; e++) {
		const uint64_t r = rows ? rows[e] : e;
		ys[e] = ar.fr + r;
		if (!dy) plan.in_ptrs[nk + e] = ar.rec + r * ar.rec_stride;
	}
	// A point shared by a missing input and a used recovery row makes A a
	// Cauchy matrix with some entries replaced by 1: invert it on the host.
	bool collide = false;
	if (!inv.ready) {
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
	const uint64_t yfirst = ar.fr + (rows ? rows[0] : 0);
	DyadicPlan dp{};
	if (dy) dp = plan_dyadic(N, k, W, ar.fi, yfirst, current_device(), SIZE_MAX, true, 0, true);
	const uint64_t bm = (1ull << dp.bits) - 1;  // points per XOR block - 1 (blocks of 8, 16 or 32)
	// workspace carve-up
	plan.ws_bytes = repair_ws_bytes(N, W, k, dy ? &dp : nullptr);
	int rc = dev_alloc(&plan.ws, plan.ws_bytes, s);
	if (rc) return rc;
	char* p = (char*)plan.ws;
	uint32_t* d_mask = nullptr;
	uint64_t *d_gp = nullptr, *d_part = nullptr, *d_idx = nullptr, *d_xm, *d_ys;
	const uint64_t** d_ptrs = nullptr;
	uint64_t* d_coef = nullptr;
	hipError_t e;
	if (dy) {
		d_mask = (uint32_t*)p;      p += align256(dp.nblocks * 4);
		d_gp = (uint64_t*)p;        p += align256(dp.nblocks * dp.Q * 8 * (bm + 1));
		d_part = (uint64_t*)p;      p += align256(dp.part_bytes);
		d_xm = (uint64_t*)p;        p += align256(k * 8) * 2;
		d_ys = d_xm + k;  // uploaded together with xs_missing
		plan.host_mask.assign(dp.nblocks, 0u);
		for (uint64_t m = 0; m < k; m++) {
			const uint64_t v = dp.pre_in + missing[m];  // virtual slot of the lost input
			plan.host_mask[v / (bm + 1)] |= 1u << (v % (bm + 1));
		}
		e = hipMemcpyAsync(d_mask, plan.host_mask.data(), dp.nblocks * 4, hipMemcpyHostToDevice, s);
		if (e == hipSuccess) e = hipMemcpyAsync(d_xm, xs_miss, 2 * k * 8, hipMemcpyHostToDevice, s);
	} else {
		d_idx = (uint64_t*)p;                p += align256(N * 8);
		d_ys = (uint64_t*)p;                 p += align256(k * 8);
		d_ptrs = (const uint64_t**)p;        p += align256(N * 8);
		d_coef = (uint64_t*)p;               p += align256(k * N * 8);
		d_xm = d_idx + nk;
		e = hipMemcpyAsync(d_idx, plan.host_idx.data(), N * 8, hipMemcpyHostToDevice, s);
		if (e == hipSuccess) e = hipMemcpyAsync(d_ys, ys, k * 8, hipMemcpyHostToDevice, s);
		if (e == hipSuccess)
			e = hipMemcpyAsync(d_ptrs, plan.in_ptrs.data(), N * 8, hipMemcpyHostToDevice, s);
		if (e == hipSuccess) e = launch_repair_coeffs(d_coef, N, d_idx, nk, d_ys, k, s);
	}
	uint64_t* d_alpha = (uint64_t*)p;        p += align256(k * 8);
	uint64_t* d_beta = (uint64_t*)p;         p += align256(k * 8);
	uint64_t* d_inv = (uint64_t*)p;          p += align256(k * k * 8);
	uint64_t* d_rhs = (uint64_t*)p;
	if (inv.d) d_inv = inv.d;
	if (e == hipSuccess && !inv.ready) {
		if (collide)
			e = hipMemcpyAsync(d_inv, plan.host_inv.data(), k * k * 8, hipMemcpyHostToDevice, s);
		else
			e = launch_cauchy_inverse(d_ys, d_xm, k, d_alpha, d_beta, d_inv, s);
	}
	if (e != hipSuccess) return hip_fail(e, "repair setup");
	if (dy) {
		// RHS = the rows used, XORed with the contraction of every input but
		// the lost ones (masked to zero) at the rows' points
		e = dp.bits == 5   ? launch_dyadic32_coeffs(d_gp, dp.nblocks, dp.Q, ar.fi & ~bm, yfirst & ~bm, s)
		    : dp.bits == 4 ? launch_dyadic16_coeffs(d_gp, dp.nblocks, dp.Q, ar.fi & ~bm, yfirst & ~bm, s)
		                   : launch_dyadic_coeffs(d_gp, dp.nblocks, dp.Q, ar.fi & ~bm, yfirst & ~bm, s);
		if (e == hipSuccess)
			e = hipMemcpy2DAsync(d_rhs, W * 8, ar.rec + (yfirst - ar.fr) * ar.rec_stride,
			                     ar.rec_stride * 8, W * 8, k, hipMemcpyDeviceToDevice, s);
		MacArgs a{};
		a.in_base = ar.data;
		a.in_stride = ar.data_stride;
		a.out_base = d_rhs;
		a.out_stride = W;
		a.N = N;
		a.R = k;
		a.W = W;
		a.accumulate = 1;
		if (e == hipSuccess) e = enqueue_dyadic(a, dp, d_gp, dp.nblocks, d_part, s, d_mask);
		if (e != hipSuccess) return hip_fail(e, "repair RHS");
	} else {
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
	}
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
// One cached A^-1: k x k words on the device, written by the request that
// missed; `ready` is recorded on that request's stream after its work, and a
// request that hits waits for it on its own stream.
struct InvEntry {
	uint64_t* d = nullptr;
	size_t bytes = 0;
	hipEvent_t ready = nullptr;
	unsigned users = 0;   // admitted requests using it (not evictable while > 0)
	uint64_t last_use = 0;
	bool filled = false;  // its producer's work is enqueued: later requests may use it
	bool valid = true;    // false: its producer failed; dropped once unused
};
typedef std::pair<int, std::vector<uint64_t>> InvKey;  // (archive, missing in request order)

struct RepairReq {
	int64_t id;
	int archive;
	std::vector<uint64_t> missing;
	uint64_t* out = nullptr;
	size_t out_bytes = 0, charge = 0;
	int stream = -1;
	hipEvent_t done = nullptr;
	RepairPlan plan;
	InvEntry* inv = nullptr;
	bool inv_hit = false;
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
	std::map<InvKey, InvEntry> inv_cache;
	size_t inv_bytes = 0;
	uint64_t inv_clock = 0, inv_hits = 0, inv_misses = 0, inv_evictions = 0;
	unsigned peak_inflight = 0;
	int64_t next_id = 0;
	bool stop = false;
	std::thread sched;
};

static const size_t kMaxInvEntries = 256;

// Free an unused cache entry (mu held).  No request uses it and every request
// that did has completed, so the memory is idle.
static void inv_drop(parhip_repair_server* S, std::map<InvKey, InvEntry>::iterator it) {
	dev_free(it->second.d, S->streams[0]);
	if (it->second.ready) hipEventDestroy(it->second.ready);
	S->used -= it->second.bytes;
	S->inv_bytes -= it->second.bytes;
	S->inv_cache.erase(it);
}

// Evict least-recently-used unused entries until `need` more bytes fit the
// budget (or none is left to evict); mu held.  Returns whether they fit.
static bool inv_make_room(parhip_repair_server* S, size_t need) {
	while (S->used + need > S->budget) {
		auto victim = S->inv_cache.end();
		for (auto it = S->inv_cache.begin(); it != S->inv_cache.end(); ++it)
			if (!it->second.users && (victim == S->inv_cache.end() || it->second.last_use < victim->second.last_use))
				victim = it;
		if (victim == S->inv_cache.end()) return false;
		inv_drop(S, victim);
		S->inv_evictions++;
	}
	return true;
}

static void finish_req(parhip_repair_server* S, RepairReq* r, int state) {
	r->t_done = std::chrono::steady_clock::now();
	r->state = state;
	// the workspace goes back to the pool on the request's stream; only the
	// output stays charged until release
	if (r->plan.ws) dev_free(r->plan.ws, S->streams[r->stream]);
	r->plan.ws = nullptr;
	S->used -= (r->charge - r->out_bytes);
	S->stream_busy[r->stream] = false;
	if (r->inv) {
		r->inv->users--;
		if (state != 2 && !r->inv_hit) r->inv->valid = false;  // never filled
		if (!r->inv->valid && !r->inv->users)
			for (auto it = S->inv_cache.begin(); it != S->inv_cache.end(); ++it)
				if (&it->second == r->inv) {
					inv_drop(S, it);
					break;
				}
		r->inv = nullptr;
	}
	if (state == 2)
		S->latencies_ms.push_back(
		    std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
}

// Admit the next pending request if the limits allow (mu held): its stream,
// its charge, its inverse (cached, or a new entry it fills).  Returns null
// when it has to wait.
static RepairReq* admit(parhip_repair_server* S) {
	if (S->pending.empty() || S->inflight.size() >= S->max_inflight) return nullptr;
	RepairReq* r = S->pending.front();
	int si = -1;
	for (size_t k = 0; k < S->streams.size(); k++)
		if (!S->stream_busy[k]) {
			si = (int)k;
			break;
		}
	if (si < 0) return nullptr;
	InvKey key(r->archive, r->missing);
	auto hit = S->inv_cache.find(key);
	// an entry whose producer is still being enqueued is not used (nor replaced)
	const bool reuse = hit != S->inv_cache.end() && hit->second.valid && hit->second.filled;
	const size_t k = r->missing.size();
	const size_t inv_need = reuse ? 0 : k * k * sizeof(uint64_t);
	if (hit != S->inv_cache.end() && !reuse && !hit->second.users) inv_drop(S, hit);
	// memory: the request's charge always, a new cache entry if it fits
	// (else the request inverts into its own workspace)
	if (!inv_make_room(S, r->charge)) return nullptr;  // wait for releases
	const bool cache_new = !reuse && hit == S->inv_cache.end() && S->inv_cache.size() < kMaxInvEntries &&
	                       inv_make_room(S, r->charge + inv_need);
	S->pending.pop_front();
	r->stream = si;
	S->stream_busy[si] = true;
	S->used += r->charge;
	if (reuse) {
		r->inv = &hit->second;
		r->inv_hit = true;
		S->inv_hits++;
	} else {
		S->inv_misses++;
		if (cache_new) {
			InvEntry& en = S->inv_cache[key];
			en.bytes = inv_need;
			S->used += inv_need;
			S->inv_bytes += inv_need;
			r->inv = &en;
		}
	}
	if (r->inv) {
		r->inv->users++;
		r->inv->last_use = ++S->inv_clock;
	}
	S->peak = std::max(S->peak, S->used);
	r->state = 1;
	return r;
}

// Enqueue an admitted request on its stream (no lock held: host index
// building, uploads and launches do not hold up submit, wait or release).
static void enqueue_req(parhip_repair_server* S, RepairReq* r, InvEntry* inv, bool hit,
                        const RepairArchive& ar) {
	hipStream_t st = S->streams[r->stream];
	int rc = dev_alloc((void**)&r->out, r->out_bytes, st);
	RepairInv ri;
	if (!rc && inv) {
		if (hit) {
			// written by an earlier request, perhaps on another stream
			if (hipStreamWaitEvent(st, inv->ready, 0) != hipSuccess) rc = PARHIP_ERR_DEVICE;
			ri.d = inv->d;
			ri.ready = true;
		} else {
			rc = dev_alloc((void**)&inv->d, inv->bytes, st);
			if (!rc && hipEventCreateWithFlags(&inv->ready, hipEventDisableTiming) != hipSuccess)
				rc = PARHIP_ERR_DEVICE;
			ri.d = inv->d;
		}
	}
	if (!rc)
		rc = repair_enqueue(ar, r->missing.data(), r->missing.size(), nullptr, r->out, ar.W, st,
		                    r->plan, ri);
	if (!rc && inv && !hit) hipEventRecord(inv->ready, st);
	r->rc = rc;
	hipEventRecord(r->done, st);
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
		// admissions, in arrival order; each enqueued with the lock released
		while (RepairReq* r = admit(S)) {
			InvEntry* inv = r->inv;
			const bool hit = r->inv_hit;
			const RepairArchive ar = S->archives[r->archive];
			lk.unlock();
			enqueue_req(S, r, inv, hit, ar);
			lk.lock();
			if (inv && !hit) {
				if (r->rc) inv->valid = false;  // never to be used; dropped at completion
				else inv->filled = true;
			}
			S->inflight.push_back(r);
			S->peak_inflight = std::max<unsigned>(S->peak_inflight, (unsigned)S->inflight.size());
		}
		if (S->stop && S->pending.empty() && S->inflight.empty()) break;
		// nothing in flight: only a submit, a release or the stop can change
		// anything, and each notifies
		if (S->inflight.empty())
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
	// points in the worst alignment of each XOR-block size (7, 15 or 31
	// virtual slots before the first input and the first recovery point): an
	// upper bound for any archive
	return std::max({repair_bytes(numInputs, blockSize64, k, 7, 7, false, 3),
	                 repair_bytes(numInputs, blockSize64, k, 15, 15, false, 4),
	                 repair_bytes(numInputs, blockSize64, k, 31, 31, false, 5)});
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
	r->charge = r->out_bytes + repair_bytes(ar.N, ar.W, k, ar.fi, ar.fr, true);
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

int parhip_repair_server_cache_stats(parhip_repair_server* S, uint64_t* hits, uint64_t* misses,
                                     uint64_t* entries, uint64_t* bytes, uint64_t* evictions) {
	if (!S) return fail(PARHIP_ERR_ARG, "repair_cache_stats: null server");
	std::lock_guard<std::mutex> lk(S->mu);
	if (hits) *hits = S->inv_hits;
	if (misses) *misses = S->inv_misses;
	if (entries) *entries = S->inv_cache.size();
	if (bytes) *bytes = S->inv_bytes;
	if (evictions) *evictions = S->inv_evictions;
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
	for (auto st : S->streams) hipStreamSynchronize(st);
	while (!S->inv_cache.empty()) inv_drop(S, S->inv_cache.begin());
	for (auto st : S->streams) {
		hipStreamSynchronize(st);
		hipStreamDestroy(st);
	}
	delete S;
}

}  // extern "C"
