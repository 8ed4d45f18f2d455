// engine.cpp -- C-ABI implementation of the GF(2^64) engine (include/parhip.h).
//
// Host side of the drop-in for GF64Controller (src/par3_engine.h): argument
// validation, device memory, the per-shape launch plan, the coefficient cache,
// and the host <-> HBM staging for the host-pointer entries.  All arithmetic
// on slices runs in the HIP kernels (gf64_kernels.hip); this file only does
// k x k pivoting on the host (SolveAndReconstruct, gf64_solve).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/parhip.h"
#include "engine_internal.h"
#include "ingest.h"
#include "gf64_host.h"
#include "kernels.h"

namespace parhip {

// ------------------------------------------------------------------ errors
static thread_local std::string t_err;
static thread_local int t_device = 0;
static thread_local int t_method = PARHIP_METHOD_AUTO;

int fail(int code, const char* fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	t_err = buf;
	return code;
}

int hip_fail(hipError_t e, const char* what) {
	return fail(e == hipErrorOutOfMemory ? PARHIP_ERR_NOMEM : PARHIP_ERR_DEVICE, "%s: %s", what,
	            hipGetErrorString(e));
}

int current_device() { return t_device; }
int current_method() { return t_method; }

// ------------------------------------------------------------------ device info
struct DevInfo {
	int cus = 256;
	bool ready = false;
};
static std::mutex g_dev_mu;
static std::map<int, DevInfo> g_dev;

int device_cus(int dev) {
	std::lock_guard<std::mutex> lk(g_dev_mu);
	DevInfo& d = g_dev[dev];
	if (!d.ready) {
		hipDeviceProp_t p;
		if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0)
			d.cus = p.multiProcessorCount;
		// keep freed stream-ordered allocations cached in the pool
		hipMemPool_t pool;
		if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
			uint64_t thr = UINT64_MAX;
			hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
		}
		d.ready = true;
	}
	return d.cus;
}

// Per-thread stream for the synchronous host-pointer entries.
struct ThreadStream {
	hipStream_t s = nullptr;
	int dev = -1;
	~ThreadStream() {
		if (s) hipStreamDestroy(s);
	}
};
static thread_local ThreadStream t_stream;

int host_stream(hipStream_t* out) {
	int dev = t_device;
	hipError_t e = hipSetDevice(dev);
	if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
	if (t_stream.s == nullptr || t_stream.dev != dev) {
		if (t_stream.s) hipStreamDestroy(t_stream.s);
		t_stream.s = nullptr;
		e = hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking);
		if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
		t_stream.dev = dev;
	}
	device_cus(dev);
	*out = t_stream.s;
	return PARHIP_OK;
}

// ------------------------------------------------------------------ memory
int dev_alloc(void** p, size_t bytes, hipStream_t s) {
	*p = nullptr;
	if (bytes == 0) return PARHIP_OK;
	hipError_t e = hipMallocAsync(p, bytes, s);
	if (e != hipSuccess) {
		*p = nullptr;
		return hip_fail(e, "hipMallocAsync");
	}
	return PARHIP_OK;
}

void dev_free(void* p, hipStream_t s) {
	if (p) hipFreeAsync(p, s);
}

// ------------------------------------------------------------------ coefficient cache
// Device-resident Cauchy matrices keyed by (device, N, R, firstInput,
// firstRecovery), LRU of 16 entries.  Mirrors the reference's s_coeffCache
// (src/par3_engine.cc:451-482) but guarded by a mutex, so concurrent callers
// are safe.  Entries are built on the device (k_cauchy); a user waits on the
// entry's `built` event, and an entry is only evicted when no caller holds it
// (reference count) and after the device has drained.
// kind 0: the R x N matrix (k_cauchy); kind 1: the XOR-block coefficients
// of the dyadic contraction (k_dyadic_coeffs, Q x blocks x 8 words).
struct CoeffKey {
	int dev;
	int kind;
	uint64_t n, r, fi, fr;
	bool operator<(const CoeffKey& o) const {
		return std::tie(dev, kind, n, r, fi, fr) < std::tie(o.dev, o.kind, o.n, o.r, o.fi, o.fr);
	}
	bool operator==(const CoeffKey& o) const {
		return std::tie(dev, kind, n, r, fi, fr) == std::tie(o.dev, o.kind, o.n, o.r, o.fi, o.fr);
	}
};
struct CoeffEntry {
	uint64_t* d = nullptr;
	int refs = 0;
	hipEvent_t built = nullptr;  // recorded once, after k_cauchy
};
static std::mutex g_coeff_mu;
static std::map<CoeffKey, CoeffEntry> g_coeff;
static std::list<CoeffKey> g_coeff_lru;
static const size_t kCoeffCacheMax = 16;

// Words of the dyadic coefficients for (N, R, fi, fr).
// kind 1: XOR-block coefficients of 8 x 8 blocks (8 words per (q, b));
// kind 2: of 16 x 16 blocks (16 words per (q, b)); kind 3: of 32 x 32 blocks.
static uint64_t dyadic_coeff_words(int kind, uint64_t n, uint64_t r, uint64_t fi, uint64_t fr) {
	const uint64_t m = (8ull << (kind - 1)) - 1;
	const uint64_t nblocks = ((fi & m) + n + m) / (m + 1), Q = ((fr & m) + r + m) / (m + 1);
	return nblocks * Q * (m + 1);
}

static int coeff_acquire_kind(int kind, uint64_t n, uint64_t r, uint64_t fi, uint64_t fr,
                              hipStream_t s, const uint64_t** out);

int coeff_acquire(uint64_t n, uint64_t r, uint64_t fi, uint64_t fr, hipStream_t s,
                  const uint64_t** out) {
	return coeff_acquire_kind(0, n, r, fi, fr, s, out);
}

// Build a coefficient set of `kind` into `d` on stream s.
static hipError_t coeff_build(int kind, uint64_t* d, uint64_t n, uint64_t r, uint64_t fi,
                              uint64_t fr, hipStream_t s) {
	if (kind == 3) {
		ProfScope ps("dyadic_coeffs", s);
		return launch_dyadic32_coeffs(d, ((fi & 31) + n + 31) / 32, ((fr & 31) + r + 31) / 32,
		                              fi & ~31ull, fr & ~31ull, s);
	}
	if (kind == 2) {
		ProfScope ps("dyadic_coeffs", s);
		return launch_dyadic16_coeffs(d, ((fi & 15) + n + 15) / 16, ((fr & 15) + r + 15) / 16,
		                              fi & ~15ull, fr & ~15ull, s);
	}
	if (kind) {
		ProfScope ps("dyadic_coeffs", s);
		return launch_dyadic_coeffs(d, ((fi & 7) + n + 7) / 8, ((fr & 7) + r + 7) / 8, fi & ~7ull,
		                            fr & ~7ull, s);
	}
	ProfScope ps("cauchy", s);
	return launch_cauchy(d, n, r, fi, fr, s);
}

// parhip_set_coeff_cache: off = every call builds its own coefficients into
// stream-ordered memory, freed by coeff_release.
static thread_local bool t_coeff_cache = true;

static int coeff_acquire_kind(int kind, uint64_t n, uint64_t r, uint64_t fi, uint64_t fr,
                              hipStream_t s, const uint64_t** out) {
	const uint64_t words = kind ? dyadic_coeff_words(kind, n, r, fi, fr) : n * r;
	if (!t_coeff_cache) {
		void* p = nullptr;
		int rc = dev_alloc(&p, words * sizeof(uint64_t), s);
		if (rc) return rc;
		hipError_t e = coeff_build(kind, (uint64_t*)p, n, r, fi, fr, s);
		if (e != hipSuccess) {
			dev_free(p, s);
			return hip_fail(e, kind ? "k_dyadic_coeffs" : "k_cauchy");
		}
		*out = (const uint64_t*)p;
		return PARHIP_OK;
	}
	CoeffKey key{t_device, kind, n, r, fi, fr};
	std::unique_lock<std::mutex> lk(g_coeff_mu);
	auto it = g_coeff.find(key);
	if (it != g_coeff.end()) {
		it->second.refs++;
		g_coeff_lru.remove(key);
		g_coeff_lru.push_front(key);
		hipStreamWaitEvent(s, it->second.built, 0);
		*out = it->second.d;
		return PARHIP_OK;
	}
	// evict unreferenced entries beyond the cap
	while (g_coeff.size() >= kCoeffCacheMax) {
		bool evicted = false;
		for (auto lit = g_coeff_lru.rbegin(); lit != g_coeff_lru.rend(); ++lit) {
			auto eit = g_coeff.find(*lit);
			if (eit != g_coeff.end() && eit->second.refs == 0) {
				hipDeviceSynchronize();
				hipFree(eit->second.d);
				hipEventDestroy(eit->second.built);
				g_coeff.erase(eit);
				g_coeff_lru.erase(std::next(lit).base());
				evicted = true;
				break;
			}
		}
		if (!evicted) break;
	}
	CoeffEntry ent;
	hipError_t e = hipMalloc((void**)&ent.d, words * sizeof(uint64_t));
	if (e != hipSuccess) return hip_fail(e, "hipMalloc(coefficients)");
	e = coeff_build(kind, ent.d, n, r, fi, fr, s);
	if (e != hipSuccess) {
		hipFree(ent.d);
		return hip_fail(e, kind ? "k_dyadic_coeffs" : "k_cauchy");
	}
	hipEventCreateWithFlags(&ent.built, hipEventDisableTiming);
	hipEventRecord(ent.built, s);
	ent.refs = 1;
	g_coeff[key] = ent;
	g_coeff_lru.push_front(key);
	*out = ent.d;
	return PARHIP_OK;
}

void coeff_release(const uint64_t* d, hipStream_t s) {
	{
		std::lock_guard<std::mutex> lk(g_coeff_mu);
		for (auto& kv : g_coeff)
			if (kv.second.d == d) {
				kv.second.refs--;
				return;
			}
	}
	// not a cache entry (entries in use are never evicted): built for one call
	dev_free((void*)d, s);
}

void coeff_cache_clear() {
	std::lock_guard<std::mutex> lk(g_coeff_mu);
	for (auto it = g_coeff.begin(); it != g_coeff.end();) {
		if (it->second.refs == 0) {
			hipDeviceSynchronize();
			hipFree(it->second.d);
			hipEventDestroy(it->second.built);
			g_coeff_lru.remove(it->first);
			it = g_coeff.erase(it);
		} else {
			++it;
		}
	}
}

// ------------------------------------------------------------------ kernel timing
// Per-thread HIP-event timing of named kernels, enabled by parhip_prof_enable.
// Events are recorded on the stream the kernel is launched on, so the figure
// is the kernel's own device duration (plus launch gaps inside the bracket).
struct ProfRec {
	std::string name;
	hipEvent_t a, b;
};
static thread_local bool t_prof = false;
static thread_local std::string t_prof_only;  // non-empty: time only this kernel
static thread_local std::vector<ProfRec> t_prof_pending;
static thread_local std::map<std::string, std::pair<double, uint64_t>> t_prof_tot;

ProfScope::ProfScope(const char* name, hipStream_t s) : s_(s) {
	if (!t_prof || (!t_prof_only.empty() && t_prof_only != name)) return;
	name_ = name;
	hipEventCreate(&a_);
	hipEventCreate(&b_);
	hipEventRecord(a_, s_);
	on_ = true;
}
ProfScope::~ProfScope() {
	if (!on_) return;
	hipEventRecord(b_, s_);
	t_prof_pending.push_back({name_, a_, b_});
}

bool prof_set(bool on) {
	const bool was = t_prof;
	t_prof = on;
	return was;
}

static void prof_collect() {
	for (auto& r : t_prof_pending) {
		hipEventSynchronize(r.b);
		float ms = 0.f;
		hipEventElapsedTime(&ms, r.a, r.b);
		auto& t = t_prof_tot[r.name];
		t.first += ms;
		t.second += 1;
		hipEventDestroy(r.a);
		hipEventDestroy(r.b);
	}
	t_prof_pending.clear();
}

// ------------------------------------------------------------------ contraction plan
// Launch-shape override for tuning (parhip_tune_override); 0 fields = automatic.
static thread_local int t_tune_rt = 0, t_tune_wpt = 0, t_tune_ib = 0;
static thread_local int64_t t_tune_groups = 0;
int64_t tune_groups_override() { return t_tune_groups; }
void tune_shape_override(int* rt, int* wpt, int* ib) {
	*rt = t_tune_rt;
	*wpt = t_tune_wpt;
	*ib = t_tune_ib;
}

Plan plan_contraction(uint64_t N, uint64_t R, uint64_t W, int dev) {
	Plan p;
	p.method = PARHIP_METHOD_TABLE;
	// output tile: the fewest padded rows, 8 on a tie.  Measured at 1024 x 1 MiB
	// the cost per computed row is flat for 8- and 4-row tiles (~0.16 ms) and
	// ~40% higher for 16-row tiles, so padding is what matters.
	if (R <= 2) {
		p.rt = 2;
	} else {
		const uint64_t pad8 = (R + 7) / 8 * 8, pad4 = (R + 3) / 4 * 4;
		p.rt = pad8 <= pad4 ? 8 : 4;
	}
	p.wpt = W >= 1024 ? 4 : 2;
	// narrow tiles (2 words per thread) amortise a table over half the words:
	// 2-input batches halve the LDS per work-group so more of them stay
	// resident (measured 262144 x 4 KiB, R = 8: 1.15 vs 1.24 ms)
	p.ib = (p.rt <= 8 && p.wpt == 4) ? 4 : 2;
	if (!mac_tab_supported(p.rt, p.wpt, p.ib)) p.ib = 4;  // 2-output tiles: IB = 4 only
	if (t_tune_rt || t_tune_wpt || t_tune_ib) {
		int rt = t_tune_rt ? t_tune_rt : p.rt, wpt = t_tune_wpt ? t_tune_wpt : p.wpt;
		int ib = t_tune_ib ? t_tune_ib : p.ib;
		if (mac_tab_supported(rt, wpt, ib)) {
			p.rt = rt;
			p.wpt = wpt;
			p.ib = ib;
		}
	}
	if (N == 0 || R == 0 || W == 0) {  // nothing to launch
		p.groups = 1;
		p.in_per_group = N;
		return p;
	}
	const uint64_t tile = 256ull * p.wpt;
	const uint64_t tiles = (W + tile - 1) / tile;
	const uint64_t rtiles = (R + p.rt - 1) / p.rt;
	const int ib = p.ib;
	// work-groups resident per CU: LDS-bound (IB*RT*2 KiB per group, 160 KiB per CU)
	const int lds_kib = ib * p.rt * 2;
	int per_cu = 160 / lds_kib;
	if (per_cu < 1) per_cu = 1;
	if (per_cu > 4) per_cu = 4;
	// two rounds of work-groups per resident slot: work-groups that start
	// later run their table builds out of step with the ones beside them,
	// which keeps the LDS busier (R = 4: 0.61 vs 0.65 ms; 262144 x 4 KiB:
	// 1.14 vs 1.18 ms including the larger reduction)
	const uint64_t target = (uint64_t)device_cus(dev) * per_cu * 2;
	uint64_t groups = 1;
	const uint64_t base = tiles * rtiles;
	if (base < target) groups = (target + base - 1) / base;
	if (t_tune_groups > 0) groups = (uint64_t)t_tune_groups;
	// at least 2*IB inputs per group, so table building stays amortised
	const uint64_t max_groups = std::max<uint64_t>(1, N / (2 * ib));
	groups = std::min(groups, max_groups);
	groups = std::min<uint64_t>(groups, 65535);
	uint64_t per = (N + groups - 1) / groups;
	per = (per + ib - 1) / ib * ib;
	groups = (N + per - 1) / per;
	p.groups = groups;
	p.in_per_group = per;
	return p;
}

size_t plan_workspace_bytes(const Plan& p, uint64_t R, uint64_t W) {
	return p.groups > 1 ? (size_t)(xor_reduce_scratch_words(p.groups, R, W) * sizeof(uint64_t)) : 0;
}

// GF(2) bit-matrix on MFMA (gf64_mfma.hip): bit-matrix rows of every
// coefficient prepared on the device, then the MFMA contraction, split-K
// partials folded as for the table kernel.  Strided slices only.
int plan_mfma(uint64_t N, uint64_t R, uint64_t W, bool fp4, bool overrides, MfmaPlan* p) {
	// measured best at 1024 x 1 MiB, R = 8 and 32: fp4 4x2 tiles with 4-input
	// stages; int8 (twice the fragments) 4x1 tiles with 2-input stages
	int rt = 4, nb = fp4 ? 2 : 1, ib = fp4 ? 4 : 2;
	if (overrides && (t_tune_rt || t_tune_wpt || t_tune_ib) &&
	    mfma_supported(t_tune_rt ? t_tune_rt : rt, t_tune_wpt ? t_tune_wpt : nb,
	                   t_tune_ib ? t_tune_ib : ib, fp4)) {
		rt = t_tune_rt ? t_tune_rt : rt;
		nb = t_tune_wpt ? t_tune_wpt : nb;
		ib = t_tune_ib ? t_tune_ib : ib;
	}
	const uint64_t span = 4ull * 32 * nb;
	const uint64_t base = ((W + span - 1) / span) * ((R + rt - 1) / rt);
	const uint64_t target = (uint64_t)device_cus(t_device) * 2;
	uint64_t groups = base < target ? (target + base - 1) / base : 1;
	if (overrides && t_tune_groups > 0) groups = (uint64_t)t_tune_groups;
	groups = std::min<uint64_t>(groups, std::max<uint64_t>(1, N / 8));
	// fp4 counts are fp32 sums of 0/1 products: exact while 64 * inputs < 2^24
	if (fp4) groups = std::max<uint64_t>(groups, (N + (1u << 17) - 1) >> 17);
	groups = std::min<uint64_t>(groups, 65535);
	uint64_t per = (N + groups - 1) / groups;
	groups = (N + per - 1) / per;
	if (fp4 && per >= (1u << 18))
		return fail(PARHIP_ERR_ARG, "mfma: %llu inputs per group exceed the exact fp32 count",
		            (unsigned long long)per);
	p->fp4 = fp4;
	p->rt = rt;
	p->nb = nb;
	p->ib = ib;
	p->groups = groups;
	p->per = per;
	p->rows_bytes = (size_t)R * N * 64 * sizeof(uint64_t);
	p->part_bytes = groups > 1 ? (size_t)xor_reduce_scratch_words(groups, R, W) * sizeof(uint64_t) : 0;
	return PARHIP_OK;
}

hipError_t enqueue_mfma(const MacArgs& m, const MfmaPlan& p, void* ws, hipStream_t s) {
	uint64_t* rows = (uint64_t*)ws;
	hipError_t e;
	{
		ProfScope ps("mfma_rows", s);
		e = launch_gf64_rows(m.coeff, m.coeff_ld, m.N, m.R, rows, s);
	}
	MfmaArgs a{};
	a.in_base = m.in_base;
	a.in_stride = m.in_stride;
	a.out_base = m.out_base;
	a.out_stride = m.out_stride;
	a.rows = rows;
	a.rows_ld = m.N;
	a.N = m.N;
	a.R = m.R;
	a.W = m.W;
	a.in_per_group = p.per;
	a.partial = p.groups > 1 ? (uint64_t*)((char*)ws + p.rows_bytes) : nullptr;
	a.accumulate = m.accumulate;
	if (e == hipSuccess) {
		ProfScope ps("mac_mfma", s);
		e = launch_mac_mfma(a, p.groups, p.rt, p.nb, p.ib, p.fp4, s);
	}
	if (e == hipSuccess && p.groups > 1) {
		ProfScope ps("xor_reduce", s);
		e = launch_xor_reduce(a.partial, p.groups, m.R, m.W, m.out_base, m.out_stride, nullptr,
		                      m.accumulate, s);
	}
	return e;
}

static int run_contraction_mfma(const MacArgs& m, bool fp4, hipStream_t s) {
	MfmaPlan p;
	int rc = plan_mfma(m.N, m.R, m.W, fp4, true, &p);
	if (rc) return rc;
	void* ws = nullptr;
	rc = dev_alloc(&ws, p.rows_bytes + p.part_bytes, s);
	if (rc) return rc;
	hipError_t e = enqueue_mfma(m, p, ws, s);
	dev_free(ws, s);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "k_mac_mfma");
}

static int run_contraction_tab8(MacArgs a, hipStream_t s) {
	int rt = (a.R > 2 && a.R <= 8) ? 4 : 2, wpt = a.W >= 2048 ? 8 : (a.W >= 1024 ? 4 : 2), ib = 1;
	if ((t_tune_rt || t_tune_wpt || t_tune_ib) &&
	    mac_tab8_supported(t_tune_rt ? t_tune_rt : rt, t_tune_wpt ? t_tune_wpt : wpt,
	                       t_tune_ib ? t_tune_ib : ib)) {
		rt = t_tune_rt ? t_tune_rt : rt;
		wpt = t_tune_wpt ? t_tune_wpt : wpt;
		ib = t_tune_ib ? t_tune_ib : ib;
	}
	const uint64_t base = ((a.W + 256ull * wpt - 1) / (256ull * wpt)) * ((a.R + rt - 1) / rt);
	const uint64_t per_cu = std::max<uint64_t>(1, 160 / ((uint64_t)ib * rt / 2 * 32));
	const uint64_t target = (uint64_t)device_cus(t_device) * per_cu;
	uint64_t groups = base < target ? (target + base - 1) / base : 1;
	if (t_tune_groups > 0) groups = (uint64_t)t_tune_groups;
	groups = std::min<uint64_t>(groups, std::max<uint64_t>(1, a.N / (2 * ib)));
	groups = std::min<uint64_t>(groups, 65535);
	uint64_t per = (a.N + groups - 1) / groups;
	per = (per + ib - 1) / ib * ib;
	groups = (a.N + per - 1) / per;
	a.in_per_group = per;
	a.partial = nullptr;
	void* ws = nullptr;
	if (groups > 1) {
		int rc = dev_alloc(&ws, (size_t)xor_reduce_scratch_words(groups, a.R, a.W) * sizeof(uint64_t), s);
		if (rc) return rc;
		a.partial = (uint64_t*)ws;
	}
	hipError_t e;
	{
		ProfScope ps("mac_tab8", s);
		e = launch_mac_tab8(a, groups, rt, wpt, ib, s);
	}
	if (e == hipSuccess && groups > 1) {
		ProfScope ps("xor_reduce", s);
		e = launch_xor_reduce(a.partial, groups, a.R, a.W, a.out_base, a.out_stride, nullptr,
		                      a.accumulate, s);
	}
	dev_free(ws, s);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "k_mac_tab8");
}

// Enqueue the full contraction described by `a` (partial/in_per_group are
// filled in here) on stream s.
int run_contraction(MacArgs a, hipStream_t s) {
	if (a.N == 0 || a.R == 0) return PARHIP_OK;
	if (a.W == 0) return PARHIP_OK;
	if ((t_method == PARHIP_METHOD_MFMA || t_method == PARHIP_METHOD_MFMA_I8) && !a.in_ptrs &&
	    !a.out_ptrs)
		return run_contraction_mfma(a, t_method == PARHIP_METHOD_MFMA, s);
	if (t_method == PARHIP_METHOD_TABLE8 && !a.in_ptrs && !a.out_ptrs) return run_contraction_tab8(a, s);
	return run_contraction_tables(a, s);
}

int run_contraction_tables(MacArgs a, hipStream_t s) {
	if (a.N == 0 || a.R == 0 || a.W == 0) return PARHIP_OK;
	Plan p = plan_contraction(a.N, a.R, a.W, t_device);
	a.in_per_group = p.in_per_group;
	a.partial = nullptr;
	void* ws = nullptr;
	if (p.groups > 1) {
		int rc = dev_alloc(&ws, plan_workspace_bytes(p, a.R, a.W), s);
		if (rc) return rc;
		a.partial = (uint64_t*)ws;
	}
	hipError_t e;
	{
		ProfScope ps("mac_tab", s);
		e = launch_mac_tab(a, p.groups, p.rt, p.wpt, p.ib, s);
	}
	if (e != hipSuccess) {
		dev_free(ws, s);
		return hip_fail(e, "k_mac_tab");
	}
	if (p.groups > 1) {
		ProfScope ps("xor_reduce", s);
		e = launch_xor_reduce(a.partial, p.groups, a.R, a.W, a.out_base, a.out_stride, a.out_ptrs,
		                      a.accumulate, s);
		dev_free(ws, s);
		if (e != hipSuccess) return hip_fail(e, "k_xor_reduce");
	}
	return PARHIP_OK;
}

// ------------------------------------------------------------------ dyadic plan
// Launch shape of the XOR-block contraction: 4 words per thread for slices of
// at least 1024 words (tables amortised over 1024 words), else 2; split-K rows
// for one round of work-groups per resident slot (2 per CU at 205 / 129
// VGPRs): at 1024 x 1 MiB, R = 8, 4 rows measured 0.530 + 0.010 ms (kernel +
// reduce) against 0.535 + 0.015 ms for 8.
// Split-K rows from which the XOR-block launches fold with atomics.
static const uint64_t kDyAtomicRows = 16;

int dyadic_block_bits(uint64_t R, uint64_t fr, int max_bits) {
	if (t_method == PARHIP_METHOD_DYADIC8) return 3;
	if (t_method == PARHIP_METHOD_DYADIC16) return std::min(4, max_bits);
	if (t_method == PARHIP_METHOD_DYADIC32) return std::min(5, max_bits);
	// table reads per input: Q_b output groups x 3^b products per 2^b inputs;
	// a larger block where it needs at least 3 % fewer
	int bits = 3;
	double best = (double)(((fr & 7) + R + 7) / 8) * 27 / 8;
	for (int b = 4; b <= max_bits; b++) {
		const uint64_t m = (1ull << b) - 1;
		const double cost = (double)(((fr & m) + R + m) >> b) * (b == 4 ? 81.0 / 16 : 243.0 / 32);
		if (cost < 0.97 * best) {
			best = cost;
			bits = b;
		}
	}
	return bits;
}

bool dyadic16_better(uint64_t R, uint64_t fr) { return dyadic_block_bits(R, fr, 4) == 4; }

DyadicPlan plan_dyadic(uint64_t N, uint64_t R, uint64_t W, uint64_t fi, uint64_t fr, int dev,
                       size_t max_part_bytes, bool allow16, int force_bits, bool allow32) {
	DyadicPlan p;
	// Block size: 32 x 32 blocks take 243 products per 1024 input-output
	// pairs, 16 x 16 blocks 81 per 256, 8 x 8 blocks 27 per 64; per input that
	// is Q32 * 243/32, Q16 * 81/16 and Q8 * 27/8 (R <= 8: 8 x 8; R = 9..16:
	// 16 x 16; R = 17..32: 32 x 32, a quarter less than 16 x 16 blocks).
	if (allow16) p.bits = dyadic_block_bits(R, fr, allow32 ? 5 : 4);
	if (force_bits >= 3 && force_bits <= 5) p.bits = force_bits;
	const uint64_t m = (1ull << p.bits) - 1;
	p.pre_in = fi & m;
	p.pre_out = fr & m;
	p.nblocks = (p.pre_in + N + m) / (m + 1);
	p.Q = (p.pre_out + R + m) / (m + 1);
	p.ib = 1;
	if (p.bits == 5) {
		p.wpt = 1;  // 512 threads, one word each
	} else if (p.bits == 4) {
		p.wpt = 1;
		if (t_tune_wpt == 1 || t_tune_wpt == 2) p.wpt = t_tune_wpt;
	} else {
		p.wpt = W >= 1024 ? 4 : 2;
		if (t_tune_wpt || t_tune_ib) {
			const int wpt = t_tune_wpt ? t_tune_wpt : p.wpt, ib = t_tune_ib ? t_tune_ib : p.ib;
			if (mac_dyadic_supported(wpt, ib)) {
				p.wpt = wpt;
				p.ib = ib;
			}
		}
	}
	const uint64_t tile_words = p.bits == 5 ? 512 : 256ull * p.wpt;
	const uint64_t tiles = (W + tile_words - 1) / tile_words;
	const uint64_t base = tiles * p.Q;
	// Resident work-groups: 2 per CU at 4 words per thread (205 VGPRs), 4 at
	// 2 words per thread (127 VGPRs, 28 KiB of LDS each; 262144 x 4 KiB,
	// R = 8: 0.56 -> 0.50 ms against 2 per CU).  The split-K row count g
	// minimises the blocks one slot works through, rounds(g) * blocks/g
	// (whole rounds: a round that is half full costs a full one), plus the
	// reduce, which reads g partial rows: 1024 x 1 MiB at R = 24 goes from
	// 2 rows (768 work-groups, 1.5 rounds: 1.95 ms) to 4 (3 rounds: 1.54 ms).
	// (16 x 16 blocks: 58 KiB of LDS, two work-groups per CU at either wpt;
	// 32 x 32 blocks: 116 KiB, one work-group of 512 threads)
	const uint64_t slots =
	    (uint64_t)device_cus(dev) * (p.bits == 5 ? 1 : (p.bits == 3 && p.wpt == 2 ? 4 : 2));
	const uint64_t gmax = std::min<uint64_t>(std::max<uint64_t>(1, p.nblocks / (2 * p.ib)), 65535);
	// a slot's block step is ~16 us of 1024-word tile work (8 x 8 blocks; a
	// 16 x 16 block of a 256-word tile is 3/4 of that per word); the reduce
	// reads R x W words per row at ~4 TB/s
	const double step_s = p.bits == 5 ? 16e-6 * 4.5 : p.bits == 4 ? 16e-6 * 0.75 * p.wpt : 16e-6;
	const double row_cost = (double)R * W * 8 / 4e12 / step_s;
	uint64_t groups = 1;
	double best = 0;
	auto consider = [&](uint64_t g) {
		if (g < 1 || g > gmax) return;
		const uint64_t rounds = (base * g + slots - 1) / slots;
		const uint64_t per = (p.nblocks + g - 1) / g;
		const double cost = (double)rounds * per + (g > 1 ? row_cost * g : 0.0);
		if (cost < best || (cost == best && g < groups)) {
			best = cost;
			groups = g;
		}
	};
	best = (double)((base + slots - 1) / slots) * p.nblocks;
	// small counts, and the counts that fill whole rounds
	for (uint64_t g = 2; g <= 64; g++) consider(g);
	for (uint64_t k = 1; k <= 16; k++) consider(k * slots / base);
	if (t_tune_groups > 0) groups = (uint64_t)t_tune_groups;
	// at least 2 batches of blocks per row, so the table builds stay amortised
	groups = std::min<uint64_t>(groups, std::max<uint64_t>(1, p.nblocks / (2 * p.ib)));
	groups = std::min<uint64_t>(groups, 65535);
	// fewer rows where the caller's partial buffer is smaller (one row needs none)
	while (groups > 1 && groups < kDyAtomicRows &&
	       (size_t)xor_reduce_scratch_words(groups, R, W) * sizeof(uint64_t) > max_part_bytes)
		groups--;
	uint64_t per = (p.nblocks + groups - 1) / groups;
	per = (per + p.ib - 1) / p.ib * p.ib;
	p.groups = (p.nblocks + per - 1) / per;
	p.blocks_per_group = per;
	p.part_bytes = p.groups > 1 && p.groups < kDyAtomicRows
	                   ? (size_t)xor_reduce_scratch_words(p.groups, R, W) * sizeof(uint64_t)
	                   : 0;
	return p;
}

bool dyadic_shape_ok(uint64_t R) {
	// 27 products per 8 x 8 block against 8 per input and output: the direct
	// tables cost less below 4 outputs
	if (R < 4) return false;
	const uint64_t Q = ((R + 7) / 8) + 1;
	return Q <= 65535;
}

bool dyadic_applies_shape(uint64_t R) {
	if (t_method != PARHIP_METHOD_AUTO && t_method != PARHIP_METHOD_DYADIC &&
	    t_method != PARHIP_METHOD_DYADIC8 && t_method != PARHIP_METHOD_DYADIC16 &&
	    t_method != PARHIP_METHOD_DYADIC32)
		return false;
	return dyadic_shape_ok(R);
}

bool dyadic_applies(const MacArgs& a) {
	return !a.in_ptrs && !a.out_ptrs && dyadic_applies_shape(a.R);
}

hipError_t enqueue_dyadic(const MacArgs& a, const DyadicPlan& p, const uint64_t* gp,
                          uint64_t gp_ld, void* ws, hipStream_t s, const uint32_t* lost,
                          uint64_t xbase, uint64_t ybase) {
	DyadicArgs d{};
	d.lost = lost;
	d.xbase = xbase;
	d.ybase = ybase;
	d.in_base = a.in_base;
	d.in_stride = a.in_stride;
	d.out_base = a.out_base;
	d.out_stride = a.out_stride;
	d.gp = gp;
	d.gp_ld = gp_ld;
	d.N = a.N;
	d.R = a.R;
	d.W = a.W;
	d.pre_in = p.pre_in;
	d.pre_out = p.pre_out;
	d.nblocks = p.nblocks;
	d.blocks_per_group = p.blocks_per_group;
	// many split-K rows (narrow slices: 262144 x 4 KiB has 1024): the rows XOR
	// into the zeroed outputs with atomics (0.514 -> 0.504 ms with the zeroing)
	// instead of partials and a reduce that reads all of them; a few rows (C2a:
	// 4): partials and the reduce
	const bool atomic = p.groups >= kDyAtomicRows;
	d.partial = p.groups > 1 && !atomic ? (uint64_t*)ws : nullptr;
	d.accumulate = a.accumulate;
	d.atomic_out = atomic;
	hipError_t e = hipSuccess;
	if (atomic && !a.accumulate) {
		ProfScope ps("zero_rows", s);
		e = launch_zero_rows(a.out_base, a.R, a.W, a.out_stride, s);
	}
	if (e == hipSuccess && p.bits == 5) {
		ProfScope ps("mac_dyadic32", s);
		e = launch_mac_dyadic32(d, p.groups, s);
	} else if (e == hipSuccess && p.bits == 4) {
		ProfScope ps("mac_dyadic16", s);
		e = launch_mac_dyadic16(d, p.groups, p.wpt, s);
	} else if (e == hipSuccess) {
		ProfScope ps("mac_dyadic", s);
		e = launch_mac_dyadic(d, p.groups, p.wpt, p.ib, s);
	}
	if (e == hipSuccess && p.groups > 1 && !atomic) {
		ProfScope ps("xor_reduce", s);
		e = launch_xor_reduce(d.partial, p.groups, a.R, a.W, a.out_base, a.out_stride, nullptr,
		                      a.accumulate, s);
	}
	return e;
}

// The plain Cauchy contraction of a call with n_total inputs at points
// (fi + i, fr + r), over its inputs [first, first + a.N) (a chunk of a streamed
// call; first = 0 and n_total = a.N for a whole call): the XOR-block kernel
// where it applies, else the cached R x n_total matrix and run_contraction.
// Either way one cached coefficient set serves every chunk of the call.
int run_cauchy_contraction(MacArgs a, uint64_t fi, uint64_t fr, hipStream_t s, uint64_t first,
                           uint64_t n_total) {
	if (a.N == 0 || a.R == 0 || a.W == 0) return PARHIP_OK;
	if (n_total == 0) n_total = a.N;
	if (dyadic_applies(a)) {
		// the chunk's own virtual blocks start at block b0 of the call's
		const uint64_t cfi = fi + first;
		const DyadicPlan p = plan_dyadic(a.N, a.R, a.W, cfi, fr, t_device, SIZE_MAX, true, 0, true);
		const uint64_t m = (1ull << p.bits) - 1;
		const uint64_t b0 = ((cfi & ~m) - (fi & ~m)) / (m + 1);
		// cached coefficients where the cache is on; else the kernel computes
		// them itself (one inversion per thread per chunk of blocks: no
		// separate launch, no coefficient memory)
		const uint64_t* gp = nullptr;
		if (t_coeff_cache) {
			int rc = coeff_acquire_kind(p.bits - 2, n_total, a.R, fi, fr, s, &gp);
			if (rc) return rc;
		}
		void* ws = nullptr;
		int rc = dev_alloc(&ws, p.part_bytes, s);
		if (rc) {
			if (gp) coeff_release(gp, s);
			return rc;
		}
		const uint64_t call_blocks = ((fi & m) + n_total + m) / (m + 1);
		const hipError_t e = enqueue_dyadic(a, p, gp ? gp + b0 * (m + 1) : nullptr, call_blocks, ws, s,
		                                    nullptr, cfi & ~m, fr & ~m);
		dev_free(ws, s);
		if (gp) coeff_release(gp, s);
		return e == hipSuccess ? PARHIP_OK
		                       : hip_fail(e, p.bits == 5 ? "k_mac_dyadic32"
		                                                 : p.bits == 4 ? "k_mac_dyadic16" : "k_mac_dyadic");
	}
	const uint64_t* coeff = nullptr;
	int rc = coeff_acquire(n_total, a.R, fi, fr, s, &coeff);
	if (rc) return rc;
	a.coeff = coeff + first;
	a.coeff_ld = n_total;
	rc = run_contraction(a, s);
	coeff_release(coeff, s);
	return rc;
}

}  // namespace parhip

using namespace parhip;

// ====================================================================== C ABI
extern "C" {

int parhip_version(void) { return 100; }

const char* parhip_last_error(void) { return t_err.c_str(); }

int parhip_set_device(int device) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
		return fail(PARHIP_ERR_ARG, "parhip_set_device: no device %d", device);
	t_device = device;
	return PARHIP_OK;
}

int parhip_set_method(int method) {
	if (method < PARHIP_METHOD_AUTO || method > PARHIP_METHOD_DYADIC32)
		return fail(PARHIP_ERR_ARG, "parhip_set_method: unknown method %d", method);
	t_method = method;
	return PARHIP_OK;
}

int parhip_prof_enable(int on) {
	t_prof = on != 0;
	return PARHIP_OK;
}

int parhip_prof_only(const char* kernel) {
	t_prof_only = kernel ? kernel : "";
	return PARHIP_OK;
}

int parhip_prof_reset(void) {
	prof_collect();
	t_prof_tot.clear();
	return PARHIP_OK;
}

int parhip_prof_get(const char* kernel, double* total_ms, uint64_t* launches) {
	if (!kernel || !total_ms || !launches) return fail(PARHIP_ERR_ARG, "prof_get: null pointer");
	prof_collect();
	auto it = t_prof_tot.find(kernel);
	*total_ms = it == t_prof_tot.end() ? 0.0 : it->second.first;
	*launches = it == t_prof_tot.end() ? 0 : it->second.second;
	return PARHIP_OK;
}

int parhip_tune_override(int rt, int wpt, int ib, int64_t groups) {
	// (rt = ib = 0, wpt 1 / 2: the 16 x 16 XOR-block form's tile widths)
	const bool dy16 = !rt && !ib && (wpt == 1 || wpt == 2);
	if ((rt || wpt || ib) && !dy16 && !mac_tab_supported(rt ? rt : 8, wpt ? wpt : 4, ib ? ib : 4) &&
	    !mac_tab8_supported(rt ? rt : 4, wpt ? wpt : 4, ib ? ib : 1) &&
	    !gf16_mfma_supported(rt ? rt : 4, wpt ? wpt : 4, true) &&
	    !gf16_mfma_supported(rt ? rt : 4, wpt ? wpt : 2, false) &&
	    !mfma_supported(rt ? rt : 4, wpt ? wpt : 2, ib ? ib : 2, true) &&
	    !mfma_supported(rt ? rt : 4, wpt ? wpt : 2, ib ? ib : 2, false))
		return fail(PARHIP_ERR_ARG, "tune_override: unsupported (rt=%d wpt=%d ib=%d)", rt, wpt, ib);
	t_tune_rt = rt;
	t_tune_wpt = wpt;
	t_tune_ib = ib;
	t_tune_groups = groups;
	return PARHIP_OK;
}

int parhip_set_coeff_cache(int enabled) {
	t_coeff_cache = enabled != 0;
	return PARHIP_OK;
}

int parhip_release_cache(void) {
	coeff_cache_clear();
	ingest_release();
	return PARHIP_OK;
}

void parhip_free(void* p) { free(p); }

uint64_t parhip_gf64_mul(uint64_t a, uint64_t b) { return host_gf64_mul(a, b); }
uint64_t parhip_gf64_inverse(uint64_t a) { return host_gf64_inv(a); }

size_t parhip_gf64_workspace_bytes(size_t numInputs, size_t numRecovery, size_t blockSize64) {
	Plan p = plan_contraction(numInputs, numRecovery, blockSize64, t_device);
	return plan_workspace_bytes(p, numRecovery, blockSize64);
}

// ------------------------------------------------------------- device entries
int parhip_d_gf64_build_cauchy(uint64_t* d_coeff, size_t numInputs, size_t numRecovery,
                               uint64_t firstInput, uint64_t firstRecovery, void* stream) {
	if (numInputs == 0 || numRecovery == 0) return PARHIP_OK;
	if (!d_coeff) return fail(PARHIP_ERR_ARG, "build_cauchy: null output");
	ProfScope ps("cauchy", (hipStream_t)stream);
	hipError_t e = launch_cauchy(d_coeff, numInputs, numRecovery, firstInput, firstRecovery,
	                             (hipStream_t)stream);
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "k_cauchy");
}

int parhip_d_gf64_recovery_with_coeff(const uint64_t* d_inputs, size_t numInputs,
                                      size_t inputStride64, uint64_t* d_recovery,
                                      size_t numRecovery, size_t recoveryStride64,
                                      size_t blockSize64, const uint64_t* d_coeff, int accumulate,
                                      void* stream) {
	if (numInputs == 0 || numRecovery == 0) return PARHIP_OK;
	if (!d_inputs || !d_recovery || !d_coeff)
		return fail(PARHIP_ERR_ARG, "recovery_with_coeff: null pointer");
	if (inputStride64 < blockSize64 || recoveryStride64 < blockSize64)
		return fail(PARHIP_ERR_ARG, "recovery_with_coeff: stride < blockSize64");
	device_cus(t_device);
	MacArgs a{};
	a.in_base = d_inputs;
	a.in_stride = inputStride64;
	a.out_base = d_recovery;
	a.out_stride = recoveryStride64;
	a.coeff = d_coeff;
	a.coeff_ld = numInputs;
	a.N = numInputs;
	a.R = numRecovery;
	a.W = blockSize64;
	a.accumulate = accumulate ? 1 : 0;
	return run_contraction(a, (hipStream_t)stream);
}

int parhip_d_gf64_recovery(const uint64_t* d_inputs, size_t numInputs, size_t inputStride64,
                           uint64_t* d_recovery, size_t numRecovery, size_t recoveryStride64,
                           size_t blockSize64, uint64_t firstInput, uint64_t firstRecovery,
                           int accumulate, void* stream) {
	if (numInputs == 0 || numRecovery == 0) return PARHIP_OK;
	if (!d_inputs || !d_recovery) return fail(PARHIP_ERR_ARG, "recovery: null pointer");
	if (inputStride64 < blockSize64 || recoveryStride64 < blockSize64)
		return fail(PARHIP_ERR_ARG, "recovery: stride < blockSize64");
	device_cus(t_device);
	MacArgs a{};
	a.in_base = d_inputs;
	a.in_stride = inputStride64;
	a.out_base = d_recovery;
	a.out_stride = recoveryStride64;
	a.N = numInputs;
	a.R = numRecovery;
	a.W = blockSize64;
	a.accumulate = accumulate ? 1 : 0;
	return run_cauchy_contraction(a, firstInput, firstRecovery, (hipStream_t)stream, 0, numInputs);
}

int parhip_d_gf64_recovery_ptrs(const uint64_t* const* d_in_ptrs, size_t numIn,
                                uint64_t* const* d_out_ptrs, size_t numOut, size_t blockSize64,
                                const uint64_t* d_coeff, size_t coeff_ld, int accumulate,
                                void* stream) {
	if (numIn == 0 || numOut == 0) return PARHIP_OK;
	if (!d_in_ptrs || !d_out_ptrs || !d_coeff)
		return fail(PARHIP_ERR_ARG, "recovery_ptrs: null pointer");
	if (coeff_ld < numIn) return fail(PARHIP_ERR_ARG, "recovery_ptrs: coeff_ld < numIn");
	device_cus(t_device);
	MacArgs a{};
	a.in_ptrs = d_in_ptrs;
	a.out_ptrs = d_out_ptrs;
	a.coeff = d_coeff;
	a.coeff_ld = coeff_ld;
	a.N = numIn;
	a.R = numOut;
	a.W = blockSize64;
	a.accumulate = accumulate ? 1 : 0;
	return run_contraction(a, (hipStream_t)stream);
}

}  // extern "C"
