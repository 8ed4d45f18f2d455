// batch.cpp -- batched PAR3 create: many independent jobs of uneven input
// count, slice size and recovery count in one persistent kernel launch.
//
// The reference runs each create as its own call (lib/par3gen.js:2441 ->
// ComputeRecoveryBlocksWithCoeff, src/par3_engine.cc:1165); here a batch of
// jobs is planned once into a device queue of tiles (job, word range, output
// range, input range), each tile using the launch variant the autotuner chose
// for its job's shape, and the whole sequence -- every job's coefficients,
// zeroing of split jobs' outputs, queue reset, persistent kernel -- is
// captured in a hipGraph and replayed.  Jobs whose slices allow it run as
// XOR-block (dyadic) tiles, the same method as the plain create
// (gf64_dyadic_dev.h); the others on the nibble-table tiles.  Jobs whose input range is split over
// several tiles accumulate with 64-bit atomic XOR (order-independent: the
// bytes are identical to the unsplit computation).  A job can instead run on
// the GF(2) bit-matrix (MFMA) strategy, as its own launch sequence after the
// persistent kernel, when the autotuner measures it faster for the job's
// shape or the caller asks for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/parhip.h"
#include "../../include/parhip_batch.h"
#include "engine_internal.h"
#include "kernels.h"

namespace parhip {

struct CfgInfo {
	int rt, wpt, ib;
};
// table variants: outputs per work-group, words per thread, inputs per LDS
// batch; MFMA: outputs per work-group, 32-word column blocks per wave, inputs
// per stage of the fp4 kernel's default shape
static const CfgInfo kCfg[BATCH_CFG_COUNT] = {
	{8, 4, 4}, {8, 2, 4}, {4, 2, 4} /* MFMA */, {16, 2, 2}, {4, 4, 4},
	{8, 4, 1} /* dyadic, one 8-output group per tile */, {8, 2, 1} /* dyadic */,
	{16, 1, 1} /* dyadic 16 x 16, one 16-output group per tile */,
	{32, 1, 1} /* dyadic 32 x 32: 512 threads, one 32-output group per tile */};

static bool is_dyadic(int cfg) {
	return cfg == BATCH_CFG_DY4 || cfg == BATCH_CFG_DY2 || cfg == BATCH_CFG_DY16 || cfg == BATCH_CFG_DY32;
}

// words per tile of a variant (a 32 x 32 tile: 512 threads of one word)
static uint64_t tile_words(int cfg) { return cfg == BATCH_CFG_DY32 ? 512 : 256ull * kCfg[cfg].wpt; }

// XOR-block tiles run the straight-line (full, 16-byte aligned tile) form of
// the kernel: slices a multiple of 512 words, inputs 16-byte aligned at an
// even stride; 27 products per 8 x 8 block pay off from 4 outputs up (the
// rule of the plain create).
static bool dyadic_ok(const BatchJob& jb) {
	return jb.R >= 4 && jb.W % 512 == 0 && (reinterpret_cast<uintptr_t>(jb.in) & 15u) == 0 &&
	       (jb.in_stride & 1) == 0;
}
static int dyadic_default(const BatchJob& jb) {
	if (jb.bits == 5) return BATCH_CFG_DY32;
	if (jb.bits == 4) return BATCH_CFG_DY16;
	return jb.W % 1024 == 0 ? BATCH_CFG_DY4 : BATCH_CFG_DY2;
}

// A dyadic job's block size, fixed when the batch is made (its coefficient
// layout depends on it): 16 x 16 or 32 x 32 blocks where they need fewer table
// reads (the plain create's rule; XOR-block jobs fill 512-word tiles).
static void set_job_geometry(BatchJob& jb) {
	jb.bits = dyadic_ok(jb) ? (uint64_t)dyadic_block_bits(jb.R, jb.fr, 5) : 3;
	const uint64_t m = (1ull << jb.bits) - 1;
	jb.pre_in = jb.fi & m;
	jb.pre_out = jb.fr & m;
	jb.nblocks = (jb.pre_in + jb.N + m) / (m + 1);
	jb.Q = (jb.pre_out + jb.R + m) / (m + 1);
}

// The job's coefficient units (8 words each) from unit `off`: its first unit
// (even for 16 x 16 blocks, a multiple of 4 for 32 x 32 blocks, so a block's
// coefficients are one aligned run of threads in k_dyadic_coeffs_batch) and
// the next free unit.
static uint64_t gp_units(const BatchJob& jb, uint64_t off, uint64_t* first) {
	if (jb.bits == 5) {
		*first = (off + 3) & ~3ull;
		return *first + 4 * jb.Q * jb.nblocks;
	}
	if (jb.bits == 4) {
		*first = (off + 1) & ~1ull;
		return *first + 2 * jb.Q * jb.nblocks;
	}
	*first = off;
	return off + jb.Q * jb.nblocks;
}

// The MFMA strategy needs the 64 bit-matrix rows of every coefficient (512 B
// each) beside the job; it is offered only where they are no larger than the
// job's input (64 R <= words per slice).
static bool mfma_eligible(const BatchJob& jb) { return 64 * jb.R <= jb.W; }

static MacArgs mfma_args(const BatchJob& jb, uint64_t* out, uint64_t out_stride) {
	MacArgs m{};
	m.in_base = jb.in;
	m.in_stride = jb.in_stride;
	m.out_base = out;
	m.out_stride = out_stride;
	m.coeff = jb.coeff;
	m.coeff_ld = jb.N;
	m.N = jb.N;
	m.R = jb.R;
	m.W = jb.W;
	m.accumulate = 0;
	return m;
}

static std::vector<int> candidates(uint64_t R, uint64_t W) {
	std::vector<int> c;
	auto add = [&](int id) {
		if (kCfg[id].wpt == 4 && W <= 512) return;  // a 1024-word tile would idle half the threads
		c.push_back(id);
	};
	if (R <= 4) add(BATCH_CFG_4_4_4);
	add(BATCH_CFG_8_4_4);
	add(BATCH_CFG_8_2_4);
	if (R > 8) add(BATCH_CFG_16_2_2);
	if (c.empty()) c.push_back(BATCH_CFG_8_2_4);
	return c;
}

// ---------------------------------------------------------------- tile planning
static const uint64_t kTileWork = 1ull << 22;  // word x output MACs per tile

// XOR-block tiles: a block is 8 inputs x 8 outputs per word; a tile is
// about twice a table tile's MACs (the method takes ~half the time per MAC).
// A 16 x 16 block has 3x the products of an 8 x 8 one: as many products per
// tile.  A 32 x 32 block has 9x: as many products per thread as an 8 x 8
// tile of 512 words, so twice per work-group (one per CU instead of two).
static uint32_t plan_job_dyadic(uint32_t jid, const BatchJob& jb, int cfg,
                                std::vector<BatchTile>& out, bool* atomic) {
	const uint64_t tw = tile_words(cfg);
	const uint64_t wt = jb.W / tw;
	uint64_t chunk = std::max<uint64_t>(2, (2 * kTileWork) / (tw * 64));
	if (cfg == BATCH_CFG_DY16) chunk = std::max<uint64_t>(2, (2 * kTileWork) / (tw * 64) * 27 / 81);
	if (cfg == BATCH_CFG_DY32) chunk = std::max<uint64_t>(2, (4 * kTileWork) / (tw * 64) * 27 / 243);
	if (chunk >= jb.nblocks) chunk = jb.nblocks;
	const uint64_t nch = (jb.nblocks + chunk - 1) / chunk;
	*atomic = nch > 1;
	uint32_t n = 0;
	for (uint64_t k = 0; k < nch; k++)
		for (uint64_t q = 0; q < jb.Q; q++)
			for (uint64_t w = 0; w < wt; w++) {
				BatchTile t;
				t.job = jid;
				t.cfg = (uint16_t)cfg;
				t.mode = *atomic ? 2 : 0;
				t.r0 = q;
				t.w0 = w * tw;
				t.i0 = k * chunk;
				t.i1 = std::min<uint64_t>(jb.nblocks, t.i0 + chunk);
				out.push_back(t);
				n++;
			}
	return n;
}

static uint32_t plan_job(uint32_t jid, const BatchJob& jb, int cfg, std::vector<BatchTile>& out,
                         bool* atomic) {
	if (is_dyadic(cfg)) return plan_job_dyadic(jid, jb, cfg, out, atomic);
	const CfgInfo& c = kCfg[cfg];
	const uint64_t tw = 256ull * c.wpt;
	const uint64_t wt = (jb.W + tw - 1) / tw;
	const uint64_t rtn = (jb.R + c.rt - 1) / c.rt;
	uint64_t chunk = kTileWork / (tw * c.rt);
	chunk = std::max<uint64_t>(chunk, 2ull * c.ib);
	chunk = (chunk + c.ib - 1) / c.ib * c.ib;
	if (chunk >= jb.N) chunk = jb.N;
	const uint64_t nch = (jb.N + chunk - 1) / chunk;
	*atomic = nch > 1;
	uint32_t n = 0;
	for (uint64_t k = 0; k < nch; k++)
		for (uint64_t rt = 0; rt < rtn; rt++)
			for (uint64_t w = 0; w < wt; w++) {
				BatchTile t;
				t.job = jid;
				t.cfg = (uint16_t)cfg;
				t.mode = *atomic ? 2 : 0;
				t.r0 = rt * c.rt;
				t.w0 = w * tw;
				t.i0 = k * chunk;
				t.i1 = std::min<uint64_t>(jb.N, t.i0 + chunk);
				out.push_back(t);
				n++;
			}
	return n;
}

// ---------------------------------------------------------------- autotuner
// Per-shape choice of the launch variant, measured on the device with the
// job's own buffers and cached by (words per slice, R, input-count bucket).
struct TuneKey {
	uint64_t W, R, nb;
	int dy;  // the job qualifies for XOR-block tiles
	bool operator<(const TuneKey& o) const {
		return std::tie(W, R, nb, dy) < std::tie(o.W, o.R, o.nb, o.dy);
	}
};
struct TuneVal {
	int cfg;
	double ns_per_mac;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneVal> g_tune;

static uint64_t n_bucket(uint64_t N) {
	uint64_t b = 1;
	while (b < N && b < (1ull << 20)) b <<= 1;
	return b;
}

static int run_single(const BatchJob& jb, int cfg, BatchJob* djob, BatchTile* dtiles,
                      uint32_t* dcounter, uint32_t* dzero, int blocks, hipStream_t s,
                      float* ms_out) {
	std::vector<BatchTile> tiles;
	bool atomic = false;
	BatchJob j = jb;
	plan_job(0, j, cfg, tiles, &atomic);
	hipMemcpyAsync(djob, &j, sizeof(BatchJob), hipMemcpyHostToDevice, s);
	hipMemcpyAsync(dtiles, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice, s);
	uint32_t zid = 0;
	hipMemcpyAsync(dzero, &zid, sizeof(uint32_t), hipMemcpyHostToDevice, s);
	hipEvent_t a, b;
	hipEventCreate(&a);
	hipEventCreate(&b);
	float best = 1e30f;
	uint32_t fam[kBatchFamilies] = {};
	fam[batch_family(cfg)] = (uint32_t)tiles.size();
	for (int rep = 0; rep < 3; rep++) {
		hipEventRecord(a, s);
		hipError_t e = launch_batch(djob, 1, dtiles, fam, dcounter, dzero, atomic ? 1 : 0, 0, 0, blocks, s);
		hipEventRecord(b, s);
		if (e != hipSuccess) {
			(void)hipStreamSynchronize(s);  // the tile copy above reads `tiles`
			hipEventDestroy(a);
			hipEventDestroy(b);
			return hip_fail(e, "autotune launch");
		}
		hipEventSynchronize(b);
		float ms = 0;
		hipEventElapsedTime(&ms, a, b);
		if (rep && ms < best) best = ms;  // first run warms up
	}
	hipEventDestroy(a);
	hipEventDestroy(b);
	*ms_out = best;
	return PARHIP_OK;
}

// Tune (or look up) the variant for a job; `jb.coeff` must hold its matrix
// (table jobs, MFMA candidates) and `jb.gp` its XOR-block coefficients
// (dyadic jobs).
static int autotune_job(const BatchJob& jb, int blocks, hipStream_t s, int* cfg_out) {
	const bool dy = dyadic_ok(jb);
	TuneKey key{jb.W, jb.R, n_bucket(jb.N), dy ? (int)jb.bits - 2 : 0};
	{
		std::lock_guard<std::mutex> lk(g_tune_mu);
		auto it = g_tune.find(key);
		if (it != g_tune.end()) {
			*cfg_out = it->second.cfg;
			return PARHIP_OK;
		}
	}
	std::vector<int> cand = candidates(jb.R, jb.W);
	if (dy) {
		// the XOR-block method does ~half the LDS work of the direct tables
		// per MAC at every R measured (4-64); only its tile width is tuned
		// (16 x 16 and 32 x 32 blocks: one width)
		cand.clear();
		if (jb.bits == 5) {
			cand.push_back(BATCH_CFG_DY32);
		} else if (jb.bits == 4) {
			cand.push_back(BATCH_CFG_DY16);
		} else {
			if (jb.W % 1024 == 0) cand.push_back(BATCH_CFG_DY4);
			cand.push_back(BATCH_CFG_DY2);
		}
	}
	const bool try_mfma = mfma_eligible(jb);
	if (cand.size() == 1 && !try_mfma) {
		*cfg_out = cand[0];
		std::lock_guard<std::mutex> lk(g_tune_mu);
		g_tune[key] = {cand[0], 0.0};
		return PARHIP_OK;
	}
	// scratch output so the caller's recovery buffer is not disturbed
	DevBuf dout, djob, dtiles, dcnt, dzero;
	BatchJob j = jb;
	int rc;
	if ((rc = dout.alloc(jb.R * jb.W * sizeof(uint64_t), s)) ||
	    (rc = djob.alloc(sizeof(BatchJob), s)) || (rc = dcnt.alloc(kBatchFamilies * sizeof(uint32_t), s)) ||
	    (rc = dzero.alloc(sizeof(uint32_t), s)))
		return rc;
	j.out = dout.as<uint64_t>();
	j.out_stride = jb.W;
	size_t max_tiles = 0;
	for (int c : cand) {
		std::vector<BatchTile> t;
		bool at;
		max_tiles = std::max<size_t>(max_tiles, plan_job(0, j, c, t, &at));
	}
	if ((rc = dtiles.alloc(max_tiles * sizeof(BatchTile), s))) return rc;
	int best = cand[0];
	double best_ns = 1e300;
	const double macs = (double)jb.N * jb.W * jb.R;
	for (int c : cand) {
		float ms = 0;
		rc = run_single(j, c, djob.as<BatchJob>(), dtiles.as<BatchTile>(), dcnt.as<uint32_t>(),
		                dzero.as<uint32_t>(), blocks, s, &ms);
		if (rc) return rc;
		const double ns = ms * 1e6 / macs;
		if (ns < best_ns) {
			best_ns = ns;
			best = c;
		}
	}
	MfmaPlan mp;
	if (try_mfma && plan_mfma(jb.N, jb.R, jb.W, true, false, &mp) == PARHIP_OK) {
		// Strategy choice on equal terms: the job alone as full-device launches
		// of each strategy (table kernel with its split-K plan, and rows + MFMA
		// kernel + fold), so neither is judged by how well one small job fills
		// a persistent grid.
		DevBuf mws;
		if ((rc = mws.alloc(mp.rows_bytes + mp.part_bytes, s))) return rc;
		const MacArgs m = mfma_args(jb, dout.as<uint64_t>(), jb.W);
		DyadicPlan dp{};
		DevBuf dws;
		if (dy) {
			dp = plan_dyadic(jb.N, jb.R, jb.W, jb.fi, jb.fr, current_device(), SIZE_MAX, true, (int)jb.bits,
			                 true);
			if ((rc = dws.alloc(dp.part_bytes, s))) return rc;
		}
		float ms_tab = 1e30f, ms_mfma = 1e30f;
		hipEvent_t a, b;
		hipEventCreate(&a);
		hipEventCreate(&b);
		const bool timing = prof_set(false);  // keep the caller's kernel timings clean
		hipError_t e = hipSuccess;
		for (int which = 0; which < 2 && !rc && e == hipSuccess; which++)
			for (int rep = 0; rep < 3 && !rc && e == hipSuccess; rep++) {
				hipEventRecord(a, s);
				if (which == 0 && dy)
					e = enqueue_dyadic(m, dp, jb.gp, jb.nblocks, dws.p, s);
				else if (which == 0)
					rc = run_contraction_tables(m, s);
				else
					e = enqueue_mfma(m, mp, mws.p, s);
				hipEventRecord(b, s);
				if (!rc && e == hipSuccess) e = hipEventSynchronize(b);
				float ms = 0;
				if (!rc && e == hipSuccess) hipEventElapsedTime(&ms, a, b);
				float& best_ms = which ? ms_mfma : ms_tab;
				if (rep && ms < best_ms) best_ms = ms;  // first run warms up
			}
		prof_set(timing);
		hipEventDestroy(a);
		hipEventDestroy(b);
		if (rc) return rc;
		if (e != hipSuccess) return hip_fail(e, "autotune mfma launch");
		if (ms_mfma < ms_tab) {
			best_ns = ms_mfma * 1e6 / macs;
			best = BATCH_CFG_MFMA;
		}
	}
	hipStreamSynchronize(s);
	std::lock_guard<std::mutex> lk(g_tune_mu);
	g_tune[key] = {best, best_ns};
	*cfg_out = best;
	return PARHIP_OK;
}

}  // namespace parhip

using namespace parhip;

// ====================================================================== batch object
struct parhip_batch {
	int dev = 0;
	std::vector<BatchJob> jobs;  // the non-empty jobs, in order
	std::vector<int> cfg;
	std::vector<int64_t> slot;   // caller's job index -> index in jobs, -1 if empty
	BatchJob* djobs = nullptr;
	BatchTile* dtiles = nullptr;
	uint32_t ntiles = 0;
	uint32_t fam_tiles[kBatchFamilies] = {};  // tiles of each family, grouped in dtiles
	uint32_t* dcounter = nullptr;                      // one queue counter per family
	uint32_t* dzero = nullptr;
	uint32_t nzero = 0;
	uint64_t* dcoeff = nullptr;
	uint64_t coeff_total = 0;   // full matrices built per run (table / MFMA jobs)
	uint64_t* dgp = nullptr;
	uint64_t dy_pairs = 0;      // XOR-block coefficient pairs built per run (dyadic jobs)
	int blocks = 0;
	hipStream_t cap = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	uint64_t tile_counts[BATCH_CFG_COUNT] = {0};
	struct MfmaJob {
		size_t job;        // index in jobs
		MfmaPlan plan;
		void* ws;          // bit-matrix rows + split-K partials, owned
	};
	std::vector<MfmaJob> mfma;
};

// The batch's launch sequence: every job's coefficients, zeroing of split
// jobs, queue reset and persistent kernel, then the MFMA jobs' launches.
static hipError_t batch_enqueue(const parhip_batch* b, hipStream_t s) {
	hipError_t e = launch_batch(b->djobs, (uint32_t)b->jobs.size(), b->dtiles, b->fam_tiles, b->dcounter,
	                            b->dzero, b->nzero, b->coeff_total, b->dy_pairs, b->blocks, s);
	for (size_t k = 0; k < b->mfma.size() && e == hipSuccess; k++) {
		const auto& mj = b->mfma[k];
		const BatchJob& jb = b->jobs[mj.job];
		e = enqueue_mfma(mfma_args(jb, jb.out, jb.out_stride), mj.plan, mj.ws, s);
	}
	return e;
}

static void batch_free(parhip_batch* b) {
	if (!b) return;
	hipSetDevice(b->dev);
	if (b->cap) hipStreamSynchronize(b->cap);
	if (b->exec) hipGraphExecDestroy(b->exec);
	if (b->graph) hipGraphDestroy(b->graph);
	if (b->djobs) hipFree(b->djobs);
	if (b->dtiles) hipFree(b->dtiles);
	if (b->dcounter) hipFree(b->dcounter);
	if (b->dzero) hipFree(b->dzero);
	if (b->dcoeff) hipFree(b->dcoeff);
	if (b->dgp) hipFree(b->dgp);
	for (auto& mj : b->mfma)
		if (mj.ws) hipFree(mj.ws);
	if (b->cap) hipStreamDestroy(b->cap);
	delete b;
}

extern "C" {

parhip_batch* parhip_batch_create(const parhip_gf64_job* jobs, size_t njobs, unsigned flags) {
	if (!jobs || njobs == 0 || njobs > (1u << 30)) {
		fail(PARHIP_ERR_ARG, "batch_create: no jobs");
		return nullptr;
	}
	auto* b = new (std::nothrow) parhip_batch;
	if (!b) return nullptr;
	b->dev = current_device();
	hipSetDevice(b->dev);
	const int cus = device_cus(b->dev);
	b->blocks = cus * 2;
	hipError_t e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
	if (e != hipSuccess) {
		hip_fail(e, "batch stream");
		batch_free(b);
		return nullptr;
	}
	// ---- job descriptors + coefficient storage
	b->slot.assign(njobs, -1);
	for (size_t j = 0; j < njobs; j++) {
		const parhip_gf64_job& in = jobs[j];
		if (in.numInputs == 0 || in.numRecovery == 0 || in.blockSize64 == 0) continue;
		if (!in.d_inputs || !in.d_recovery || in.inputStride64 < in.blockSize64 ||
		    in.recoveryStride64 < in.blockSize64) {
			fail(PARHIP_ERR_ARG, "batch_create: bad job %zu", j);
			batch_free(b);
			return nullptr;
		}
		BatchJob jb{};
		jb.in = in.d_inputs;
		jb.in_stride = in.inputStride64;
		jb.out = in.d_recovery;
		jb.out_stride = in.recoveryStride64;
		jb.N = in.numInputs;
		jb.R = in.numRecovery;
		jb.W = in.blockSize64;
		jb.fi = in.firstInput;
		jb.fr = in.firstRecovery;
		set_job_geometry(jb);
		// storage for every form the job may run in: XOR-block coefficients
		// where it qualifies, the full matrix where tables or MFMA may run it
		const bool dy = dyadic_ok(jb);
		const bool full = !dy || ((flags & (PARHIP_BATCH_MFMA | PARHIP_BATCH_AUTOTUNE)) && mfma_eligible(jb));
		jb.coeff_off = b->coeff_total;
		if (full) b->coeff_total += jb.N * jb.R;
		jb.gp_off = b->dy_pairs;
		if (dy) b->dy_pairs = gp_units(jb, b->dy_pairs, &jb.gp_off);
		b->slot[j] = (int64_t)b->jobs.size();
		b->jobs.push_back(jb);
	}
	if (b->jobs.empty()) return b;  // nothing to do: run() is a no-op
	e = hipSuccess;
	if (b->coeff_total) e = hipMalloc((void**)&b->dcoeff, b->coeff_total * sizeof(uint64_t));
	if (e == hipSuccess && b->dy_pairs) e = hipMalloc((void**)&b->dgp, b->dy_pairs * 8 * sizeof(uint64_t));
	if (e == hipSuccess) e = hipMalloc((void**)&b->djobs, b->jobs.size() * sizeof(BatchJob));
	if (e == hipSuccess) e = hipMalloc((void**)&b->dcounter, kBatchFamilies * sizeof(uint32_t));
	if (e != hipSuccess) {
		hip_fail(e, "batch allocation");
		batch_free(b);
		return nullptr;
	}
	for (auto& jb : b->jobs) {
		jb.coeff = b->dcoeff ? b->dcoeff + jb.coeff_off : nullptr;
		jb.gp = b->dgp ? b->dgp + jb.gp_off * 8 : nullptr;
	}
	e = hipMemcpy(b->djobs, b->jobs.data(), b->jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice);
	// coefficients first: the autotuner runs the real kernels on them
	if (e == hipSuccess)
		e = launch_batch(b->djobs, (uint32_t)b->jobs.size(), nullptr, nullptr, b->dcounter, nullptr, 0,
		                 b->coeff_total, b->dy_pairs, 0, b->cap);
	if (e == hipSuccess) e = hipStreamSynchronize(b->cap);
	if (e != hipSuccess) {
		hip_fail(e, "batch coefficients");
		batch_free(b);
		return nullptr;
	}
	// ---- per-job launch variant
	b->cfg.resize(b->jobs.size());
	for (size_t j = 0; j < b->jobs.size(); j++) {
		const BatchJob& jb = b->jobs[j];
		int cfg = jb.R <= 4 ? BATCH_CFG_4_4_4 : (jb.W <= 512 ? BATCH_CFG_8_2_4 : BATCH_CFG_8_4_4);
		if (dyadic_ok(jb)) cfg = dyadic_default(jb);
		if ((flags & PARHIP_BATCH_MFMA) && mfma_eligible(jb)) {
			cfg = BATCH_CFG_MFMA;
		} else if (flags & PARHIP_BATCH_AUTOTUNE) {
			if (autotune_job(jb, b->blocks, b->cap, &cfg)) {
				batch_free(b);
				return nullptr;
			}
		}
		if (cfg == BATCH_CFG_MFMA) {
			parhip_batch::MfmaJob mj{j, MfmaPlan{}, nullptr};
			if (plan_mfma(jb.N, jb.R, jb.W, true, false, &mj.plan) != PARHIP_OK) {
				batch_free(b);
				return nullptr;
			}
			e = hipMalloc(&mj.ws, mj.plan.rows_bytes + mj.plan.part_bytes);
			if (e != hipSuccess) {
				hip_fail(e, "batch mfma workspace");
				batch_free(b);
				return nullptr;
			}
			b->mfma.push_back(mj);
		}
		b->cfg[j] = cfg;
	}
	// ---- per-run coefficient layout: full matrices only for the table and
	// MFMA jobs, XOR-block coefficients only for the dyadic jobs
	b->coeff_total = 0;
	b->dy_pairs = 0;
	for (size_t j = 0; j < b->jobs.size(); j++) {
		BatchJob& jb = b->jobs[j];
		jb.coeff_off = b->coeff_total;
		jb.gp_off = b->dy_pairs;
		if (is_dyadic(b->cfg[j])) {
			b->dy_pairs = gp_units(jb, b->dy_pairs, &jb.gp_off);
		} else {
			jb.coeff = b->dcoeff + jb.coeff_off;
			b->coeff_total += jb.N * jb.R;
		}
		if (jb.gp) jb.gp = b->dgp + jb.gp_off * 8;
	}
	e = hipMemcpy(b->djobs, b->jobs.data(), b->jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice);
	if (e != hipSuccess) {
		hip_fail(e, "batch job upload");
		batch_free(b);
		return nullptr;
	}
	// ---- tiles: interleave jobs so concurrent work-groups mix shapes
	std::vector<std::vector<BatchTile>> per(b->jobs.size());
	std::vector<uint32_t> zero_ids;
	for (size_t j = 0; j < b->jobs.size(); j++) {
		if (b->cfg[j] == BATCH_CFG_MFMA) continue;  // own launches, no tiles
		bool atomic = false;
		plan_job((uint32_t)j, b->jobs[j], b->cfg[j], per[j], &atomic);
		b->tile_counts[b->cfg[j]] += per[j].size();
		if (atomic) zero_ids.push_back((uint32_t)j);
	}
	std::vector<BatchTile> tiles;
	size_t total = 0;
	for (auto& v : per) total += v.size();
	if (total > 0xFFFFFFF0ull) {
		fail(PARHIP_ERR_RANGE, "batch_create: too many tiles");
		batch_free(b);
		return nullptr;
	}
	tiles.reserve(total);
	// grouped by family (each its own persistent kernel), jobs interleaved
	// inside a family
	for (int f = 0; f < kBatchFamilies; f++) {
		size_t fam_total = 0;
		for (size_t j = 0; j < per.size(); j++)
			if (b->cfg[j] != BATCH_CFG_MFMA && batch_family(b->cfg[j]) == f) fam_total += per[j].size();
		std::vector<size_t> pos(per.size(), 0);
		const size_t start = tiles.size();
		while (tiles.size() - start < fam_total)
			for (size_t j = 0; j < per.size(); j++)
				if (batch_family(b->cfg[j]) == f && pos[j] < per[j].size()) tiles.push_back(per[j][pos[j]++]);
		b->fam_tiles[f] = (uint32_t)fam_total;
	}
	b->ntiles = (uint32_t)tiles.size();
	b->nzero = (uint32_t)zero_ids.size();
	if (!tiles.empty()) e = hipMalloc((void**)&b->dtiles, tiles.size() * sizeof(BatchTile));
	if (e == hipSuccess && b->nzero) e = hipMalloc((void**)&b->dzero, b->nzero * sizeof(uint32_t));
	if (e != hipSuccess) {
		hip_fail(e, "batch allocation");
		batch_free(b);
		return nullptr;
	}
	if (!tiles.empty())
		e = hipMemcpy(b->dtiles, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice);
	if (e == hipSuccess && b->nzero)
		e = hipMemcpy(b->dzero, zero_ids.data(), zero_ids.size() * sizeof(uint32_t),
		              hipMemcpyHostToDevice);
	if (e != hipSuccess) {
		hip_fail(e, "batch tile upload");
		batch_free(b);
		return nullptr;
	}
	// ---- graph: Cauchy build, zeroing, queue reset, persistent kernel
	if (flags & PARHIP_BATCH_GRAPH) {
		const bool timing = prof_set(false);  // no event nodes in the graph
		e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
		if (e == hipSuccess) e = batch_enqueue(b, b->cap);
		hipError_t e2 = hipStreamEndCapture(b->cap, &b->graph);
		prof_set(timing);
		if (e == hipSuccess) e = e2;
		if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
		if (e != hipSuccess) {
			hip_fail(e, "batch graph capture");
			batch_free(b);
			return nullptr;
		}
	}
	return b;
}

int parhip_batch_run(parhip_batch* b, void* stream) {
	if (!b) return fail(PARHIP_ERR_ARG, "batch_run: null batch");
	if (b->jobs.empty()) return PARHIP_OK;
	hipStream_t s = (hipStream_t)stream;
	hipError_t e;
	ProfScope ps("batch", s);
	if (b->exec) {
		e = hipGraphLaunch(b->exec, s);
	} else {
		e = batch_enqueue(b, s);
	}
	return e == hipSuccess ? PARHIP_OK : hip_fail(e, "batch_run");
}

int parhip_batch_info(const parhip_batch* b, uint64_t* ntiles, uint64_t* natomic_jobs,
                      uint64_t* coeff_bytes) {
	if (!b) return fail(PARHIP_ERR_ARG, "batch_info: null batch");
	if (ntiles) *ntiles = b->ntiles;
	if (natomic_jobs) *natomic_jobs = b->nzero;
	if (coeff_bytes) *coeff_bytes = (b->coeff_total + b->dy_pairs * 8) * sizeof(uint64_t);
	return PARHIP_OK;
}

int parhip_batch_job_config(const parhip_batch* b, size_t job, int* rt, int* wpt, int* ib) {
	if (!b || job >= b->slot.size()) return fail(PARHIP_ERR_ARG, "batch_job_config: bad job");
	if (b->slot[job] < 0) return fail(PARHIP_ERR_ARG, "batch_job_config: job %zu is empty", job);
	const CfgInfo& c = kCfg[b->cfg[(size_t)b->slot[job]]];
	if (rt) *rt = c.rt;
	if (wpt) *wpt = c.wpt;
	if (ib) *ib = c.ib;
	return PARHIP_OK;
}

int parhip_batch_job_method(const parhip_batch* b, size_t job) {
	if (!b || job >= b->slot.size()) return fail(PARHIP_ERR_ARG, "batch_job_method: bad job");
	if (b->slot[job] < 0) return fail(PARHIP_ERR_ARG, "batch_job_method: job %zu is empty", job);
	const int cfg = b->cfg[(size_t)b->slot[job]];
	return cfg == BATCH_CFG_MFMA ? PARHIP_METHOD_MFMA
	                             : (is_dyadic(cfg) ? PARHIP_METHOD_DYADIC : PARHIP_METHOD_TABLE);
}

void parhip_batch_destroy(parhip_batch* b) { batch_free(b); }

// The entry of a shape: the XOR-block one where the shape qualifies for it
// (aligned slices of a multiple of 512 words, R >= 4), else the other.
static std::map<TuneKey, TuneVal>::iterator tune_find(size_t N, size_t R, size_t W) {
	auto it = g_tune.find(TuneKey{W, R, n_bucket(N), 3});
	if (it == g_tune.end()) it = g_tune.find(TuneKey{W, R, n_bucket(N), 2});
	if (it == g_tune.end()) it = g_tune.find(TuneKey{W, R, n_bucket(N), 1});
	if (it == g_tune.end()) it = g_tune.find(TuneKey{W, R, n_bucket(N), 0});
	return it;
}

int parhip_autotune_lookup(size_t numInputs, size_t numRecovery, size_t blockSize64, int* rt,
                           int* wpt, int* ib, double* ns_per_mac) {
	std::lock_guard<std::mutex> lk(g_tune_mu);
	auto it = tune_find(numInputs, numRecovery, blockSize64);
	if (it == g_tune.end()) return fail(PARHIP_ERR_ARG, "autotune: shape not tuned");
	const CfgInfo& c = kCfg[it->second.cfg];
	if (rt) *rt = c.rt;
	if (wpt) *wpt = c.wpt;
	if (ib) *ib = c.ib;
	if (ns_per_mac) *ns_per_mac = it->second.ns_per_mac;
	return PARHIP_OK;
}

int parhip_autotune_lookup_method(size_t numInputs, size_t numRecovery, size_t blockSize64) {
	std::lock_guard<std::mutex> lk(g_tune_mu);
	auto it = tune_find(numInputs, numRecovery, blockSize64);
	if (it == g_tune.end()) return fail(PARHIP_ERR_ARG, "autotune: shape not tuned");
	const int cfg = it->second.cfg;
	return cfg == BATCH_CFG_MFMA ? PARHIP_METHOD_MFMA
	                             : (is_dyadic(cfg) ? PARHIP_METHOD_DYADIC : PARHIP_METHOD_TABLE);
}

int parhip_autotune_save(const char* path) {
	if (!path) return fail(PARHIP_ERR_ARG, "autotune_save: null path");
	FILE* f = fopen(path, "w");
	if (!f) return fail(PARHIP_ERR_ARG, "autotune_save: cannot open %s", path);
	std::lock_guard<std::mutex> lk(g_tune_mu);
	fprintf(f, "# words_per_slice recovery input_bucket xor_block cfg ns_per_mac\n");
	for (auto& kv : g_tune)
		fprintf(f, "%llu %llu %llu %d %d %.6g\n", (unsigned long long)kv.first.W,
		        (unsigned long long)kv.first.R, (unsigned long long)kv.first.nb, kv.first.dy,
		        kv.second.cfg, kv.second.ns_per_mac);
	fclose(f);
	return PARHIP_OK;
}

int parhip_autotune_load(const char* path) {
	if (!path) return fail(PARHIP_ERR_ARG, "autotune_load: null path");
	FILE* f = fopen(path, "r");
	if (!f) return fail(PARHIP_ERR_ARG, "autotune_load: cannot open %s", path);
	char line[256];
	std::lock_guard<std::mutex> lk(g_tune_mu);
	while (fgets(line, sizeof line, f)) {
		if (line[0] == '#') continue;
		unsigned long long W, R, nb;
		int dy, cfg;
		double ns;
		// an XOR-block variant only under a key that qualifies for it
		if (sscanf(line, "%llu %llu %llu %d %d %lg", &W, &R, &nb, &dy, &cfg, &ns) == 6 && cfg >= 0 &&
		    cfg < BATCH_CFG_COUNT && kCfg[cfg].rt && dy >= 0 && dy <= 3 &&
		    (is_dyadic(cfg) ? dy == (cfg == BATCH_CFG_DY32 ? 3 : cfg == BATCH_CFG_DY16 ? 2 : 1) : true))
			g_tune[TuneKey{W, R, nb, dy}] = {cfg, ns};
	}
	fclose(f);
	return PARHIP_OK;
}

int parhip_autotune_clear(void) {
	std::lock_guard<std::mutex> lk(g_tune_mu);
	g_tune.clear();
	return PARHIP_OK;
}

}  // extern "C"
