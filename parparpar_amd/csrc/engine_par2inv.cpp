// engine_par2inv.cpp -- C ABI of the PAR2 repair matrix
// (Galois16RecMatrix::Compute, gf16/gfmat_inv.cpp:562-676; gf16_inv.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/parhip.h"
#include "../../include/parhip_par2.h"
#include "engine_internal.h"
#include "kernels.h"

using namespace parhip;

// Compute's argument contract (asserts at gfmat_inv.cpp:563-567).
static int rec_matrix_args(const uint8_t* valid, uint32_t n, uint32_t validCount,
                           const uint16_t* rec, const uint32_t* cnt) {
	if (!valid || !rec || !cnt) return fail(PARHIP_ERR_ARG, "gf16_rec_matrix: null pointer");
	if (n == 0 || n > 32768) return fail(PARHIP_ERR_RANGE, "gf16_rec_matrix: numInputs must be 1..32768");
	if (validCount >= n)
		return fail(PARHIP_ERR_ARG, "gf16_rec_matrix: nothing to recover (validCount >= numInputs)");
	if (*cnt == 0 || *cnt > 65535)
		return fail(PARHIP_ERR_RANGE, "gf16_rec_matrix: recovery count must be 1..65535");
	uint32_t v = 0;
	for (uint32_t i = 0; i < n; i++) v += valid[i] ? 1u : 0u;
	if (v != validCount) return fail(PARHIP_ERR_ARG, "gf16_rec_matrix: validCount does not match inputValid");
	return PARHIP_OK;
}

extern "C" {

int parhip_d_gf16_rec_matrix_compute(const uint8_t* inputValid, uint32_t numInputs,
                                     uint32_t validCount, uint16_t* recovery,
                                     uint32_t* recoveryCount, uint16_t* d_factors, void* stream) {
	int rc = rec_matrix_args(inputValid, numInputs, validCount, recovery, recoveryCount);
	if (rc) return rc;
	if (!d_factors) return fail(PARHIP_ERR_ARG, "gf16_rec_matrix: null factor matrix");
	hipStream_t s = (hipStream_t)stream;
	const uint32_t numRec = numInputs - validCount;
	uint32_t cnt = *recoveryCount;
	std::sort(recovery, recovery + cnt);  // gfmat_inv.cpp:594
	std::vector<uint16_t> vidx, midx;
	for (uint32_t i = 0; i < numInputs; i++) (inputValid[i] ? vidx : midx).push_back((uint16_t)i);
	const uint16_t *et, *lt;
	if ((rc = gf16_device_log_exp(&et, &lt))) return rc;
	const size_t nn = (size_t)numRec * numRec;
	DevBuf dA, dB, dIdx, dExp, dBad;
	if ((rc = dA.alloc(nn * 2, s)) || (rc = dB.alloc(nn * 2, s)) ||
	    (rc = dIdx.alloc((size_t)numInputs * 2, s)) || (rc = dExp.alloc((size_t)numRec * 2, s)) ||
	    (rc = dBad.alloc(sizeof(int), s)))
		return rc;
	uint16_t* d_vidx = dIdx.as<uint16_t>();
	uint16_t* d_midx = d_vidx + vidx.size();
	hipError_t e = hipMemcpyAsync(d_vidx, vidx.data(), vidx.size() * 2, hipMemcpyHostToDevice, s);
	if (e == hipSuccess)
		e = hipMemcpyAsync(d_midx, midx.data(), midx.size() * 2, hipMemcpyHostToDevice, s);
	// vidx / midx are this frame's: let the copies finish before any return
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
	uint16_t* inv = nullptr;
	for (;;) {  // invert_loop: a zero pivot drops that exponent and retries
		if (numRec > cnt) {
			*recoveryCount = cnt;
			return 0;
		}
		e = hipMemcpyAsync(dExp.p, recovery, (size_t)numRec * 2, hipMemcpyHostToDevice, s);
		if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
		// A[j][m] = coeff(missing_m, recovery_j)
		if ((rc = gf16_coeffs_device(dA.as<uint16_t>(), d_midx, numRec, dExp.as<uint16_t>(), numRec, s)))
			return rc;
		{
			ProfScope ps("gf16_invert", s);
			e = launch_gf16_invert(dA.as<uint16_t>(), dB.as<uint16_t>(), numRec, et, lt,
			                       dBad.as<int>(), &inv, s);
		}
		if (e != hipSuccess) return hip_fail(e, "k_gf16_gj_step");
		int bad = -1;
		e = hipMemcpyAsync(&bad, dBad.p, sizeof(int), hipMemcpyDeviceToHost, s);
		if (e == hipSuccess) e = hipStreamSynchronize(s);
		if (e != hipSuccess) return hip_fail(e, "gf16_rec_matrix");
		if (bad < 0) break;
		std::copy(recovery + bad + 1, recovery + cnt, recovery + bad);  // recovery.erase
		cnt--;
	}
	// missing columns: A^-1
	e = hipMemcpy2DAsync(d_factors + validCount, (size_t)numInputs * 2, inv, (size_t)numRec * 2,
	                     (size_t)numRec * 2, numRec, hipMemcpyDeviceToDevice, s);
	if (e != hipSuccess) return hip_fail(e, "hipMemcpy2D");
	if (validCount) {  // valid columns: A^-1 x C_valid, C_valid[j][v] = coeff(valid_v, recovery_j)
		DevBuf dC;
		if ((rc = dC.alloc((size_t)numRec * validCount * 2, s))) return rc;
		if ((rc = gf16_coeffs_device(dC.as<uint16_t>(), d_vidx, validCount, dExp.as<uint16_t>(), numRec, s)))
			return rc;
		Gf16Args a{};
		a.in_base = dC.as<uint16_t>();
		a.in_stride = validCount;
		a.out_base = d_factors;
		a.out_stride = numInputs;
		a.coeff = inv;
		a.coeff_ld = numRec;
		a.N = numRec;
		a.R = numRec;
		a.W = validCount;
		a.accumulate = 0;
		if ((rc = run_gf16_contraction(a, s))) return rc;
	}
	*recoveryCount = numRec;  // recovery.resize(numRec)
	return 1;
}

int parhip_gf16_rec_matrix_compute(const uint8_t* inputValid, uint32_t numInputs,
                                   uint32_t validCount, uint16_t* recovery,
                                   uint32_t* recoveryCount, uint16_t* factors) {
	int rc = rec_matrix_args(inputValid, numInputs, validCount, recovery, recoveryCount);
	if (rc) return rc;
	if (!factors) return fail(PARHIP_ERR_ARG, "gf16_rec_matrix: null factor matrix");
	hipStream_t s;
	if ((rc = host_stream(&s))) return rc;
	const uint32_t numRec = numInputs - validCount;
	DevBuf dF;
	if ((rc = dF.alloc((size_t)numRec * numInputs * 2, s))) return rc;
	rc = parhip_d_gf16_rec_matrix_compute(inputValid, numInputs, validCount, recovery, recoveryCount,
	                                      dF.as<uint16_t>(), s);
	if (rc != 1) {
		hipStreamSynchronize(s);
		return rc;
	}
	hipError_t e = hipMemcpyAsync(factors, dF.p, (size_t)numRec * numInputs * 2,
	                              hipMemcpyDeviceToHost, s);
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	return e == hipSuccess ? 1 : hip_fail(e, "factors D2H");
}

}  // extern "C"
