"""The reference addon's call shapes over the gfx950 engine.

ParParPar's JS front end calls ``parpar_gf64.node`` (src/gf64_addon.cc:3271-3612)
with Buffers; this module offers the same functions with the same argument
order over writable numpy buffers (uint8 or uint64), so tests can replay the
reference's own call sequences.  Every call goes to libparhip.so.

    compute_recovery(inputs, outputs, numInputs, numRecovery, blockSize,
                     firstInput, firstRecovery[, numThreads])       gf64_addon.cc:1266
    compute_recovery_full(...same...)                               gf64_addon.cc:1443
    compute_recovery_accumulate(...same...)  (accumulate=true)      gf64_addon.cc:1799
    compute_recovery_barycentric(...same...[, accumulate])          gf64_addon.cc:1636
    compute_recovery_fenger(...same...)                             gf64_addon.cc:1980
    compute_recovery_streaming(inputBatches, outputBuffer, numInputs, numRecovery,
                     blockSize, firstInput, firstRecovery[, numThreads])
                                                                    gf64_addon.cc:2349
    compute_recovery_with_coeff(coeffMatrix, inputs, outputs, numInputs,
                     numRecovery, blockSize, firstInput, firstRecovery[, numThreads])
                                                                    gf64_addon.cc:2189
    build_coefficient_matrix(numInputs, numRecovery, firstInput, firstRecovery)
                                                                    gf64_addon.cc:2124
    solve_and_reconstruct(A, rhsBlocks, n, blockSize) -> bool       gf64_addon.cc:2655
    gf64_solve(A, b, n) -> x or None                                gf64_addon.cc:1127
    mul_arr(out, inBuf, coeff, len_words, n_coeff)                  gf64_addon.cc:3038
    Gf64Encoder().{mul, mul_arr, coupled_muladd_arr, fused_output_muladd_arr,
                   two_d_muladd_arr}                                gf64_addon.cc:168-1054

`blockSize` is in bytes and must be a multiple of 8, and the counts must be
positive, as the addon requires (gf64_addon.cc:1553-1578); violations raise
ValueError (the addon's RangeError).  The C entries underneath treat empty
problems as no-ops, as GF64Controller does (src/par3_engine.cc:1069).
"""
import ctypes

import numpy as np

from . import check, lib


def _buf(a, name):
    if not isinstance(a, np.ndarray):
        raise TypeError("%s must be a numpy buffer" % name)
    if not a.flags["C_CONTIGUOUS"]:
        raise TypeError("%s must be contiguous" % name)
    return a.ctypes.data


def _check_sizes(buf, count, block, name):
    if buf.nbytes < count * block:
        raise ValueError("%s too small: %d < %d" % (name, buf.nbytes, count * block))


def _check_block(blockSize):
    if blockSize <= 0 or blockSize % 8:
        raise ValueError("blockSize must be a positive multiple of 8")


def _check_counts(numInputs, numRecovery):
    # the addon rejects empty problems (gf64_addon.cc:1383-1391, :1554-1562,
    # :1917-1922, :2153-2156); the engine underneath treats them as no-ops
    if numInputs <= 0:
        raise ValueError("numInputs must be positive")
    if numRecovery <= 0:
        raise ValueError("numRecovery must be positive")


def compute_recovery(inputs, outputs, numInputs, numRecovery, blockSize, firstInput, firstRecovery,
                     numThreads=0, accumulate=False):
    _check_counts(numInputs, numRecovery)
    _check_block(blockSize)
    _check_sizes(inputs, numInputs, blockSize, "inputs")
    _check_sizes(outputs, numRecovery, blockSize, "outputs")
    if outputs.flags["WRITEABLE"] is False:
        raise TypeError("outputs must be writable")
    check(lib().parhip_gf64_compute_recovery_blocks(
        _buf(inputs, "inputs"), numInputs, _buf(outputs, "outputs"), numRecovery, blockSize // 8,
        firstInput, firstRecovery, numThreads, 1 if accumulate else 0))
    return outputs


def compute_recovery_full(inputs, outputs, numInputs, numRecovery, blockSize, firstInput,
                          firstRecovery, numThreads=0, accumulate=False):
    _check_counts(numInputs, numRecovery)
    _check_block(blockSize)
    _check_sizes(inputs, numInputs, blockSize, "inputs")
    _check_sizes(outputs, numRecovery, blockSize, "outputs")
    check(lib().parhip_gf64_compute_recovery_blocks_full(
        _buf(inputs, "inputs"), numInputs, _buf(outputs, "outputs"), numRecovery, blockSize // 8,
        firstInput, firstRecovery, numThreads, 1 if accumulate else 0))
    return outputs


def compute_recovery_accumulate(inputs, outputs, numInputs, numRecovery, blockSize, firstInput,
                                firstRecovery, numThreads=0):
    return compute_recovery_full(inputs, outputs, numInputs, numRecovery, blockSize, firstInput,
                                 firstRecovery, numThreads, accumulate=True)


def compute_recovery_barycentric(inputs, outputs, numInputs, numRecovery, blockSize, firstInput,
                                 firstRecovery, numThreads=0, accumulate=False):
    """Counts must be positive, as the addon requires (gf64_addon.cc:1746-1755)."""
    _check_counts(numInputs, numRecovery)
    _check_block(blockSize)
    _check_sizes(inputs, numInputs, blockSize, "inputs")
    _check_sizes(outputs, numRecovery, blockSize, "outputs")
    check(lib().parhip_gf64_compute_recovery_blocks_barycentric(
        _buf(inputs, "inputs"), numInputs, _buf(outputs, "outputs"), numRecovery, blockSize // 8,
        firstInput, firstRecovery, numThreads, 1 if accumulate else 0))
    return outputs


def compute_recovery_fenger(inputs, outputs, numInputs, numRecovery, blockSize, firstInput,
                            firstRecovery, numThreads=0):
    """Negative counts are a RangeError, zero counts a no-op (gf64_addon.cc:2017-2055,
    src/par3_engine_fenger.cc:118-119)."""
    if numInputs < 0 or numRecovery < 0:
        raise ValueError("numInputs and numRecovery must be non-negative")
    _check_block(blockSize)
    _check_sizes(inputs, numInputs, blockSize, "inputs")
    _check_sizes(outputs, numRecovery, blockSize, "outputs")
    check(lib().parhip_gf64_compute_recovery_blocks_fenger(
        _buf(inputs, "inputs"), numInputs, _buf(outputs, "outputs"), numRecovery, blockSize // 8,
        firstInput, firstRecovery, numThreads))
    return outputs


def compute_recovery_streaming(inputBatches, outputBuffer, numInputs, numRecovery, blockSize,
                               firstInput, firstRecovery, numThreads=0):
    """All input batches in one call; the argument rules of gf64_addon.cc:2349-2525
    (TypeError for a non-list or a non-buffer element, RangeError -> ValueError for
    negative counts, a bad blockSize, a short output, a batch that is not whole
    slices, or batches that do not hold exactly numInputs slices)."""
    if not isinstance(inputBatches, (list, tuple)):
        raise TypeError("inputBatches must be an array of Buffers")
    if numInputs < 0:
        raise ValueError("numInputs must be non-negative")
    if numRecovery < 0:
        raise ValueError("numRecovery must be non-negative")
    _check_block(blockSize)
    if numInputs == 0 or numRecovery == 0:
        return outputBuffer
    _check_sizes(outputBuffer, numRecovery, blockSize, "outputBuffer")
    counts, total = [], 0
    for b in inputBatches:
        if not isinstance(b, np.ndarray):
            raise TypeError("inputBatches element must be a Buffer")
        if b.nbytes % blockSize:
            raise ValueError("inputBatches element length not a multiple of blockSize")
        k = b.nbytes // blockSize
        if total + k > numInputs:
            raise ValueError("inputBatches total blocks exceeds numInputs")
        _buf(b, "inputBatches element")
        counts.append(k)
        total += k
    if total != numInputs:
        raise ValueError("inputBatches total blocks does not equal numInputs")
    n = len(inputBatches)
    ptrs = (ctypes.c_void_p * max(n, 1))(*[b.ctypes.data for b in inputBatches])
    cnt = (ctypes.c_size_t * max(n, 1))(*counts)
    check(lib().parhip_gf64_compute_recovery_batches(
        ptrs, cnt, n, _buf(outputBuffer, "outputBuffer"), numInputs, numRecovery, blockSize // 8,
        firstInput, firstRecovery, numThreads))
    return outputBuffer


def compute_recovery_with_coeff(coeffMatrix, inputs, outputs, numInputs, numRecovery, blockSize,
                                firstInput=0, firstRecovery=0, numThreads=0, accumulate=False):
    _check_counts(numInputs, numRecovery)
    _check_block(blockSize)
    _check_sizes(inputs, numInputs, blockSize, "inputs")
    _check_sizes(outputs, numRecovery, blockSize, "outputs")
    _check_sizes(coeffMatrix, numInputs * numRecovery, 8, "coeffMatrix")
    check(lib().parhip_gf64_compute_recovery_blocks_with_coeff(
        _buf(inputs, "inputs"), numInputs, _buf(outputs, "outputs"), numRecovery, blockSize // 8,
        _buf(coeffMatrix, "coeffMatrix"), numThreads, 1 if accumulate else 0))
    return outputs


def build_coefficient_matrix(numInputs, numRecovery, firstInput, firstRecovery):
    if numInputs <= 0 or numRecovery <= 0:
        raise ValueError("numInputs and numRecovery must be positive")
    out = np.empty(numInputs * numRecovery, dtype=np.uint64)
    check(lib().parhip_gf64_build_cauchy_matrix(out.ctypes.data, numInputs, numRecovery,
                                                firstInput, firstRecovery))
    return out


def solve_and_reconstruct(A, rhsBlocks, n, blockSize):
    """Returns True on success, False on a singular matrix (A, rhs modified in place)."""
    if blockSize % 8:
        raise ValueError("blockSize must be a multiple of 8")
    _check_sizes(A, n * n, 8, "A")
    _check_sizes(rhsBlocks, n, blockSize, "rhsBlocks")
    rc = lib().parhip_gf64_solve_and_reconstruct(_buf(A, "A"), _buf(rhsBlocks, "rhsBlocks"), n,
                                                 blockSize // 8, 0)
    if rc == -1:
        return False
    check(rc)
    return True


def gf64_solve(A, b, n):
    x = np.zeros(n, dtype=np.uint64)
    A = np.ascontiguousarray(A, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    rc = lib().parhip_gf64_solve(A.ctypes.data, b.ctypes.data, x.ctypes.data, n)
    if rc == -1:
        return None
    check(rc)
    return x


def _check_words(buf, count, what):
    if count > buf.nbytes // 8:
        raise ValueError("%s buffer too small for %d words" % (what, count))


def _check_region(out, inBuf, coeff, len_words, n_coeff):
    # MulArr_NAPI / Gf64Encoder.mul_arr rules (gf64_addon.cc:3087-3106, :408-416):
    # non-negative counts, n_coeff > 0 when len_words > 0, every count within
    # its buffer (compared in words, so huge counts cannot wrap)
    if len_words < 0 or n_coeff < 0:
        raise ValueError("len_words and n_coeff must be non-negative")
    if len_words > 0 and n_coeff == 0:
        raise ValueError("n_coeff must be positive when len_words > 0")
    _check_words(out, len_words, "out")
    _check_words(inBuf, len_words, "in")
    _check_words(coeff, n_coeff, "coeff")


def _check_count(v, name):
    # Gf64Encoder's K / G bound (gf64_addon.cc:497-498, :655-656, :878-890)
    if v < 0 or v > 65536:
        raise ValueError("%s out of range [0, 65536]" % name)


def mul_arr(out, inBuf, coeff, len_words, n_coeff):
    _check_region(out, inBuf, coeff, len_words, n_coeff)
    check(lib().parhip_gf64_region_mul_arr(_buf(out, "out"), _buf(inBuf, "in"), _buf(coeff, "coeff"),
                                           len_words, n_coeff))
    return out


def _ptr_array(bufs):
    arr = (ctypes.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
    return arr


class Gf64Encoder:
    """Region operations (gf64/gf64_global.h:23-29) as the addon's Gf64Encoder exposes them."""

    def mul(self, out, inBuf, constant, len_words):
        if len_words < 0:
            raise ValueError("Length must be non-negative")
        _check_words(out, len_words, "out")
        _check_words(inBuf, len_words, "in")
        check(lib().parhip_gf64_region_mul(_buf(out, "out"), _buf(inBuf, "in"), len_words, constant))
        return out

    def mul_arr(self, out, inBuf, coeff, len_words, n_coeff):
        return mul_arr(out, inBuf, coeff, len_words, n_coeff)

    def muladd_arr(self, out, inBuf, coeff, len_words, n_coeff):
        _check_region(out, inBuf, coeff, len_words, n_coeff)
        check(lib().parhip_gf64_region_muladd_arr(_buf(out, "out"), _buf(inBuf, "in"),
                                                  _buf(coeff, "coeff"), len_words, n_coeff))
        return out

    def coupled_muladd_arr(self, out, in_blocks, coeff_blocks, len_words):
        _check_count(len(in_blocks), "G")
        if len_words < 0:
            raise ValueError("Length must be non-negative")
        _check_words(out, len_words, "out")
        _check_words(coeff_blocks, len(in_blocks), "coeff")
        for b in in_blocks:
            _buf(b, "in_blocks element")
            _check_words(b, len_words, "in_blocks element")
        ptrs = _ptr_array(in_blocks)
        check(lib().parhip_gf64_region_coupled_muladd_arr(
            _buf(out, "out"), ctypes.addressof(ptrs), _buf(coeff_blocks, "coeff"), len_words,
            len(in_blocks)))
        return out

    def fused_output_muladd_arr(self, outs, inBuf, coeffs, len_words):
        """outs[k] ^= in * coeffs[k] (coeffs: uint64 array; one pointer per output)."""
        _check_count(len(outs), "K")
        if len_words < 0:
            raise ValueError("Length must be non-negative")
        _check_words(inBuf, len_words, "in")
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64)
        if coeffs.size < len(outs):
            raise ValueError("Coefficients array length must equal K")
        for o in outs:
            _buf(o, "outs element")
            _check_words(o, len_words, "outs element")
        optrs = _ptr_array(outs)
        cptrs = (ctypes.c_void_p * len(outs))(*[coeffs.ctypes.data + 8 * k for k in range(len(outs))])
        check(lib().parhip_gf64_region_fused_output_muladd_arr(
            ctypes.addressof(optrs), _buf(inBuf, "in"), ctypes.addressof(cptrs), len_words, len(outs)))
        return outs

    def two_d_muladd_arr(self, outs, in_blocks, coeff2d, K_stride, len_words):
        _check_count(len(outs), "K")
        _check_count(len(in_blocks), "G")
        if len_words < 0:
            raise ValueError("len must be non-negative")
        if K_stride < len(in_blocks):
            raise ValueError("K_stride must be >= G")
        _check_count(K_stride, "K_stride")
        coeff2d = np.ascontiguousarray(coeff2d, dtype=np.uint64)
        if coeff2d.size < len(outs) * K_stride:
            raise ValueError("coeff_block_2d buffer too small for K * K_stride")
        for o in outs:
            _buf(o, "outs element")
            _check_words(o, len_words, "outs element")
        for b in in_blocks:
            _buf(b, "in_blocks element")
            _check_words(b, len_words, "in_blocks element")
        optrs = _ptr_array(outs)
        iptrs = _ptr_array(in_blocks)
        check(lib().parhip_gf64_region_2d_muladd_arr(
            ctypes.addressof(optrs), len(outs), ctypes.addressof(iptrs), len(in_blocks),
            coeff2d.ctypes.data, K_stride, len_words))
        return outs


def par3_create_streaming(sourcePath, options):
    """par3_create_streaming(sourcePath, options) as the addon's callback result.

    src/gf64_create_streaming.cc:327-470 (argument rules) and :93-240 (work):
    options has recoverySlices, blockSize and optional firstInput,
    firstRecovery, numThreads, useMmap.  Returns {recoveryBytes,
    throughputMBps, durationMs, recoveryBuffer (bytes)}; raises on error with
    the engine's message (Node-style errno code first for open failures).
    """
    if not isinstance(sourcePath, str) or not sourcePath:
        raise TypeError("sourcePath must be a non-empty string")
    if not isinstance(options, dict):
        raise TypeError("options must be an object")
    for k in ("recoverySlices", "blockSize"):
        if k not in options:
            raise TypeError("options.%s is required" % k)
    out = ctypes.POINTER(ctypes.c_uint8)()
    nbytes = ctypes.c_size_t(0)
    ms = ctypes.c_double(0)
    mbps = ctypes.c_double(0)
    L = lib()
    check(L.parhip_par3_create_streaming(
        sourcePath.encode(), int(options["recoverySlices"]), int(options["blockSize"]),
        int(options.get("firstInput") or 0), int(options.get("firstRecovery") or 0),
        int(options.get("numThreads") or 0), 1 if options.get("useMmap") else 0,
        ctypes.byref(out), ctypes.byref(nbytes), ctypes.byref(ms), ctypes.byref(mbps)))
    try:
        buf = ctypes.string_at(out, nbytes.value)
    finally:
        L.parhip_free(ctypes.cast(out, ctypes.c_void_p))
    return {"recoveryBytes": nbytes.value, "throughputMBps": mbps.value, "durationMs": ms.value,
            "recoveryBuffer": buf}


def set_stream_chunk(nbytes):
    """Staging chunk of the host-pointer entries for this thread (0 = default)."""
    check(lib().parhip_set_stream_chunk(int(nbytes)))


def assemble_recovery_packets(recoveryBlocksBuf, inputSetIdBuf, firstRecoveryIndex, numRecovery,
                              blockSize):
    """assemble_recovery_packets (src/gf64_addon.cc:2790-2938): bytes of
    numRecovery * (64 + blockSize) PAR3 REC packets, sealed with BLAKE3."""
    rec = bytes(recoveryBlocksBuf)
    isid = bytes(inputSetIdBuf)
    if len(isid) != 8:
        raise TypeError("inputSetIdBuf must be a 8-byte Buffer")
    if numRecovery <= 0:
        raise ValueError("numRecovery must be a positive integer")
    if blockSize <= 0:
        raise ValueError("blockSize must be a positive integer")
    if len(rec) != numRecovery * blockSize:
        raise ValueError("recoveryBlocksBuf length must equal numRecovery * blockSize")
    out = ctypes.create_string_buffer(numRecovery * (64 + blockSize))
    check(lib().parhip_assemble_recovery_packets(rec, isid, int(firstRecoveryIndex), numRecovery,
                                                 blockSize, out))
    return out.raw
