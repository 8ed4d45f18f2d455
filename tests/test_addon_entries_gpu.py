"""The addon's barycentric and streaming recovery entries against the oracle.

compute_recovery_barycentric (src/gf64_addon.cc:1636-1796) routes to
GF64Controller::ComputeRecoveryBlocksBarycentric, whose two branches (R <= 32
delegation, R > 32 per-row inverse + muladd_arr,
src/par3_engine_barycentric.cc:121-229) both compute the Cauchy product;
compute_recovery_streaming (src/gf64_addon.cc:2349-2650) concatenates a list of
input batches and computes the same product once; compute_recovery_fenger
(:1980-2078) is documented by the reference as bit-exact with
ComputeRecoveryBlocks.  Cases: R on both sides of
32, accumulate, uneven and empty batches, batches that straddle staging chunks,
and the argument errors of both bindings.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_words(rng, *shape):
    return rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64)


# the last two are test/par3-barycentric-parallel-parity.js:167-168 (64 x 64 x 64 words at
# points (1000, 100000); 32768 inputs of 8 words, R = 256)
@pytest.mark.parametrize("n,w,r,fi,fr", [(5, 128, 2, 0, 5), (40, 64, 33, 0, 40), (17, 96, 48, 100, 3),
                                         (64, 512, 8, 0, 64), (64, 64, 64, 1000, 100000),
                                         (32768, 8, 256, 0, 32768)])
def test_barycentric(gpu, oracle, n, w, r, fi, fr):
    from parparpar_amd import binding
    rng = np.random.default_rng(n * r)
    inp = rand_words(rng, n, w)
    out = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery_barycentric(inp, out, n, r, w * 8, fi, fr)
    want = oracle.recovery(inp, r, fi, fr)
    assert np.array_equal(out, want)
    prev = rand_words(rng, r, w)
    acc = prev.copy()
    binding.compute_recovery_barycentric(inp, acc, n, r, w * 8, fi, fr, 0, True)
    assert np.array_equal(acc, prev ^ want)


def test_fenger(gpu, oracle):
    """compute_recovery_fenger writes the Cauchy product (R = 9, 16, and 8192 where the
    reference starts to choose it); zero counts leave the outputs alone."""
    from parparpar_amd import binding
    rng = np.random.default_rng(5)
    for n, w, r in [(12, 64, 9), (33, 40, 16), (4, 16, 8192)]:
        inp = rand_words(rng, n, w)
        out = rand_words(rng, r, w)
        binding.compute_recovery_fenger(inp, out, n, r, w * 8, 0, n)
        assert np.array_equal(out, oracle.recovery(inp, r, 0, n))
    keep = np.full((2, 8), 9, dtype=np.uint64)
    binding.compute_recovery_fenger(keep, keep, 0, 2, 64, 0, 0)
    assert (keep == 9).all()
    with pytest.raises(ValueError):
        binding.compute_recovery_fenger(keep, keep, -1, 2, 64, 0, 0)


# test/par3-fenger-padded-engine.js:103-127: odd and power-of-two N and R at
# points (0x10000, 0x1000000), N = 1, and the two overlapping point ranges
# (recovery points inside the input range: zero denominators take the 0 -> 1 rule)
@pytest.mark.parametrize("n,r,fi,fr", [(1000, 512, 0x10000, 0x1000000), (1000, 600, 0x10000, 0x1000000),
                                       (1024, 600, 0x10000, 0x1000000), (1, 512, 0x10000, 0x1000000),
                                       (1000, 512, 0x10000, 66300), (1024, 512, 0x10000, 66400)])
def test_fenger_padded_engine_cases(gpu, oracle, n, r, fi, fr):
    from parparpar_amd import binding
    rng = np.random.default_rng(n + r + fr)
    w = 128 // 8
    inp = rand_words(rng, n, w)
    out = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery_fenger(inp, out, n, r, 128, fi, fr)
    assert np.array_equal(out, oracle.recovery(inp, r, fi, fr))


def test_fenger_rejects_impossible_sizes(gpu):
    """The same file's guards: blockSize 2^32 + 8 and numInputs 2^63 are refused
    cleanly, before any work."""
    from parparpar_amd import binding
    b = np.zeros(1, dtype=np.uint64)
    with pytest.raises(ValueError):
        binding.compute_recovery_fenger(b, b.copy(), 1, 1, 2 ** 32 + 8, 0x10000, 0x1000000)
    with pytest.raises((ValueError, OverflowError)):
        binding.compute_recovery_fenger(b, b.copy(), 2 ** 63, 2, 8, 0x10000, 0x1000000)


def test_barycentric_errors(gpu):
    from parparpar_amd import binding
    buf = np.zeros(16, dtype=np.uint64)
    with pytest.raises(ValueError):
        binding.compute_recovery_barycentric(buf, buf, 0, 1, 8, 0, 1)
    with pytest.raises(ValueError):
        binding.compute_recovery_barycentric(buf, buf, 1, 0, 8, 0, 1)
    with pytest.raises(ValueError):
        binding.compute_recovery_barycentric(buf, buf, 1, 1, 12, 0, 1)
    with pytest.raises(ValueError):
        binding.compute_recovery_barycentric(buf, buf, 3, 1, 64, 0, 3)


@pytest.mark.parametrize("sizes", [[7], [1, 2, 3, 4], [0, 5, 0, 0, 6, 0], [10, 0]])
def test_streaming_batches(gpu, oracle, sizes):
    from parparpar_amd import binding
    rng = np.random.default_rng(sum(sizes) * 13 + len(sizes))
    n, w, r = sum(sizes), 80, 5
    inp = rand_words(rng, n, w)
    batches, o = [], 0
    for k in sizes:
        batches.append(inp[o:o + k].copy())
        o += k
    out = np.full((r, w), 0xAB, dtype=np.uint64)  # written, not accumulated
    binding.compute_recovery_streaming(batches, out, n, r, w * 8, 3, n + 3)
    assert np.array_equal(out, oracle.recovery(inp, r, 3, n + 3))


def test_streaming_through_the_staging_ring(gpu, oracle):
    """Small staging chunks (3 slices) so chunks start and end inside batches."""
    from parparpar_amd import binding
    rng = np.random.default_rng(77)
    sizes = [2, 5, 0, 1, 7, 4]
    n, w, r = sum(sizes), 256, 4
    inp = rand_words(rng, n, w)
    batches, o = [], 0
    for k in sizes:
        batches.append(inp[o:o + k].copy())
        o += k
    binding.set_stream_chunk(3 * w * 8)
    try:
        out = np.zeros((r, w), dtype=np.uint64)
        binding.compute_recovery_streaming(batches, out, n, r, w * 8, 0, n)
    finally:
        binding.set_stream_chunk(0)
    assert np.array_equal(out, oracle.recovery(inp, r, 0, n))


def test_streaming_errors(gpu):
    import parparpar_amd as P
    from parparpar_amd import binding
    out = np.zeros((2, 8), dtype=np.uint64)
    b = np.zeros((3, 8), dtype=np.uint64)
    with pytest.raises(TypeError):
        binding.compute_recovery_streaming(b, out, 3, 2, 64, 0, 3)
    with pytest.raises(TypeError):
        binding.compute_recovery_streaming([b, b"x" * 64], out, 4, 2, 64, 0, 4)
    with pytest.raises(ValueError):
        binding.compute_recovery_streaming([b], out, -1, 2, 64, 0, 3)
    with pytest.raises(ValueError):
        binding.compute_recovery_streaming([b], out, 3, 2, 60, 0, 3)
    with pytest.raises(ValueError):
        binding.compute_recovery_streaming([b], out, 2, 2, 64, 0, 3)  # exceeds numInputs
    with pytest.raises(ValueError):
        binding.compute_recovery_streaming([b], out, 4, 2, 64, 0, 4)  # does not equal numInputs
    with pytest.raises(ValueError):
        binding.compute_recovery_streaming([np.zeros(9, dtype=np.uint8)], out, 1, 2, 8, 0, 1)
    # zero counts: outputs untouched
    keep = np.full((2, 8), 7, dtype=np.uint64)
    binding.compute_recovery_streaming([], keep, 0, 2, 64, 0, 0)
    assert (keep == 7).all()
    # the C entry checks the batch counts itself
    import ctypes
    ptrs = (ctypes.c_void_p * 1)(b.ctypes.data)
    cnt = (ctypes.c_size_t * 1)(2)
    assert P.lib().parhip_gf64_compute_recovery_batches(ptrs, cnt, 1, out.ctypes.data, 3, 2, 8, 0, 3,
                                                        0) == P.PARHIP_ERR_RANGE


def test_workspace_query_empty_shapes(gpu):
    """The workspace query is defined for empty shapes (no division by an empty tile count)."""
    import parparpar_amd as P
    L = P.lib()
    assert L.parhip_gf64_workspace_bytes(0, 8, 128) == 0
    assert L.parhip_gf64_workspace_bytes(8, 0, 128) == 0
    assert L.parhip_gf64_workspace_bytes(8, 8, 0) == 0
    assert L.parhip_gf64_workspace_bytes(1024, 8, 131072) >= 0
    assert L.parhip_gf64_repair_workspace_bytes(64, 512, 0) >= 0
    assert L.parhip_gf64_repair_workspace_bytes(8, 512, 8) > 0
