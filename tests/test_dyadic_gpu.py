"""The XOR-block (dyadic) form of the plain Cauchy contraction against the
oracle and against the direct nibble-table kernel.

The plain create's coefficients c(r, i) = f((fi + i) XOR (fr + r))
(src/par3_engine.cc:358-429) depend only on t XOR s inside aligned 8 x 8
blocks, which the kernel evaluates as subset convolutions.  Every shape here is
compared byte for byte: with the oracle over the whole output at small sizes,
and, at the 1 GiB shapes, the whole output of the two independent GPU kernels
(dyadic and direct tables) plus oracle word columns."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_words(rng, n, w):
    return np.frombuffer(rng.bytes(n * w * 8), dtype=np.uint64).reshape(n, w).copy()


def _torch_words(seed, n, w):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)


def _method(P, m):
    P.set_method(m)


DY_SHAPES = [
    # (N, words, R, firstInput, firstRecovery, accumulate)
    (8, 512, 8, 0, 8, False),          # one aligned block, one tile
    (64, 8192, 8, 0, 64, False),       # the 64 x 64 KiB case
    (9, 1025, 4, 0, 9, True),          # ragged block, partial output group, odd words
    (13, 511, 5, 3, 21, False),        # unaligned points on both sides
    (100, 1023, 7, 7, 1001, True),
    (33, 2049, 9, 1, 40, False),       # two output groups, the second holds one row
    (70, 300, 15, 6, 70, True),
    (40, 64, 16, 0, 40, False),
    (17, 1000, 17, 5, 2, False),       # overlapping point ranges: 0 -> 1
    (200, 777, 24, 0, 200, True),
    (5, 4096, 33, 4, 9, False),
    (1000, 32, 64, 0, 1000, False),
    (1, 1, 4, 0, 1, False),
    (3, 7, 12, 2 ** 64 - 5, 2 ** 63 + 3, True),  # points that wrap past 2^64
]


@pytest.mark.parametrize("n,w,r,fi,fr,acc", DY_SHAPES)
def test_dyadic_equals_oracle(gpu, oracle, n, w, r, fi, fr, acc):
    import torch
    from parparpar_amd import device
    rng = np.random.default_rng(n * 131 + w * 7 + r)
    inp = rand_words(rng, n, w)
    init = rand_words(rng, r, w)
    d_in = torch.from_numpy(inp.view(np.int64)).cuda()
    d_out = torch.from_numpy(init.view(np.int64).copy()).cuda()
    expect = init.copy() if acc else np.zeros((r, w), dtype=np.uint64)
    oracle.recovery_with_coeff(inp, oracle.cauchy(n, r, fi, fr), expect, acc)
    from parparpar_amd import lib
    try:
        for cache in (1, 0):  # coefficients from the cache / computed in the kernel
            gpu.check(lib().parhip_set_coeff_cache(cache))
            for m in (gpu.METHOD_DYADIC, gpu.METHOD_DYADIC8, gpu.METHOD_DYADIC16, gpu.METHOD_DYADIC32,
                      gpu.METHOD_TABLE, gpu.METHOD_AUTO):
                _method(gpu, m)
                out = d_out.clone()
                device.recovery(d_in, r, fi, fr, out=out, accumulate=acc)
                torch.cuda.synchronize()
                assert np.array_equal(out.cpu().numpy().view(np.uint64), expect), (cache, m, n, w, r, fi, fr)
    finally:
        gpu.check(lib().parhip_set_coeff_cache(1))
        _method(gpu, gpu.METHOD_AUTO)


@pytest.mark.parametrize("wpt,ib", [(4, 1), (2, 1), (2, 2)])
@pytest.mark.parametrize("groups", [1, 3, 40])
def test_dyadic_launch_shapes(gpu, oracle, wpt, ib, groups):
    """Every kernel variant and split-K row count gives the same bytes."""
    import torch
    from parparpar_amd import device, lib
    n, w, r, fi, fr = 331, 2560, 12, 5, 331
    rng = np.random.default_rng(wpt * 10 + ib + groups)
    inp = rand_words(rng, n, w)
    expect = oracle.recovery(inp, r, fi, fr)
    d_in = torch.from_numpy(inp.view(np.int64)).cuda()
    _method(gpu, gpu.METHOD_DYADIC)
    try:
        gpu.check(lib().parhip_tune_override(0, wpt, ib, groups))
        out = device.recovery(d_in, r, fi, fr)
        torch.cuda.synchronize()
    finally:
        gpu.check(lib().parhip_tune_override(0, 0, 0, 0))
        _method(gpu, gpu.METHOD_AUTO)
    assert np.array_equal(out.cpu().numpy().view(np.uint64), expect)


@pytest.mark.parametrize("wpt", [1, 2])
@pytest.mark.parametrize("groups", [1, 3, 40])
@pytest.mark.parametrize("cache", [1, 0])
def test_dyadic16_launch_shapes(gpu, oracle, wpt, groups, cache):
    """The 16 x 16 form: both tile widths, split-K rows through partials
    (3) and atomics (40), coefficients from the cache or the kernel; ragged
    and full tiles (2560 = 5 x 512 words, 1000 words)."""
    import torch
    from parparpar_amd import device, lib
    for n, w, r, fi, fr in ((331, 2560, 20, 5, 331), (97, 1000, 33, 13, 2 ** 40 + 7)):
        rng = np.random.default_rng(wpt * 100 + groups + n)
        inp = rand_words(rng, n, w)
        expect = oracle.recovery(inp, r, fi, fr)
        d_in = torch.from_numpy(inp.view(np.int64)).cuda()
        _method(gpu, gpu.METHOD_DYADIC16)
        try:
            gpu.check(lib().parhip_set_coeff_cache(cache))
            gpu.check(lib().parhip_tune_override(0, wpt, 0, groups))
            out = device.recovery(d_in, r, fi, fr)
            torch.cuda.synchronize()
        finally:
            gpu.check(lib().parhip_tune_override(0, 0, 0, 0))
            gpu.check(lib().parhip_set_coeff_cache(1))
            _method(gpu, gpu.METHOD_AUTO)
        assert np.array_equal(out.cpu().numpy().view(np.uint64), expect), (n, w, r, fi, fr)


@pytest.mark.parametrize("groups", [1, 3, 40])
@pytest.mark.parametrize("cache", [1, 0])
def test_dyadic32_launch_shapes(gpu, oracle, groups, cache):
    """The 32 x 32 form: split-K rows through partials (3) and atomics (40),
    coefficients from the cache or the kernel; full tiles (2560 = 5 x 512
    words) and ragged ones (1000 words); one and two output groups, points off
    the 32-point grid on both sides and past 2^40."""
    import torch
    from parparpar_amd import device, lib
    for n, w, r, fi, fr in ((331, 2560, 20, 5, 331), (97, 1000, 33, 13, 2 ** 40 + 7),
                            (130, 1536, 64, 31, 1)):
        rng = np.random.default_rng(groups * 1000 + n + cache)
        inp = rand_words(rng, n, w)
        expect = oracle.recovery(inp, r, fi, fr)
        d_in = torch.from_numpy(inp.view(np.int64)).cuda()
        _method(gpu, gpu.METHOD_DYADIC32)
        try:
            gpu.check(lib().parhip_set_coeff_cache(cache))
            gpu.check(lib().parhip_tune_override(0, 0, 0, groups))
            out = device.recovery(d_in, r, fi, fr)
            torch.cuda.synchronize()
        finally:
            gpu.check(lib().parhip_tune_override(0, 0, 0, 0))
            gpu.check(lib().parhip_set_coeff_cache(1))
            _method(gpu, gpu.METHOD_AUTO)
        assert np.array_equal(out.cpu().numpy().view(np.uint64), expect), (n, w, r, fi, fr)


def test_dyadic_reference_collision_shape(gpu, oracle):
    """The reference's overlapping-range case (test/par3-fenger-padded-engine.js:110-112):
    N=1000, R=512, 128-byte blocks, inputs at 0x10000, recovery at 66300."""
    import torch
    from parparpar_amd import device
    n, r, w, fi, fr = 1000, 512, 16, 0x10000, 66300
    rng = np.random.default_rng(66300)
    inp = rand_words(rng, n, w)
    out = device.recovery(torch.from_numpy(inp.view(np.int64)).cuda(), r, fi, fr)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint64), oracle.recovery(inp, r, fi, fr))


def test_dyadic_streamed_host_entry(gpu, oracle):
    """Host inputs larger than a staging chunk run chunk by chunk on one
    coefficient set; a chunk of 75 slices starts off the 8-input grid."""
    from parparpar_amd import binding
    n, w, r = 600, 512, 8
    rng = np.random.default_rng(75)
    inp = rand_words(rng, n, w)
    out = np.zeros((r, w), dtype=np.uint64)
    binding.set_stream_chunk(75 * w * 8)
    try:
        binding.compute_recovery(inp, out, n, r, w * 8, 3, n + 5)
    finally:
        binding.set_stream_chunk(0)
    assert np.array_equal(out, oracle.recovery(inp, r, 3, n + 5))


def test_full_size_dyadic32_equals_dyadic16(gpu, oracle):
    """1 GiB as 1024 x 1 MiB at R = 32: the whole outputs of the 32 x 32 and
    16 x 16 XOR-block kernels agree, and sampled word columns equal the
    oracle."""
    import torch
    from parparpar_amd import device
    n, w, r = 1024, 131072, 32
    print("full-size dyadic32 vs dyadic16 N=%d W=%d R=%d" % (n, w, r), flush=True)
    inp = _torch_words(n + w + r, n, w)
    try:
        _method(gpu, gpu.METHOD_DYADIC16)
        d16 = device.recovery(inp, r, 0, n)
        _method(gpu, gpu.METHOD_DYADIC32)
        d32 = device.recovery(inp, r, 0, n)
    finally:
        _method(gpu, gpu.METHOD_AUTO)
    torch.cuda.synchronize()
    assert torch.equal(d16, d32)
    rng = np.random.default_rng(w + r)
    cols = np.unique(np.concatenate([[0, w - 1], rng.integers(0, w, size=6)])).astype(np.int64)
    tc = torch.as_tensor(cols, device="cuda")
    sample = np.ascontiguousarray(inp[:, tc].cpu().numpy().view(np.uint64))
    ref = oracle.recovery_with_coeff(sample, oracle.cauchy(n, r, 0, n))
    assert np.array_equal(d32[:, tc].cpu().numpy().view(np.uint64), ref)
    del inp, d16, d32
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n,w,r", [(1024, 131072, 8), (262144, 512, 8)])
def test_full_size_dyadic_equals_tables(gpu, oracle, n, w, r):
    """1 GiB at R=8 in both geometries: the whole outputs of three independent
    kernels agree -- the dyadic (XOR-block) kernel, the direct nibble-table
    kernel and the GF(2) bit-matrix on fp4 MFMA -- and sampled word columns
    equal the oracle."""
    import torch
    from parparpar_amd import device
    print("full-size dyadic vs tables vs fp4 MFMA N=%d W=%d R=%d" % (n, w, r), flush=True)
    inp = _torch_words(n + w, n, w)
    try:
        _method(gpu, gpu.METHOD_TABLE)
        tab = device.recovery(inp, r, 0, n)
        _method(gpu, gpu.METHOD_DYADIC)
        dy = device.recovery(inp, r, 0, n)
        _method(gpu, gpu.METHOD_MFMA)
        mf = device.recovery(inp, r, 0, n)
    finally:
        _method(gpu, gpu.METHOD_AUTO)
    torch.cuda.synchronize()
    assert torch.equal(tab, dy)
    assert torch.equal(tab, mf)
    rng = np.random.default_rng(w)
    cols = np.unique(np.concatenate([[0, w - 1], rng.integers(0, w, size=14)])).astype(np.int64)
    tc = torch.as_tensor(cols, device="cuda")
    sample = np.ascontiguousarray(inp[:, tc].cpu().numpy().view(np.uint64))
    ref = oracle.recovery_with_coeff(sample, oracle.cauchy(n, r, 0, n))
    assert np.array_equal(dy[:, tc].cpu().numpy().view(np.uint64), ref)
    del inp, tab, dy, mf
    torch.cuda.empty_cache()
