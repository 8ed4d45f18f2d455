"""GF(2^64) parity: the gfx950 engine against the oracle, through the C ABI.

Bit-exact is the bar everywhere (integer field arithmetic).  Small shapes are
compared in full; the 1 GiB shapes through exact samples of word columns (each
column of the contraction is independent, so a sampled column is a full
check of that column) plus size-independent properties.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "par3-golden.bin")


def rand_words(rng, *shape):
    return rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64)


def test_golden_vector(gpu, oracle):
    from parparpar_amd import binding
    n, b = 5, 1024
    inp = np.zeros((n, b // 8), dtype=np.uint64)
    for f in range(n):
        inp[f] = np.arange(b // 8, dtype=np.uint64) * 8 + f * b
    out = np.zeros((2, b // 8), dtype=np.uint64)
    binding.compute_recovery(inp, out, n, 2, b, 0, n, 0)
    assert out.tobytes() == open(GOLDEN, "rb").read()
    assert "%08X" % oracle.crc32(out.tobytes()) == "6A941C15"


def test_kat(gpu):
    from parparpar_amd import binding
    out = np.zeros(1, dtype=np.uint64)
    binding.compute_recovery(np.array([1, 2], dtype=np.uint64), out, 2, 1, 8, 0, 2)
    assert int(out[0]) == 0x7ffffffffffffffa


def test_native_fidelity_cauchy_consistency(gpu, oracle):
    # test/par3-native-fidelity.js:302-360: inputs [1, 2, 3], R = 2, points (0, 3)
    from parparpar_amd import binding
    inp = np.array([1, 2, 3], dtype=np.uint64)
    out = np.zeros(2, dtype=np.uint64)
    binding.compute_recovery(inp, out, 3, 2, 8, 0, 3)
    co = oracle.cauchy(3, 2, 0, 3)
    expect = [oracle.gf64_mul(1, int(co[r, 0])) ^ oracle.gf64_mul(2, int(co[r, 1])) ^
              oracle.gf64_mul(3, int(co[r, 2])) for r in range(2)]
    assert [int(x) for x in out] == expect


def test_config1_64x64k_r8(gpu, oracle):
    # 64 input slices x 64 KiB of seeded random bytes, R=8, points (0, 64)
    from parparpar_amd import binding
    rng = np.random.default_rng(0xC1)
    n, words, r = 64, 65536 // 8, 8
    inp = np.frombuffer(rng.bytes(n * 65536), dtype=np.uint64).reshape(n, words).copy()
    out = np.zeros((r, words), dtype=np.uint64)
    binding.compute_recovery(inp, out, n, r, 65536, 0, n)
    ref = oracle.recovery(inp, r, 0, n)
    assert np.array_equal(out, ref)
    # negative trap: a one-bit change in one input changes exactly that column
    inp[17, 1234] ^= np.uint64(1 << 63)
    out2 = np.zeros_like(out)
    binding.compute_recovery(inp, out2, n, r, 65536, 0, n)
    diff = np.nonzero((out2 != out).any(axis=0))[0]
    assert diff.tolist() == [1234]


SHAPES = [
    # (N, words, R, firstInput, firstRecovery, accumulate)
    (1, 1, 1, 0, 1, False),
    (2, 3, 1, 0, 2, True),
    (3, 64, 2, 5, 100, False),
    (7, 513, 3, 0, 7, True),
    (9, 1025, 5, 0, 9, False),
    (17, 2048, 8, 1, 40, True),
    (33, 4099, 9, 0, 33, False),
    (64, 777, 16, 0, 64, True),
    (10, 100, 17, 0, 10, False),
    (50, 300, 33, 3, 1000, True),
    (200, 64, 8, 0, 200, False),
    (1000, 32, 4, 0, 1000, True),
    (4, 8, 6, 2, 3, False),   # overlapping point ranges: zero denominator -> 1
]


@pytest.mark.parametrize("n,w,r,fi,fr,acc", SHAPES)
def test_shapes(gpu, oracle, n, w, r, fi, fr, acc):
    from parparpar_amd import binding
    rng = np.random.default_rng(n * 7919 + w * 31 + r)
    inp = rand_words(rng, n, w)
    init = rand_words(rng, r, w)
    out = init.copy()
    binding.compute_recovery(inp, out, n, r, w * 8, fi, fr, 0, accumulate=acc)
    expect = init.copy() if acc else np.zeros((r, w), dtype=np.uint64)
    oracle.recovery_with_coeff(inp, oracle.cauchy(n, r, fi, fr), expect, acc)
    assert np.array_equal(out, expect)


def test_with_coeff_and_matrix(gpu, oracle):
    from parparpar_amd import binding
    rng = np.random.default_rng(5)
    n, r, w = 37, 11, 200
    m = binding.build_coefficient_matrix(n, r, 3, 500)
    assert np.array_equal(m.reshape(r, n), oracle.cauchy(n, r, 3, 500))
    big = binding.build_coefficient_matrix(1000, 40, 0, 1000)
    assert np.array_equal(big.reshape(40, 1000), oracle.cauchy(1000, 40, 0, 1000))
    # larger matrices: wide, narrow (rows shorter than a wave) and ragged
    # shapes, and colliding points (0 -> 1)
    for (bn, br, bfi, bfr) in [(1000, 100, 0, 1000), (37, 2000, 5, 9), (65537, 1, 0, 65537),
                               (3, 30000, 0, 1), (4099, 17, 100, 100)]:
        got = binding.build_coefficient_matrix(bn, br, bfi, bfr).reshape(br, bn)
        assert np.array_equal(got, oracle.cauchy(bn, br, bfi, bfr)), (bn, br, bfi, bfr)
    co = rand_words(rng, r, n)  # arbitrary coefficients, including structure-free ones
    co[0, :] = 0
    co[1, :] = 1
    inp = rand_words(rng, n, w)
    out = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery_with_coeff(co, inp, out, n, r, w * 8)
    assert np.array_equal(out, oracle.recovery_with_coeff(inp, co))
    assert not out[0].any()
    x = np.zeros(w, dtype=np.uint64)
    for i in range(n):
        x ^= inp[i]
    assert np.array_equal(out[1], x)  # coefficient 1 -> plain XOR


def test_chunked_accumulate_equals_single(gpu, oracle):
    # lib/par3gen.js:2306-2316: compute_recovery_full per chunk with
    # firstInput=chunkStart, firstRecovery=N, accumulate=true
    from parparpar_amd import binding
    rng = np.random.default_rng(6)
    n, w, r = 100, 257, 8
    inp = rand_words(rng, n, w)
    single = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery_full(inp, single, n, r, w * 8, 0, n)
    acc = np.zeros((r, w), dtype=np.uint64)
    for s0 in range(0, n, 30):
        chunk = np.ascontiguousarray(inp[s0:s0 + 30])
        binding.compute_recovery_accumulate(chunk, acc, len(chunk), r, w * 8, s0, n)
    assert np.array_equal(acc, single)
    assert np.array_equal(single, oracle.recovery(inp, r, 0, n))


def test_solve_and_reconstruct(gpu, oracle):
    from parparpar_amd import binding
    rng = np.random.default_rng(7)
    for k, w in [(1, 5), (3, 64), (8, 1000), (20, 129)]:
        A = oracle.cauchy(k, k, 0, 300)
        rhs = rand_words(rng, k, w)
        rc, A_o, x_o = oracle.solve_and_reconstruct(A, rhs)
        A_g, x_g = A.copy(), rhs.copy()
        assert binding.solve_and_reconstruct(A_g, x_g, k, w * 8) is True
        assert np.array_equal(A_g, A_o) and np.array_equal(x_g, x_o)
    # pivoting: zero leading entry forces a row swap
    A = np.array([[0, 5, 7], [3, 0, 1], [2, 9, 0]], dtype=np.uint64)
    rhs = rand_words(rng, 3, 17)
    rc, A_o, x_o = oracle.solve_and_reconstruct(A, rhs)
    A_g, x_g = A.copy(), rhs.copy()
    assert binding.solve_and_reconstruct(A_g, x_g, 3, 17 * 8) is (rc == 0)
    assert np.array_equal(A_g, A_o) and np.array_equal(x_g, x_o)
    # singular: return false, leave the reference's partially eliminated state
    A = np.array([[2, 3, 5], [4, 6, 10], [7, 1, 9]], dtype=np.uint64)
    A[1] = [oracle.gf64_mul(int(v), 9) for v in A[0]]
    rhs = rand_words(rng, 3, 9)
    rc, A_o, x_o = oracle.solve_and_reconstruct(A, rhs)
    assert rc == -1
    A_g, x_g = A.copy(), rhs.copy()
    assert binding.solve_and_reconstruct(A_g, x_g, 3, 9 * 8) is False
    assert np.array_equal(A_g, A_o) and np.array_equal(x_g, x_o)


def test_gf64_solve(gpu, oracle):
    from parparpar_amd import binding
    rng = np.random.default_rng(8)
    A = oracle.cauchy(6, 6, 0, 50)
    b = rand_words(rng, 6)
    x = binding.gf64_solve(A, b, 6)
    rc, _, xo = oracle.solve_and_reconstruct(A, b.reshape(6, 1))
    assert np.array_equal(x, xo.reshape(6))
    assert binding.gf64_solve(np.zeros((2, 2), dtype=np.uint64), b[:2], 2) is None


def test_region_ops(gpu, oracle):
    from parparpar_amd import binding
    L = oracle.lib()
    P = oracle.p64
    enc = binding.Gf64Encoder()
    rng = np.random.default_rng(9)
    ln = 1031
    x = rand_words(rng, ln)
    co = rand_words(rng, 7)
    # mul
    out = np.zeros(ln, dtype=np.uint64)
    enc.mul(out, x, int(co[0]), ln)
    ref = np.zeros(ln, dtype=np.uint64)
    L.or_gf64_region_mul(P(ref), P(x), ln, int(co[0]))
    assert np.array_equal(out, ref)
    # mul_arr cycling, n = 1, 3, 7 (gf64/test/test_gf64_region_mularr.c)
    for nc in (1, 3, 7):
        out = np.zeros(ln, dtype=np.uint64)
        binding.mul_arr(out, x, co, ln, nc)
        ref = np.zeros(ln, dtype=np.uint64)
        L.or_gf64_region_mul_arr(P(ref), P(x), P(co), ln, nc)
        assert np.array_equal(out, ref), nc
    # muladd_arr: sum over all coefficients
    init = rand_words(rng, ln)
    out = init.copy()
    enc.muladd_arr(out, x, co, ln, 5)
    ref = init.copy()
    L.or_gf64_region_muladd_arr(P(ref), P(x), P(co), ln, 5)
    assert np.array_equal(out, ref)
    # coupled / fused / 2d
    G, K = 5, 4
    ins = [rand_words(rng, ln) for _ in range(G)]
    outs0 = [rand_words(rng, ln) for _ in range(K)]
    c2d = rand_words(rng, K, G + 3)  # K_stride = G + 3
    o = outs0[0].copy()
    enc.coupled_muladd_arr(o, ins, np.ascontiguousarray(c2d[0, :G]), ln)
    r = outs0[0].copy()
    for g in range(G):
        for i in range(ln):
            r[i] ^= np.uint64(oracle.gf64_mul(int(ins[g][i]), int(c2d[0, g])))
    assert np.array_equal(o, r)
    outs = [a.copy() for a in outs0]
    enc.fused_output_muladd_arr(outs, ins[0], c2d[:, 0].copy(), ln)
    for k in range(K):
        r = outs0[k].copy()
        for i in range(ln):
            r[i] ^= np.uint64(oracle.gf64_mul(int(ins[0][i]), int(c2d[k, 0])))
        assert np.array_equal(outs[k], r)
    outs = [a.copy() for a in outs0]
    enc.two_d_muladd_arr(outs, ins, c2d, G + 3, ln)
    co2 = np.ascontiguousarray(c2d[:, :G])
    exp = np.stack(outs0).copy()
    oracle.recovery_with_coeff(np.stack(ins), co2, exp, True)
    for k in range(K):
        assert np.array_equal(outs[k], exp[k])


def test_input_errors(gpu):
    from parparpar_amd import binding
    with pytest.raises(ValueError):
        binding.compute_recovery(np.zeros(2, np.uint64), np.zeros(2, np.uint64), 2, 1, 7, 0, 2)
    # N == 0 / R == 0: the addon throws RangeError and leaves the output alone
    # (test/par3-native-edge.js scenarios 1-2); the engine entry is a silent
    # no-op (src/par3_engine.cc:1069)
    import parparpar_amd as P
    out = np.full(4, 5, dtype=np.uint64)
    for n, r in [(0, 1), (1, 0), (-1, 1)]:
        with pytest.raises(ValueError):
            binding.compute_recovery(np.zeros(4, np.uint64), out, n, r, 32, 0, 0)
        with pytest.raises(ValueError):
            binding.compute_recovery_full(np.zeros(4, np.uint64), out, n, r, 32, 0, 0)
    with pytest.raises(ValueError):
        binding.build_coefficient_matrix(0, 2, 0, 0)
    assert (out == 5).all()
    assert P.lib().parhip_gf64_compute_recovery_blocks(out.ctypes.data, 0, out.ctypes.data, 1, 4, 0, 0, 0,
                                                       0) == 0
    assert (out == 5).all()
    # scenarios 5-8, 13: bad block sizes, an impossible numInputs, a short output
    with pytest.raises(ValueError):
        binding.compute_recovery(np.zeros(2, np.uint64), np.zeros(1, np.uint64), 2, 1, 0, 0, 2)
    with pytest.raises(ValueError):
        binding.compute_recovery(np.zeros(128, np.uint64), np.zeros(128, np.uint64), 2147483647, 1, 8, 0, 1)
    with pytest.raises(ValueError):
        binding.compute_recovery(np.zeros(2, np.uint64), np.zeros(4, np.uint8), 2, 1, 8, 0, 2)


def test_output_larger_than_needed_and_state_isolation(gpu, oracle):
    """test/par3-native-edge.js scenarios 14 and 18: bytes past numRecovery x
    blockSize are not written; a small call between two large ones changes
    nothing in the second."""
    from parparpar_amd import binding
    rng = np.random.default_rng(18)
    inp = rand_words(rng, 2, 1)
    out = np.full(2, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    binding.compute_recovery(inp, out, 2, 1, 8, 0, 2, 1)
    assert out[1] == 0xFFFFFFFFFFFFFFFF and out[0] == oracle.recovery(inp, 1, 0, 2)[0, 0]
    big = rand_words(rng, 32, 131072)
    o1 = np.zeros((10, 131072), dtype=np.uint64)
    binding.compute_recovery(big, o1, 32, 10, 1 << 20, 0, 32)
    small = np.zeros((1, 1), dtype=np.uint64)
    binding.compute_recovery(rand_words(rng, 2, 1), small, 2, 1, 8, 0, 2, 1)
    o2 = np.zeros_like(o1)
    binding.compute_recovery(big, o2, 32, 10, 1 << 20, 0, 32)
    assert np.array_equal(o1, o2)
    cols = [0, 1, 65535, 131071]
    assert np.array_equal(o1[:, cols], oracle.recovery_columns(big, oracle.cauchy(32, 10, 0, 32), cols))


# ----------------------------------------------------------- device-resident, full size

def _torch_words(rng, n, w, device="cuda"):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(rng.integers(0, 2 ** 31)))
    return torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device=device, generator=g)


@pytest.mark.parametrize("n,w,r", [(1024, 131072, 8), (262144, 512, 8)])
def test_full_size_create_columns(gpu, oracle, n, w, r):
    """1 GiB at R=8 in both geometries: exact check of sampled word columns."""
    import torch
    from parparpar_amd import device
    rng = np.random.default_rng(n + w)
    print("full-size create N=%d W=%d R=%d" % (n, w, r), flush=True)
    inp = _torch_words(rng, n, w)
    out = device.recovery(inp, r, 0, n)
    torch.cuda.synchronize()
    cols = np.unique(np.concatenate([[0, w - 1], rng.integers(0, w, size=14)])).astype(np.int64)
    sample = inp[:, torch.as_tensor(cols, device="cuda")].cpu().numpy().view(np.uint64)
    sample = np.ascontiguousarray(sample)
    co = oracle.cauchy(n, r, 0, n)
    ref = oracle.recovery_with_coeff(sample, co)
    got = out[:, torch.as_tensor(cols, device="cuda")].cpu().numpy().view(np.uint64)
    assert np.array_equal(got, ref)
    # size-independent property: linearity over the whole 1 GiB input
    out2 = device.recovery(inp, r, 0, n)
    assert torch.equal(out, out2)  # deterministic
    del inp, out, out2
    torch.cuda.empty_cache()


def test_device_strided_and_unaligned(gpu, oracle):
    import torch
    from parparpar_amd import device
    rng = np.random.default_rng(21)
    n, w, r = 13, 301, 5
    base = _torch_words(rng, n, w + 3)
    view = base[:, 1:w + 1]  # 8-byte aligned rows, odd stride: exercises the 64-bit load path
    out = torch.zeros((r, w), dtype=torch.int64, device="cuda")
    device.recovery(view, r, 0, n, out=out)
    host = np.ascontiguousarray(view.cpu().numpy().view(np.uint64))
    assert np.array_equal(out.cpu().numpy().view(np.uint64), oracle.recovery(host, r, 0, n))


@pytest.mark.parametrize("shift_in,shift_out,acc", [(1, 3, False), (4, 0, True), (7, 5, True)])
def test_host_buffers_not_8_byte_aligned(gpu, oracle, shift_in, shift_out, acc):
    """Host buffers at any byte address: the addon bounce-copies 8-byte
    misaligned Buffers (src/gf64_addon.cc:1466-1505); the library takes them as
    they are. Small (one copy) and multi-chunk (streamed) sizes."""
    import parparpar_amd as P
    from parparpar_amd import binding
    rng = np.random.default_rng(shift_in * 10 + shift_out)
    L = P.lib()
    for n, w, r in [(5, 129, 3), (40, 8192, 6)]:
        inp = rand_words(rng, n, w)
        init = rand_words(rng, r, w)
        raw_in = np.zeros(n * w * 8 + 16, dtype=np.uint8)
        raw_out = np.zeros(r * w * 8 + 16, dtype=np.uint8)
        vin = raw_in[shift_in:shift_in + n * w * 8]
        vout = raw_out[shift_out:shift_out + r * w * 8]
        vin[:] = inp.view(np.uint8).reshape(-1)
        vout[:] = init.view(np.uint8).reshape(-1)
        assert vin.ctypes.data % 8 == shift_in % 8
        P.check(L.parhip_set_stream_chunk(1 << 20))   # the 2.5 MiB case streams in chunks
        try:
            binding.compute_recovery(vin, vout, n, r, w * 8, 0, n, 0, accumulate=acc)
        finally:
            P.check(L.parhip_set_stream_chunk(0))
        expect = init.copy() if acc else np.zeros((r, w), dtype=np.uint64)
        oracle.recovery_with_coeff(inp, oracle.cauchy(n, r, 0, n), expect, acc)
        got = np.frombuffer(vout.tobytes(), dtype=np.uint64).reshape(r, w)
        assert np.array_equal(got, expect), (n, w, r)
        assert not raw_out[:shift_out].any() and not raw_out[shift_out + r * w * 8:].any()


# Compute scenarios of test/par3-native-edge.js (3, 4, 9-12, 15, 16):
# (numInputs, numRecovery, blockSize, firstInput, firstRecovery, numThreads)
EDGE = [(1, 1, 8, 0, 2, 1), (2, 1, 8, 0, 2, 0), (4, 3, 64, 0, 4, 9999), (4, 3, 64, 0, 4, 0),
        (5, 100, 256, 0, 5, 1), (6, 4, 128, 1000, 2000, 2), (10, 100, 512, 0, 10, 0)]


@pytest.mark.parametrize("n,r,bs,fi,fr,threads", EDGE)
def test_reference_edge_scenarios(gpu, oracle, n, r, bs, fi, fr, threads):
    from parparpar_amd import binding
    rng = np.random.default_rng(n * 131 + r * 7 + bs)
    w = bs // 8
    inp = rand_words(rng, n, w)
    if (n, r, bs) == (1, 1, 8):
        inp[0, 0] = np.uint64(0xDEADBEEFCAFEBABE)     # scenario 3's word
    out = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery(inp, out, n, r, bs, fi, fr, threads)
    assert np.array_equal(out, oracle.recovery_with_coeff(inp, oracle.cauchy(n, r, fi, fr)))
    assert out.any()
    if n == 1:
        assert not np.array_equal(out[0], inp[0])     # Cauchy coefficient 1/(0 ^ 2) != 1
    again = np.zeros_like(out)
    binding.compute_recovery(inp, again, n, r, bs, fi, fr, threads)   # scenario 16
    assert np.array_equal(out, again)


def test_reference_edge_64mib_blocks(gpu, oracle):
    """Scenario 7: 2 inputs x 2 recovery of 64 MiB blocks (host buffers, streamed)."""
    from parparpar_amd import binding
    rng = np.random.default_rng(7)
    w = (64 << 20) // 8
    inp = np.frombuffer(rng.bytes(2 * w * 8), dtype=np.uint64).reshape(2, w).copy()
    out = np.zeros((2, w), dtype=np.uint64)
    binding.compute_recovery(inp, out, 2, 2, 64 << 20, 0, 2, 0)
    cols = np.unique(np.concatenate([[0, 1, w // 2, w - 1], rng.integers(0, w, size=60)]))
    exp = oracle.recovery_with_coeff(np.ascontiguousarray(inp[:, cols]), oracle.cauchy(2, 2, 0, 2))
    assert np.array_equal(out[:, cols], exp)
    assert out[0, 0] != 0


@pytest.mark.parametrize("ln", [0, 1, 2, 3, 4, 5, 8, 16, 17, 33])
def test_region_mul_arr_reference_grid(gpu, oracle, ln):
    """gf64/test/test_gf64_region_mularr.c:50-85: every length in its list with
    n_coeff in {1, 2, 3, 4, 7}; out[i] = in[i] * coeff[i % n_coeff] (cycling, not
    a dot product), and nothing past len is written."""
    from parparpar_amd import binding
    rng = np.random.default_rng(0xD3D3 + ln)
    x = rand_words(rng, max(ln, 1))
    for nc in (1, 2, 3, 4, 7):
        co = rand_words(rng, nc)
        out = np.full(ln + 2, 0xAAAAAAAAAAAAAAAA, dtype=np.uint64)
        binding.mul_arr(out, x, co, ln, nc)
        expect = [oracle.gf64_mul(int(x[i]), int(co[i % nc])) for i in range(ln)]
        assert [int(v) for v in out[:ln]] == expect, (ln, nc)
        assert (out[ln:] == np.uint64(0xAAAAAAAAAAAAAAAA)).all()


@pytest.mark.parametrize("G", [1, 2, 4, 8, 32, 64, 248])
def test_repair_rhs_kernel_grid(gpu, oracle, G):
    """test/par3-repair-rhs-regression.js:52-128: coupled_muladd_arr per row and
    two_d_muladd_arr for all rows, over G in its list, K in {1, 2, 4, 8} and
    len in {1, 7, 8, 32, 64, 128} words, against the oracle's contraction; plus
    its negative trap (one corrupted coefficient changes the result)."""
    from parparpar_amd import binding
    enc = binding.Gf64Encoder()
    rng = np.random.default_rng(90 + G)
    for K in (1, 2, 4, 8):
        for ln in (1, 7, 8, 32, 64, 128):
            ins = rand_words(rng, G, ln)
            c2d = rand_words(rng, K, G)
            ref = oracle.recovery_with_coeff(ins, c2d)
            for k in range(K):
                o = np.zeros(ln, dtype=np.uint64)
                enc.coupled_muladd_arr(o, list(ins), np.ascontiguousarray(c2d[k]), ln)
                assert np.array_equal(o, ref[k]), ("coupled", G, K, ln, k)
            outs = [np.zeros(ln, dtype=np.uint64) for _ in range(K)]
            enc.two_d_muladd_arr(outs, list(ins), c2d, G, ln)
            assert np.array_equal(np.stack(outs), ref), ("2d", G, K, ln)
    bad = c2d.copy()
    bad[0, 0] ^= np.uint64(1)
    outs = [np.zeros(ln, dtype=np.uint64) for _ in range(K)]
    enc.two_d_muladd_arr(outs, list(ins), bad, G, ln)
    assert not np.array_equal(np.stack(outs), ref)


def test_past_4gib_boundary(gpu, oracle):
    """test/par3-repair-4100-hash.js guards offsets past 4 GiB: 4100 slices of
    1 MiB (4.0 GiB + 4 MiB) on the device and, streamed, from host memory; the
    columns sampled against the oracle include every input, those past the
    boundary too, and a one-word change in input 4099 shows up in its column."""
    import torch
    from parparpar_amd import binding, device
    n, w, r = 4100, (1 << 20) // 8, 2
    g = torch.Generator(device="cuda")
    g.manual_seed(4100)
    d_in = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
    d_out = device.recovery(d_in, r, 0, n)
    torch.cuda.synchronize()
    cols = np.array([0, 1, w // 2, w - 2, w - 1], dtype=np.int64)
    tc = torch.as_tensor(cols, device="cuda")
    samp = np.ascontiguousarray(d_in[:, tc].cpu().numpy().view(np.uint64))
    co = oracle.cauchy(n, r, 0, n)
    exp = oracle.recovery_with_coeff(samp, co)
    assert np.array_equal(d_out[:, tc].cpu().numpy().view(np.uint64), exp)
    # host entry over the same bytes (a 4.0 GiB pageable buffer, streamed)
    host = d_in.cpu().numpy().view(np.uint64)
    out = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery(host, out, n, r, w * 8, 0, n)
    assert np.array_equal(out, d_out.cpu().numpy().view(np.uint64))
    # the last input sits past the boundary: flipping one of its words moves its column only
    d_in[n - 1, w - 1] ^= 1
    d_out2 = device.recovery(d_in, r, 0, n)
    torch.cuda.synchronize()
    diff = (d_out2 != d_out).any(dim=0).nonzero().flatten().cpu().tolist()
    assert diff == [w - 1]
    del d_in, d_out, d_out2, host
    torch.cuda.empty_cache()


def test_mul_arr_semantics_reference_cases(gpu, oracle):
    """gf64/test/test.c:292-346: n_coeff = 1 is a scalar multiply, n_coeff > 1
    cycles, coefficient 0 zeroes, 1 is the identity, and in == out works."""
    from parparpar_amd import binding
    x = np.array([1, 2, 3, 4], dtype=np.uint64)
    m = oracle.gf64_mul
    out = np.zeros(4, dtype=np.uint64)
    binding.mul_arr(out, x, np.array([0xDEADBEEF12345678], dtype=np.uint64), 4, 1)
    assert [int(v) for v in out] == [m(int(v), 0xDEADBEEF12345678) for v in x]
    co = np.array([0xAA00AA00AA00AA00, 0xBB00BB00BB00BB00], dtype=np.uint64)
    binding.mul_arr(out, x, co, 4, 2)
    assert [int(v) for v in out] == [m(int(x[i]), int(co[i % 2])) for i in range(4)]
    out[:] = 0xDEADBEEF
    binding.mul_arr(out, x, np.array([0], dtype=np.uint64), 4, 1)
    assert not out.any()
    binding.mul_arr(out, x, np.array([1], dtype=np.uint64), 4, 1)
    assert np.array_equal(out, x)
    buf = x.copy()
    co = np.array([0xCC00CC00CC00CC00, 0xDD00DD00DD00DD00], dtype=np.uint64)
    binding.mul_arr(buf, buf, co, 4, 2)
    assert [int(v) for v in buf] == [m(int(x[i]), int(co[i % 2])) for i in range(4)]


@pytest.mark.gpu
def test_measure_lds_gather(gpu):
    """The LDS gather-rate measurement returns plausible per-read times (ns per
    wave-level read per CU), 16-byte reads no cheaper than 8-byte ones."""
    import ctypes
    import parparpar_amd as P
    a, b = ctypes.c_double(), ctypes.c_double()
    P.check(P.lib().parhip_measure_lds_gather(ctypes.byref(a), ctypes.byref(b)))
    assert 0.2 < a.value < 50.0 and 0.1 < b.value < 50.0, (a.value, b.value)
    assert a.value >= b.value * 0.9, (a.value, b.value)
    assert P.lib().parhip_measure_lds_gather(None, ctypes.byref(b)) == P.PARHIP_ERR_ARG
