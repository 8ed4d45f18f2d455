"""The reference's randomized kernel parity grid, replayed against the oracle.

test/par3-kernel-parity.js draws its cases from mulberry32 (seeds 0xDEADBEEF,
0xFEEDFACE, 0xBADDCAFE) and checks compute_recovery and Gf64Encoder.mul_arr
against a BigInt restatement.  The same draws are made here with a restatement
of mulberry32 (the published 32-bit generator: add 0x6D2B79F5, two
xor-shift-multiply rounds), so the case list follows the reference's; the
expected bytes come from the oracle, so the checks do not depend on the
generator matching bit for bit.

  A  1000 cases: numInputs {4, 16, 64}, numRecovery {1, 4, 16}, blockSize
     {8, 64, 1024} bytes, firstInput [0, 100], firstRecovery [200, 300]
  B  the first 50 of them at numThreads {0, 1, 2, 4}: identical bytes
  C  mul_arr with n_coeff {2, 4, 8, 16, 32} x blockSize {64, 256, 1024}: cycling
     coefficients out[w] = in[w] * coeff[w % n_coeff]
  D  mul_arr with n_coeff {1, 4, 64} x {64, 1024, 4096} bytes x 30 trials

and, in the same way, the grids of test/par3-small-r-fused-parity.js,
par3-2d-decomposition-parity.js, par3-barycentric-parity.js,
par3-accumulate-binding-parity.js, par3-full-recovery-parity.js,
par3-input-parallel-parity.js, par3-gf64-mularr-parity.js and
par3-repair-parity.js.
"""
import numpy as np
import pytest

M32 = 0xFFFFFFFF


def mulberry32(seed):
    state = [seed & M32]

    def nxt():  # next uint32 (the reference's rng() * 2^32)
        s = (state[0] + 0x6D2B79F5) & M32
        state[0] = s
        t = ((s ^ (s >> 15)) * (1 | s)) & M32
        t = ((t + (((t ^ (t >> 7)) * (61 | t)) & M32)) & M32) ^ t
        return (t ^ (t >> 14)) & M32

    return nxt


def pick(nxt, pool):
    return pool[(nxt() * len(pool)) >> 32]


def words(nxt, n):
    out = np.empty(n, dtype=np.uint64)
    for w in range(n):
        hi = nxt()
        lo = nxt()
        out[w] = (hi << 32) | lo
    return out


def section_a_cases(count):
    nxt = mulberry32(0xDEADBEEF)
    cases = []
    for _ in range(count):
        ni = pick(nxt, [4, 16, 64])
        nr = pick(nxt, [1, 4, 16])
        bs = pick(nxt, [8, 64, 1024])
        fi = (nxt() * 101) >> 32
        fr = ((nxt() * 101) >> 32) + 200
        inp = words(nxt, ni * bs // 8).reshape(ni, bs // 8)
        cases.append((ni, nr, bs, fi, fr, inp))
    return cases


def test_mulberry32_draw_ranges():
    """The restated generator yields uint32 values and the case fields stay in
    the reference's pools (CPU only)."""
    cases = section_a_cases(200)
    assert {c[0] for c in cases} == {4, 16, 64}
    assert {c[1] for c in cases} == {1, 4, 16}
    assert {c[2] for c in cases} == {8, 64, 1024}
    assert all(0 <= c[3] <= 100 and 200 <= c[4] <= 300 for c in cases)


@pytest.mark.gpu
def test_section_a_compute_recovery(gpu, oracle):
    from parparpar_amd import binding
    bad = []
    for k, (ni, nr, bs, fi, fr, inp) in enumerate(section_a_cases(1000)):
        out = np.zeros((nr, bs // 8), dtype=np.uint64)
        binding.compute_recovery(inp, out, ni, nr, bs, fi, fr, 0)
        if not np.array_equal(out, oracle.recovery(inp, nr, fi, fr)):
            bad.append((k + 1, ni, nr, bs, fi, fr))
    assert not bad, bad[:5]


@pytest.mark.gpu
def test_section_b_thread_counts(gpu):
    from parparpar_amd import binding
    for (ni, nr, bs, fi, fr, inp) in section_a_cases(50):
        ref = np.zeros((nr, bs // 8), dtype=np.uint64)
        binding.compute_recovery(inp, ref, ni, nr, bs, fi, fr, 0)
        for nt in (0, 1, 2, 4):
            out = np.zeros_like(ref)
            binding.compute_recovery(inp, out, ni, nr, bs, fi, fr, nt)
            assert np.array_equal(out, ref), (ni, nr, bs, nt)


def _mul_arr_expected(oracle, inp, coeff):
    exp = np.zeros_like(inp)
    oracle.lib().or_gf64_region_mul_arr(oracle.p64(exp), oracle.p64(inp), oracle.p64(coeff), inp.size,
                                        coeff.size)
    return exp


@pytest.mark.gpu
def test_section_c_mul_arr_cycling(gpu, oracle):
    from parparpar_amd import binding
    nxt = mulberry32(0xFEEDFACE)
    for n_coeff in (2, 4, 8, 16, 32):
        for bs in (64, 256, 1024):
            nw = bs // 8
            inp = words(nxt, nw)
            coeff = words(nxt, n_coeff)
            out = np.zeros(nw, dtype=np.uint64)
            binding.Gf64Encoder().mul_arr(out, inp, coeff, nw, n_coeff)
            assert np.array_equal(out, _mul_arr_expected(oracle, inp, coeff)), (n_coeff, bs)


@pytest.mark.gpu
def test_section_d_mul_arr_trials(gpu, oracle):
    from parparpar_amd import binding
    nxt = mulberry32(0xBADDCAFE)
    enc = binding.Gf64Encoder()
    for n_coeff in (1, 4, 64):
        for nbytes in (64, 1024, 4096):
            nw = nbytes // 8
            for trial in range(30):
                inp = words(nxt, nw)
                coeff = words(nxt, n_coeff)
                out = np.zeros(nw, dtype=np.uint64)
                enc.mul_arr(out, inp, coeff, nw, n_coeff)
                assert np.array_equal(out, _mul_arr_expected(oracle, inp, coeff)), (n_coeff, nbytes, trial)


# test/par3-small-r-fused-parity.js:77-82 (compute_recovery_full at several
# thread counts, then compute_recovery_barycentric on fresh inputs)
SMALL_R = [(32, 1, 1024), (32, 2, 1024), (64, 4, 2048), (128, 8, 4096), (128, 16, 2048), (64, 32, 1024)]


@pytest.mark.gpu
def test_small_r_fused_configs(gpu, oracle):
    from parparpar_amd import binding
    nxt = mulberry32(0x5A110F05ED)
    for (n, r, b) in SMALL_R:
        inp = words(nxt, n * b // 8).reshape(n, b // 8)
        ref = np.zeros((r, b // 8), dtype=np.uint64)
        binding.compute_recovery_full(inp, ref, n, r, b, 0, n, 1)
        assert np.array_equal(ref, oracle.recovery(inp, r, 0, n)), (n, r, b)
        for threads in (2, 4, 8, 16):
            out = np.zeros_like(ref)
            binding.compute_recovery_full(inp, out, n, r, b, 0, n, threads)
            assert np.array_equal(out, ref), (n, r, b, threads)
    for (n, r, b) in SMALL_R:
        inp = words(nxt, n * b // 8).reshape(n, b // 8)
        full = np.zeros((r, b // 8), dtype=np.uint64)
        bary = np.zeros_like(full)
        binding.compute_recovery_full(inp, full, n, r, b, 0, n, 8)
        binding.compute_recovery_barycentric(inp, bary, n, r, b, 0, n, 8)
        assert np.array_equal(full, bary) and np.array_equal(bary, oracle.recovery(inp, r, 0, n))


def fill_pattern(nbytes):
    """test/par3-2d-decomposition-parity.js fillPattern: the uint32 at byte i is
    (i * 1103515245 + 12345) >>> 0, evaluated in IEEE doubles as JavaScript does."""
    i = np.arange(0, nbytes, 4, dtype=np.float64)
    v = (i * 1103515245.0 + 12345.0).astype(np.uint64) & np.uint64(M32)
    return v.astype(np.uint32).view(np.uint64)


# test/par3-2d-decomposition-parity.js:42-54: the reference's S_R x S_B
# decomposition at many thread counts, incl. R > 32 (K x G tiles)
DECOMP = [(8, 2, 128 << 10, (1, 2, 4, 8, 16, 32)), (16, 4, 256 << 10, (1, 2, 4, 8, 16, 32, 64)),
          (32, 8, 512 << 10, (1, 2, 8, 16, 32, 64)), (16, 2, 256 << 10, (3, 5, 6, 7, 12)),
          (4, 2, 1 << 20, (1, 4, 8, 16)), (16, 64, 256 << 10, (1, 2, 32, 64, 128))]


@pytest.mark.gpu
@pytest.mark.parametrize("n,r,b,threads", DECOMP)
def test_2d_decomposition_shapes(gpu, oracle, n, r, b, threads):
    from parparpar_amd import binding
    inp = fill_pattern(n * b).reshape(n, b // 8)
    want = oracle.recovery(inp, r, 0, n)
    for t in threads:
        out = np.zeros((r, b // 8), dtype=np.uint64)
        binding.compute_recovery_full(inp, out, n, r, b, 0, n, t)
        assert np.array_equal(out, want), t


@pytest.mark.gpu
def test_barycentric_grid_and_bit_flip(gpu, oracle):
    """test/par3-barycentric-parity.js:216-231 grid (N {16, 64, 256} x R {1, 4, 8}
    x blockSize {8, 64, 1024}) and its negative trap (:426-470): one flipped
    input bit changes the output, which still equals the oracle."""
    from parparpar_amd import binding
    nxt = mulberry32(0xBA7C)
    for n in (16, 64, 256):
        for r in (1, 4, 8):
            for b in (8, 64, 1024):
                inp = words(nxt, n * b // 8).reshape(n, b // 8)
                out = np.zeros((r, b // 8), dtype=np.uint64)
                binding.compute_recovery_barycentric(inp, out, n, r, b, 0, n + 1)
                assert np.array_equal(out, oracle.recovery(inp, r, 0, n + 1)), (n, r, b)
    n, r, b = 64, 4, 64
    inp = words(nxt, n * b // 8).reshape(n, b // 8)
    base = np.zeros((r, b // 8), dtype=np.uint64)
    binding.compute_recovery_barycentric(inp, base, n, r, b, 0, n + 1)
    inp[0, 0] ^= np.uint64(1 << 17)
    flipped = np.zeros_like(base)
    binding.compute_recovery_barycentric(inp, flipped, n, r, b, 0, n + 1)
    assert not np.array_equal(flipped, base)
    assert np.array_equal(flipped, oracle.recovery(inp, r, 0, n + 1))


# test/par3-accumulate-binding-parity.js:52-60: chunked accumulation over
# dirty outputs (first chunk writes, later chunks XOR), for the full and
# barycentric entries, across the R = 32 threshold
ACCUM = [(16, 2, 64, 2), (32, 4, 512, 4), (64, 8, 4096, 4), (128, 8, 16384, 8), (256, 16, 4096, 4),
         (64, 32, 1024, 2), (48, 34, 512, 3)]


@pytest.mark.gpu
@pytest.mark.parametrize("entry", ["full", "barycentric"])
def test_chunked_accumulation(gpu, oracle, entry):
    from parparpar_amd import binding
    fn = {"full": binding.compute_recovery_full, "barycentric": binding.compute_recovery_barycentric}[entry]
    nxt = mulberry32(0x110BADC0DE if entry == "full" else 0x110BADC0DF)
    for (ni, nr, bs, chunks) in ACCUM:
        inp = words(nxt, ni * bs // 8).reshape(ni, bs // 8)
        single = np.zeros((nr, bs // 8), dtype=np.uint64)
        fn(inp, single, ni, nr, bs, 0, ni, 0, False)
        assert np.array_equal(single, oracle.recovery(inp, nr, 0, ni)), (ni, nr, bs)
        out = np.full((nr, bs // 8), 0xAAAAAAAAAAAAAAAA, dtype=np.uint64)
        per = ni // chunks
        for c in range(chunks):
            s = c * per
            e = ni if c == chunks - 1 else (c + 1) * per
            sub = np.ascontiguousarray(inp[s:e])
            fn(sub, out, e - s, nr, bs, s, ni, 0, c > 0)
        assert np.array_equal(out, single), (entry, ni, nr, bs, chunks)


@pytest.mark.gpu
def test_accumulate_alias_and_trap(gpu):
    """:139-176: compute_recovery_accumulate on zeroed outputs equals
    compute_recovery_full; a flipped input bit is seen by the chunked path, and
    a chunk that overwrites instead of accumulating is caught."""
    from parparpar_amd import binding
    nxt = mulberry32(0x7AA9)
    t = words(nxt, 8 * 8).reshape(8, 8)
    a = np.zeros((2, 8), dtype=np.uint64)
    b = np.zeros((2, 8), dtype=np.uint64)
    binding.compute_recovery_full(t, a, 8, 2, 64, 0, 8, 0, False)
    binding.compute_recovery_accumulate(t, b, 8, 2, 64, 0, 8, 0)
    assert np.array_equal(a, b)
    clean = words(nxt, 16 * 8).reshape(16, 8)
    ref = np.zeros((4, 8), dtype=np.uint64)
    binding.compute_recovery_full(clean, ref, 16, 4, 64, 0, 16, 0, False)
    bad = clean.copy()
    bad.view(np.uint8).reshape(-1)[9 * 64 + 3] ^= 1
    single = np.zeros_like(ref)
    binding.compute_recovery_full(bad, single, 16, 4, 64, 0, 16, 0, False)
    chunked = np.zeros_like(ref)
    binding.compute_recovery_full(np.ascontiguousarray(bad[:8]), chunked, 8, 4, 64, 0, 16, 0, False)
    binding.compute_recovery_full(np.ascontiguousarray(bad[8:]), chunked, 8, 4, 64, 8, 16, 0, True)
    flawed = np.zeros_like(ref)
    binding.compute_recovery_full(np.ascontiguousarray(bad[:8]), flawed, 8, 4, 64, 0, 16, 0, False)
    binding.compute_recovery_full(np.ascontiguousarray(bad[8:]), flawed, 8, 4, 64, 8, 16, 0, False)
    assert not np.array_equal(ref, single)
    assert np.array_equal(chunked, single)
    assert not np.array_equal(chunked, flawed)


@pytest.mark.gpu
def test_full_recovery_random_configs(gpu, oracle):
    """test/par3-full-recovery-parity.js:127-160: 16 draws of numInputs {4, 16, 64},
    numRecovery {2, 8, 16}, blockSize {4096, 16384}, firstInput [0, 99],
    firstRecovery [200, 299]."""
    from parparpar_amd import binding
    nxt = mulberry32(0xC0FFEE42)
    for _ in range(16):
        ni = pick(nxt, [4, 16, 64])
        nr = pick(nxt, [2, 8, 16])
        bs = pick(nxt, [4096, 16384])
        fi = (nxt() * 100) >> 32
        fr = ((nxt() * 100) >> 32) + 200
        inp = words(nxt, ni * bs // 8).reshape(ni, bs // 8)
        out = np.zeros((nr, bs // 8), dtype=np.uint64)
        binding.compute_recovery_full(inp, out, ni, nr, bs, fi, fr, 0)
        assert np.array_equal(out, oracle.recovery(inp, nr, fi, fr)), (ni, nr, bs, fi, fr)


@pytest.mark.gpu
@pytest.mark.parametrize("N,n,bs", [(12, 3, 256), (40, 8, 1024), (64, 16, 512)])
def test_repair_native_matrix_flow(gpu, oracle, N, n, bs):
    """The repair flow of test/par3-repair-native-matrix.js:100-160 (and
    lib/par3gen.js:3840-4012) through the host entries: one
    build_coefficient_matrix (n rows x N columns), the A submatrix at the
    missing columns, RHS = recovery ^ sum over known inputs of c * D (here one
    compute_recovery_with_coeff over the known columns), solve_and_reconstruct;
    the solution is the lost inputs."""
    from parparpar_amd import binding
    rng = np.random.default_rng(N * n)
    fi, fr = 0, N
    data = rng.integers(0, 2 ** 64, size=(N, bs // 8), dtype=np.uint64)
    rec = np.zeros((n, bs // 8), dtype=np.uint64)
    binding.compute_recovery(data, rec, N, n, bs, fi, fr)
    missing = np.sort(rng.choice(N, n, replace=False))
    known = np.setdiff1d(np.arange(N), missing)
    M = binding.build_coefficient_matrix(N, n, fi, fr).reshape(n, N)
    assert np.array_equal(M, oracle.cauchy(N, n, fi, fr))
    A = np.ascontiguousarray(M[:, missing])
    ck = np.ascontiguousarray(M[:, known])
    rhs = rec.copy()
    binding.compute_recovery_with_coeff(ck, np.ascontiguousarray(data[known]), rhs, len(known), n, bs,
                                        accumulate=True)
    assert binding.solve_and_reconstruct(A, rhs, n, bs) is True
    assert np.array_equal(rhs, data[missing])


# test/par3-repair-parity.js:427-433: (totalInput, recoveryCount, blockSize,
# missingCounts), firstInput [0, 100], firstRecovery [200, 300]
REPAIR = [(8, 3, 64, [1, 2, 3]), (16, 4, 128, [1, 2, 3, 4]), (8, 2, 1024, [1, 2]),
          (32, 5, 256, [1, 2, 3, 4, 5]), (10, 3, 512, [1, 2, 3])]


@pytest.mark.gpu
def test_repair_parity_scenarios(gpu, oracle):
    """Every scenario and missing count of the reference's repair parity test,
    repaired twice: through the host entries (coefficient matrix, RHS by
    compute_recovery_with_coeff, solve_and_reconstruct), and on the device
    (parhip_d_gf64_repair); both give back the lost slices, and the host solve
    equals the oracle's Gauss-Jordan."""
    import torch
    from parparpar_amd import binding, repair
    nxt = mulberry32(0x4E9A1)
    for (N, R, bs, counts) in REPAIR:
        fi = (nxt() * 101) >> 32
        fr = ((nxt() * 101) >> 32) + 200
        data = words(nxt, N * bs // 8).reshape(N, bs // 8)
        rec = np.zeros((R, bs // 8), dtype=np.uint64)
        binding.compute_recovery(data, rec, N, R, bs, fi, fr)
        M = binding.build_coefficient_matrix(N, R, fi, fr).reshape(R, N)
        d_data = torch.from_numpy(data.view(np.int64)).cuda()
        d_rec = torch.from_numpy(rec.view(np.int64)).cuda()
        for k in counts:
            order = np.argsort(np.array([nxt() for _ in range(N)], dtype=np.uint64), kind="stable")
            missing = np.sort(order[:k])  # k distinct positions
            known = np.setdiff1d(np.arange(N), missing)
            A = np.ascontiguousarray(M[:k][:, missing])
            rhs = rec[:k].copy()
            binding.compute_recovery_with_coeff(np.ascontiguousarray(M[:k][:, known]),
                                                np.ascontiguousarray(data[known]), rhs, len(known), k, bs,
                                                accumulate=True)
            rc, A_o, x_o = oracle.solve_and_reconstruct(A.copy(), rhs.copy())
            assert rc == 0
            assert binding.solve_and_reconstruct(A, rhs, k, bs) is True
            assert np.array_equal(rhs, x_o) and np.array_equal(rhs, data[missing]), (N, R, bs, k)
            out = repair.repair(d_data, d_rec, missing, fi, fr)
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy().view(np.uint64), data[missing]), (N, R, bs, k)


@pytest.mark.gpu
def test_mul_arr_cells_and_guards(gpu, oracle):
    """test/par3-gf64-mularr-parity.js: n_coeff {2, 4, 8, 16} x len {8, 64, 256,
    1024} bytes x 10 trials of the cycling mul_arr, then its guards (:217-233):
    n_coeff = 0 with len > 0 and counts of 2^61 against 64-byte buffers are
    rejected on both entries; len = n_coeff = 0 stays a no-op."""
    from parparpar_amd import binding
    nxt = mulberry32(0x3A11)
    enc = binding.Gf64Encoder()
    for n_coeff in (2, 4, 8, 16):
        for nbytes in (8, 64, 256, 1024):
            for _ in range(10):
                inp = words(nxt, nbytes // 8)
                co = words(nxt, n_coeff)
                out = np.zeros_like(inp)
                enc.mul_arr(out, inp, co, nbytes // 8, n_coeff)
                assert np.array_equal(out, _mul_arr_expected(oracle, inp, co)), (n_coeff, nbytes)
    g = np.zeros(8, dtype=np.uint64)
    for fn in (binding.mul_arr, enc.mul_arr):
        for args in ((8, 0), (2 ** 61, 1), (1, 2 ** 61), (-1, 1)):
            with pytest.raises(ValueError):
                fn(g, g.copy(), g.copy(), *args)
        fn(g, g.copy(), g.copy(), 0, 0)
    with pytest.raises(ValueError):
        enc.muladd_arr(g, g.copy(), g.copy(), 9, 1)
    with pytest.raises(ValueError):
        enc.mul(g, g.copy(), 3, 9)
    with pytest.raises(ValueError):
        enc.coupled_muladd_arr(g, [g.copy()] * 2, g[:1].copy(), 8)  # coeff shorter than G
    with pytest.raises(ValueError):
        enc.fused_output_muladd_arr([g, np.zeros(4, dtype=np.uint64)], g.copy(), [1, 2], 8)
    with pytest.raises(ValueError):
        enc.two_d_muladd_arr([g], [g.copy(), g.copy()], np.ones(2, dtype=np.uint64), 1, 8)  # K_stride < G


@pytest.mark.gpu
def test_input_parallel_cases(gpu, oracle):
    """test/par3-input-parallel-parity.js:57-62: the shapes the reference splits
    over the input domain, at its thread counts; identical bytes, equal to the
    oracle."""
    from parparpar_amd import binding
    nxt = mulberry32(0x7800E3D)
    for (N, R, B, threads) in [(64, 2, 512, (1, 2, 4, 8, 16)), (256, 4, 1024, (1, 2, 8, 16)),
                               (512, 8, 4096, (1, 8, 12, 16)), (1024, 8, 8192, (1, 8, 16))]:
        inp = words(nxt, N * B // 8).reshape(N, B // 8)
        want = oracle.recovery(inp, R, 0, N)
        for t in threads:
            out = np.zeros((R, B // 8), dtype=np.uint64)
            binding.compute_recovery_full(inp, out, N, R, B, 0, N, t, False)
            assert np.array_equal(out, want), (N, R, B, t)
