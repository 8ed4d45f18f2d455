"""Pins of the oracle (CPU restatement) before its results are trusted.

Each oracle function is checked against the reference's own golden vectors and
known answers, against an independent method (the inverse and the product are
unique, so any correct method must agree), against the reference engine built
in place under oracle/_ref when it exists, and against invariants chosen so
that a dropped term, a wrong index or a transposed operand fails.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "par3-golden.bin")


def golden_inputs():
    # test/par3-golden.js:72-95: 5 blocks of 1024 B, each u64 word holds its
    # global byte offset fileIndex*1024 + i (little endian).
    n, b = 5, 1024
    w = b // 8
    inp = np.zeros((n, w), dtype=np.uint64)
    for f in range(n):
        inp[f] = np.arange(w, dtype=np.uint64) * 8 + f * b
    return inp


def test_golden_par3_bin_crc(oracle):
    data = open(GOLDEN, "rb").read()
    assert len(data) == 2048
    assert "%08X" % oracle.crc32(data) == "6A941C15"  # test/par3-golden.js:35


def test_golden_recovery_matches_reference_vector(oracle):
    # R=2, firstInput=0, firstRecovery=N=5 (lib/par3gen.js:2071)
    rec = oracle.recovery(golden_inputs(), 2, 0, 5)
    assert rec.tobytes() == open(GOLDEN, "rb").read()


def test_golden_negative_trap(oracle):
    # a planted mistake (one flipped input bit) must change the golden bytes
    inp = golden_inputs()
    inp[3, 17] ^= np.uint64(1 << 40)
    assert oracle.recovery(inp, 2, 0, 5).tobytes() != open(GOLDEN, "rb").read()
    # so must a transposed recovery point (firstRecovery off by one)
    assert oracle.recovery(golden_inputs(), 2, 0, 6).tobytes() != open(GOLDEN, "rb").read()


def test_kat_native_fidelity(oracle):
    # test/par3-native-fidelity.js:240-244: compute_recovery([1,2], R=1, B=8, 0, 2)
    out = oracle.recovery(np.array([[1], [2]], dtype=np.uint64), 1, 0, 2)
    assert int(out[0, 0]) == 0x7ffffffffffffffa


def _peasant(a, b):
    """Textbook GF(2^64) product: shift-and-add over b's bits, reducing by
    x^64 = x^4 + x^3 + x + 1 at every overflow."""
    r = 0
    for _ in range(64):
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if a >> 64:
            a = (a & (2 ** 64 - 1)) ^ 0x1B
    return r


def _inverse(a):
    """a^(2^64 - 2) by square-and-multiply over _peasant."""
    r, e = 1, 2 ** 64 - 2
    while e:
        if e & 1:
            r = _peasant(r, a)
        a = _peasant(a, a)
        e >>= 1
    return r


def test_native_fidelity_cauchy_consistency(oracle):
    """test/par3-native-fidelity.js:302-360 (Section C): 3 inputs [1, 2, 3] x 2
    recovery, blockSize 8, points (0, 3); the matrix inv(c ^ (3 + r)) and the
    multiply-accumulate computed independently in Python."""
    inputs = [1, 2, 3]
    expected = []
    for r in range(2):
        acc = 0
        for c, v in enumerate(inputs):
            acc ^= _peasant(v, _inverse(c ^ (3 + r)))
        expected.append(acc)
    got = oracle.recovery(np.array([[v] for v in inputs], dtype=np.uint64), 2, 0, 3)
    assert [int(x) for x in got[:, 0]] == expected
    # the KAT of Section A through the same independent path
    assert _peasant(1, _inverse(0 ^ 2)) ^ _peasant(2, _inverse(1 ^ 2)) == 0x7ffffffffffffffa


def test_kat_mul(oracle):
    # test/gf64-js.test.js:286-296
    assert oracle.gf64_mul(2, 3) == 6
    assert oracle.gf64_mul(0x8000000000000000, 3) == 0x800000000000001B
    # x^63 * x = x^64 == x^4+x^3+x+1
    assert oracle.gf64_mul(1 << 63, 2) == 0x1B


def test_mul_vs_independent_peasant(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(11)
    a = rng.integers(0, 2 ** 64, size=2000, dtype=np.uint64)
    b = rng.integers(0, 2 ** 64, size=2000, dtype=np.uint64)
    for x, y in zip(a.tolist(), b.tolist()):
        assert L.or_gf64_mul(x, y) == L.or_gf64_mul_peasant(x, y)


def test_mul_field_axioms(oracle):
    m = oracle.gf64_mul
    rng = np.random.default_rng(12)
    v = rng.integers(0, 2 ** 64, size=(300, 3), dtype=np.uint64).tolist()
    for a, b, c in v:
        assert m(a, b) == m(b, a)
        assert m(a, b ^ c) == m(a, b) ^ m(a, c)
        assert m(m(a, b), c) == m(a, m(b, c))
        assert m(a, 1) == a and m(a, 0) == 0


def test_inverse(oracle):
    L = oracle.lib()
    assert oracle.gf64_inv(0) == 0 and oracle.gf64_inv(1) == 1  # gf64/test/test.c:112-120
    rng = np.random.default_rng(13)
    for a in rng.integers(2, 2 ** 64, size=300, dtype=np.uint64).tolist():
        inv = L.or_gf64_inv(a)
        assert L.or_gf64_inv_fermat(a) == inv
        assert oracle.gf64_mul(a, inv) == 1


def test_cauchy_definition(oracle):
    n, r, fi, fr = 7, 5, 3, 4  # overlapping ranges: exercise the 0 -> 1 rule
    m = oracle.cauchy(n, r, fi, fr)
    for row in range(r):
        for col in range(n):
            d = (fi + col) ^ (fr + row)
            expect = 1 if d == 0 else oracle.gf64_inv(d)
            if d == 0:
                expect = oracle.gf64_inv(1)
            assert int(m[row, col]) == expect
            if d:
                assert oracle.gf64_mul(int(m[row, col]), d) == 1


def test_recovery_linearity_and_accumulate(oracle):
    rng = np.random.default_rng(14)
    n, w, r = 9, 33, 4
    a = rng.integers(0, 2 ** 64, size=(n, w), dtype=np.uint64)
    b = rng.integers(0, 2 ** 64, size=(n, w), dtype=np.uint64)
    co = oracle.cauchy(n, r, 0, n)
    ra = oracle.recovery_with_coeff(a, co)
    rb = oracle.recovery_with_coeff(b, co)
    assert np.array_equal(oracle.recovery_with_coeff(a ^ b, co), ra ^ rb)
    # accumulate=true XORs into the outputs
    acc = rb.copy()
    oracle.recovery_with_coeff(a, co, recovery=acc, accumulate=True)
    assert np.array_equal(acc, ra ^ rb)
    # chunked over inputs (lib/par3gen.js:2306-2316) == single call
    acc = np.zeros_like(ra)
    oracle.recovery_with_coeff(np.ascontiguousarray(a[:4]), np.ascontiguousarray(co[:, :4]), acc, True)
    oracle.recovery_with_coeff(np.ascontiguousarray(a[4:]), np.ascontiguousarray(co[:, 4:]), acc, True)
    assert np.array_equal(acc, ra)


def test_columns_equal_full(oracle):
    rng = np.random.default_rng(15)
    n, w, r = 6, 50, 3
    a = rng.integers(0, 2 ** 64, size=(n, w), dtype=np.uint64)
    co = oracle.cauchy(n, r, 0, n)
    full = oracle.recovery_with_coeff(a, co)
    cols = np.array([0, 7, 49, 3], dtype=np.uint64)
    assert np.array_equal(oracle.recovery_columns(a, co, cols), full[:, cols.astype(np.int64)])


def test_solve_round_trip(oracle):
    # encode, erase k inputs, rebuild them from k recovery blocks (the repair
    # path, lib/par3gen.js:3840-4012): the solution must equal the originals
    rng = np.random.default_rng(16)
    n, w, r = 12, 40, 5
    data = rng.integers(0, 2 ** 64, size=(n, w), dtype=np.uint64)
    co = oracle.cauchy(n, r, 0, n)
    rec = oracle.recovery_with_coeff(data, co)
    missing = [1, 4, 9]
    known = [i for i in range(n) if i not in missing]
    k = len(missing)
    A = np.ascontiguousarray(co[:k][:, missing])
    rhs = rec[:k].copy()
    for eq in range(k):
        for j in known:
            for wi in range(w):
                rhs[eq, wi] ^= np.uint64(oracle.gf64_mul(int(co[eq, j]), int(data[j, wi])))
    rc, A2, sol = oracle.solve_and_reconstruct(A, rhs)
    assert rc == 0
    assert np.array_equal(sol, data[missing])
    assert np.array_equal(A2, np.eye(k, dtype=np.uint64))
    # singular matrix -> -1 (src/par3_engine.cc:995)
    S = np.array([[1, 2], [1, 2]], dtype=np.uint64)
    rc, _, _ = oracle.solve_and_reconstruct(S, np.zeros((2, 3), dtype=np.uint64))
    assert rc == -1


def test_region_semantics(oracle):
    L = oracle.lib()
    P = oracle.p64
    rng = np.random.default_rng(17)
    ln = 37
    x = rng.integers(0, 2 ** 64, size=ln, dtype=np.uint64)
    co = rng.integers(0, 2 ** 64, size=5, dtype=np.uint64)
    out = np.zeros(ln, dtype=np.uint64)
    L.or_gf64_region_mul_arr(P(out), P(x), P(co), ln, 3)  # cycling
    for i in range(ln):
        assert int(out[i]) == oracle.gf64_mul(int(x[i]), int(co[i % 3]))
    out2 = rng.integers(0, 2 ** 64, size=ln, dtype=np.uint64)
    keep = out2.copy()
    L.or_gf64_region_muladd_arr(P(out2), P(x), P(co), ln, 5)  # sum over all coefficients
    s = 0
    for c in co.tolist():
        s ^= c
    for i in range(ln):
        assert int(out2[i]) == int(keep[i]) ^ oracle.gf64_mul(int(x[i]), s)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "libref_par3.so")),
                    reason="reference engine not built here (oracle/_ref)")
def test_oracle_equals_reference_engine(oracle):
    R = oracle.ref()
    L = oracle.lib()
    rng = np.random.default_rng(18)
    for a, b in rng.integers(0, 2 ** 64, size=(500, 2), dtype=np.uint64).tolist():
        assert R.ref_gf64_mul(a, b) == L.or_gf64_mul(a, b)
        assert R.ref_gf64_inv(a) == L.or_gf64_inv(a)
    for (n, w, r, fi, fr, acc) in [(64, 512, 8, 0, 64, 0), (13, 129, 3, 7, 200, 1), (40, 8, 33, 0, 40, 0)]:
        inp = rng.integers(0, 2 ** 64, size=(n, w), dtype=np.uint64)
        init = rng.integers(0, 2 ** 64, size=(r, w), dtype=np.uint64)
        mine = init.copy() if acc else np.zeros((r, w), dtype=np.uint64)
        co = oracle.cauchy(n, r, fi, fr)
        oracle.recovery_with_coeff(inp, co, mine, bool(acc))
        theirs = init.copy() if acc else np.zeros((r, w), dtype=np.uint64)
        R.ref_compute_recovery(oracle.p64(inp), n, oracle.p64(theirs), r, w, fi, fr, 4, acc)
        assert np.array_equal(mine, theirs)
        rco = np.zeros((r, n), dtype=np.uint64)
        R.ref_build_cauchy(oracle.p64(rco), n, r, fi, fr)
        assert np.array_equal(rco, co)
    # solve: identical in-place results
    k, w = 6, 16
    A = oracle.cauchy(k, k, 0, 100)
    rhs = rng.integers(0, 2 ** 64, size=(k, w), dtype=np.uint64)
    rc, A_mine, x_mine = oracle.solve_and_reconstruct(A, rhs)
    A_ref, x_ref = A.copy(), rhs.copy()
    assert R.ref_solve_and_reconstruct(oracle.p64(A_ref), oracle.p64(x_ref), k, w, 1) == rc == 0
    assert np.array_equal(A_mine, A_ref) and np.array_equal(x_mine, x_ref)
    # PAR2 coefficients
    for i in range(0, 32768, 331):
        assert R.gfmat_input_log(i) == L.or_gf16_input_log(i)
        for e in (0, 1, 2, 255, 511, 4097, 65534):
            assert R.gfmat_coeff(i, e) == L.or_gf16_coeff(i, e)


# ------------------------------------------------------------------ GF(2^16)

def test_gf16_mul_vs_peasant(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(19)
    for a, b in rng.integers(0, 65536, size=(3000, 2)).tolist():
        assert L.or_gf16_mul(a, b) == L.or_gf16_mul_peasant(a, b)
    for a in range(0, 65536, 7):
        assert L.or_gf16_mul(a, 1) == a
        assert L.or_gf16_mul(a, 2) == L.or_gf16_mul_peasant(a, 2)


def test_gf16_generator_order(oracle):
    L = oracle.lib()
    seen = set()
    for e in range(65535):
        seen.add(L.or_gf16_exp(e))
    assert len(seen) == 65535 and 0 not in seen


def test_gf16_input_log_sequence(oracle):
    L = oracle.lib()
    assert [L.or_gf16_input_log(i) for i in range(12)] == [1, 2, 4, 7, 8, 11, 13, 14, 16, 19, 22, 23]
    prev = 0
    for i in range(0, 32768):
        v = L.or_gf16_input_log(i)
        assert v > prev
        assert v % 3 and v % 5 and v % 17 and v % 257
        prev = v


def test_gf16_coeff(oracle):
    L = oracle.lib()
    for i in range(0, 32768, 1021):
        assert L.or_gf16_coeff(i, 0) == 1
        base = L.or_gf16_exp(L.or_gf16_input_log(i))
        acc = 1
        for e in range(1, 20):
            acc = L.or_gf16_mul(acc, base)
            assert L.or_gf16_coeff(i, e) == acc


def test_par2_odd_tail(oracle):
    rng = np.random.default_rng(20)
    s = 10
    full = rng.integers(0, 256, size=s, dtype=np.uint8)
    odd = full[:7].copy()
    out = oracle.par2_create([full, odd], [3, 9], [0, 5], s)
    # exponent 0: all coefficients 1 -> plain XOR of the (zero-padded) inputs
    pad = np.zeros(s, dtype=np.uint8)
    pad[:7] = odd
    assert np.array_equal(out[0], full ^ pad)
    # exponent 5 at the odd tail word: product of the zero-extended byte
    c = oracle.gf16_coeff(9, 5)
    t = oracle.lib().or_gf16_mul(int(odd[6]), c)
    c0 = oracle.gf16_coeff(3, 5)
    h = oracle.lib().or_gf16_mul(int(full[6]) | (int(full[7]) << 8), c0)
    assert int(out[1][6]) | (int(out[1][7]) << 8) == t ^ h


@pytest.mark.parametrize("n,words,r,threads", [(5, 64, 3, 2), (17, 1000, 9, 4), (40, 16384, 16, 8),
                                               (3, 31, 2, 1)])
def test_cpu_baseline_par2_equals_oracle(oracle, n, words, r, threads):
    """The PAR2 CPU baseline (split-nibble shuffles, threads over outputs; the
    bench's cpu_baseline for PAR2) gives the oracle's create bytes, including a
    slice whose words are not a multiple of 32 (scalar tail)."""
    rng = np.random.default_rng(n * 1000 + words)
    x = np.frombuffer(rng.bytes(n * words * 2), dtype=np.uint16).reshape(n, words).copy()
    inums = rng.choice(32768, n, replace=False).astype(np.uint16)
    exps = np.sort(rng.choice(65535, r, replace=False)).astype(np.uint16)
    exps[0] = 0  # exponent 0: plain XOR parity
    got, used = oracle.baseline_par2_create(x, inums, exps, threads)
    assert used == min(threads, r)
    ref = oracle.par2_create([row.view(np.uint8) for row in x], inums, exps, words * 2)
    assert np.array_equal(got.view(np.uint8), ref)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "libref_par3.so")),
                    reason="reference engine not built here (oracle/_ref)")
def test_region_ops_equal_reference_scalar(oracle):
    """The six region operations against the reference's own scalar versions
    (gf64/gf64_region_scalar.c, compiled into oracle/_ref), byte for byte, on
    random words, coefficient cycling (n_coeff < len) and several blocks."""
    import ctypes
    L, R = oracle.lib(), oracle.ref()
    P = oracle.p64
    rng = np.random.default_rng(0x5CA1)
    u64p = ctypes.POINTER(ctypes.c_uint64)

    def rnd(*shape):
        return rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64)

    def ptrs(rows):
        return (u64p * len(rows))(*[P(r) for r in rows])

    ln = 61
    x = rnd(ln)
    co = rnd(7)
    for fo, fr_ in (("or_gf64_region_mul_arr", "gf64_region_mul_scalar_arr"),
                    ("or_gf64_region_muladd_arr", "gf64_region_muladd_scalar_arr")):
        for nc in (1, 3, 7):
            base = rnd(ln)
            a, b = base.copy(), base.copy()
            getattr(L, fo)(P(a), P(x), P(co), ln, nc)
            getattr(R, fr_)(P(b), P(x), P(co), ctypes.c_size_t(ln), ctypes.c_size_t(nc))
            assert np.array_equal(a, b), (fo, nc)
    a, b = rnd(ln), np.zeros(ln, dtype=np.uint64)
    L.or_gf64_region_mul(P(a), P(x), ln, int(co[0]))
    R.gf64_region_mul_scalar(P(b), P(x), ctypes.c_size_t(ln), ctypes.c_uint64(int(co[0])))
    assert np.array_equal(a, b)
    # coupled: G input blocks, one coefficient each
    G, K = 5, 4
    blocks = [rnd(ln) for _ in range(G)]
    cb = rnd(G)
    base = rnd(ln)
    a, b = base.copy(), base.copy()
    L.or_gf64_region_coupled_muladd_arr(P(a), ptrs(blocks), P(cb), ln, G)
    R.gf64_region_coupled_muladd_scalar_arr(P(b), ptrs(blocks), P(cb), ctypes.c_size_t(ln), ctypes.c_size_t(G))
    assert np.array_equal(a, b)
    # fused output: one input, K outputs, coefficient per output
    cs = [rnd(1) for _ in range(K)]
    outs_a = [rnd(ln) for _ in range(K)]
    outs_b = [o.copy() for o in outs_a]
    L.or_gf64_region_fused_output_muladd_arr(ptrs(outs_a), P(x), ptrs(cs), ln, K)
    R.gf64_region_fused_output_muladd_scalar_arr(ptrs(outs_b), P(x), ptrs(cs), ctypes.c_size_t(ln),
                                                 ctypes.c_size_t(K))
    assert all(np.array_equal(p, q) for p, q in zip(outs_a, outs_b))
    # 2D: K outputs x G input blocks, coefficient matrix with a row stride
    stride = G + 3
    c2 = rnd(K * stride)
    outs_a = [rnd(ln) for _ in range(K)]
    outs_b = [o.copy() for o in outs_a]
    L.or_gf64_region_2d_muladd_arr(ptrs(outs_a), K, ptrs(blocks), G, P(c2), stride, ln)
    R.gf64_region_2d_muladd_scalar_arr(ptrs(outs_b), ctypes.c_size_t(K), ptrs(blocks), ctypes.c_size_t(G),
                                       P(c2), ctypes.c_size_t(stride), ctypes.c_size_t(ln))
    assert all(np.array_equal(p, q) for p, q in zip(outs_a, outs_b))


def test_par2_columns_equal_create(oracle):
    """The column evaluator the GPU tests use at full size equals the full PAR2
    create on every column of a small case (odd slice size included)."""
    rng = np.random.default_rng(0x7C01)
    for n, s, r in ((5, 34, 3), (9, 1001, 4)):
        ins = rng.integers(0, 256, size=(n, s + (s & 1)), dtype=np.uint8)
        if s & 1:
            ins[:, s] = 0  # the odd tail byte's partner is zero
        inums = rng.choice(32768, n, replace=False).astype(np.uint16)
        exps = rng.choice(65535, r, replace=False).astype(np.uint16)
        full = oracle.par2_create([ins[i, :s] for i in range(n)], inums.tolist(), exps.tolist(), s)
        words = (s + 1) // 2
        cols = oracle.par2_columns(np.ascontiguousarray(ins), ins.shape[1], inums, exps,
                                   np.arange(words, dtype=np.uint64))
        # (an odd slice's last word: its low byte is the slice's last byte)
        assert np.array_equal(np.ascontiguousarray(cols).view(np.uint8)[:, :s], full)
