"""Pins of the oracle's PAR2 repair matrix (Galois16RecMatrix::Compute).

par2cmdline's inversion vendored in the reference tests (test/gf16/p2c-inv)
cannot be built here: galois.cpp includes libpar2internal.h, which is not
vendored.  The restatement is pinned instead by
- reconstruction: with the factors, sum(F[r][v] * D_v) + sum(F[r][j] * R_j)
  over the valid inputs v and the used recovery slices j must give back each
  dropped input exactly (the inverse is unique, so this fixes every factor);
- the PAR2 non-invertible case: inputs 0 and 1927 have bases whose ratio has
  multiplicative order 17, so with both missing, recovery exponents e and
  e + 17 make the leading 2 x 2 block singular; the second pivot is zero and
  that exponent is dropped (gfmat_inv.cpp:638-646), or Compute fails when no
  spare exponent is left;
- sorting of the exponents and their resize to numRec;
- the fixed cases of the reference's own inversion test
  (test/gf16/test-inv.cpp:87-147) and random cases of the same shape, whose
  success is the field fact par2cmdline computes there (every leading pivot
  nonzero) and whose factors must satisfy the repair identity, both computed
  by tests/gf16_math.py from the PAR2 specification without the oracle.
"""
import numpy as np
import pytest


def gf16_matvec(oracle, coeffs, slices):
    """XOR_k coeffs[k] * slices[k] over 16-bit LE words."""
    out = np.zeros(slices.shape[1] + 1, dtype=np.uint8)
    L = oracle.lib()
    for c, s in zip(coeffs, slices):
        b = np.ascontiguousarray(s, dtype=np.uint8)
        L.or_gf16_muladd_region(out.ctypes.data_as(oracle._u8p), oracle.p8(b), b.size, int(c))
    return out[:slices.shape[1]]


@pytest.mark.parametrize("n,missing,extra,seed", [(12, 3, 0, 1), (40, 7, 2, 2), (100, 1, 0, 3),
                                                  (65, 16, 5, 4)])
def test_reconstruction_round_trip(oracle, n, missing, extra, seed):
    rng = np.random.default_rng(seed)
    slice_size = 64
    data = rng.integers(0, 256, size=(n, slice_size), dtype=np.uint8)
    exps = rng.choice(np.arange(0, 300), size=missing + extra, replace=False).astype(np.uint16)
    rec_all = oracle.par2_create(list(data), np.arange(n), exps, slice_size)
    lost = rng.choice(n, size=missing, replace=False)
    valid = np.ones(n, dtype=bool)
    valid[lost] = False
    ok, used, F = oracle.gf16_rec_matrix(valid, exps)
    assert ok and len(used) == missing
    assert list(used) == sorted(used)
    by_exp = {int(e): rec_all[i] for i, e in enumerate(exps)}
    vidx = np.nonzero(valid)[0]
    midx = np.nonzero(~valid)[0]
    slices = np.concatenate([data[vidx], np.stack([by_exp[int(e)] for e in used])])
    for r in range(missing):
        got = gf16_matvec(oracle, F[r], slices)
        assert np.array_equal(got, data[midx[r]]), r


def test_bad_row_dropped_and_failure(oracle):
    n = 1928
    valid = np.ones(n, dtype=bool)
    valid[[0, 1927]] = False
    ok, used, F = oracle.gf16_rec_matrix(valid, [30, 22, 5])
    assert ok and list(used) == [5, 30]          # 22 = 5 + 17 dropped
    ok2, used2, F2 = oracle.gf16_rec_matrix(valid, [5, 30])
    assert ok2 and np.array_equal(F, F2)         # same as asking for the survivors
    ok3, used3, _ = oracle.gf16_rec_matrix(valid, [5, 22])
    assert not ok3                               # no spare exponent left


def test_not_enough_recovery(oracle):
    valid = np.ones(10, dtype=bool)
    valid[[1, 2, 3]] = False
    ok, used, _ = oracle.gf16_rec_matrix(valid, [0, 1])
    assert not ok


from par2inv_cases import (FLAW, REFERENCE_FIXED, REFERENCE_FIXED_OK, expected_ok,
                           random_cases)


@pytest.mark.parametrize("case", range(len(REFERENCE_FIXED)))
def test_reference_fixed_cases(oracle, case):
    import gf16_math as G
    valid, exps = REFERENCE_FIXED[case]
    want = expected_ok(valid, exps)
    assert want == REFERENCE_FIXED_OK[case]
    ok, used, F = oracle.gf16_rec_matrix(valid, exps)
    assert ok == want
    if ok:
        assert list(used) == sorted(exps)
        assert G.check_factors(valid, used, F)


def test_reference_flaw_is_singular():
    """Inputs 0 and 6554 have constants 2^1 and 2^13108: their ratio 2^13107 has
    order 65535 / 13107 = 5, so with both missing, exponents 0 and 5 give two
    proportional rows, a true rank deficiency, while 0 and 6 do not."""
    import gf16_math as G
    assert G.INPUT_LOGS[6554] - G.INPUT_LOGS[0] == 65535 // 5
    M = G.coeff_matrix([0, 6554], [0, 5])
    assert G.rank(M) == 1
    assert G.rank(G.coeff_matrix([0, 6554], [0, 6])) == 2


def test_reference_flaw_drops_an_exponent(oracle):
    """test/gf16/test-inv.cpp:103-129: with exponents 0, 5, 6 the flawed pair
    still inverts after dropping 0 or 5, and 6 is kept."""
    import gf16_math as G
    ok, used, F = oracle.gf16_rec_matrix(FLAW, [0, 5, 6])
    assert ok and len(used) == 2
    assert used[0] in (0, 5) and used[1] == 6
    assert G.check_factors(FLAW, used, F)


def test_reference_random_small(oracle):
    import gf16_math as G
    cases = random_cases(0x01020304, [2, 100], lambda n: [0.1, 0.5, 0.9], lambda n: 10)
    assert len(cases) > 40
    for valid, exps in cases:
        ok, used, F = oracle.gf16_rec_matrix(valid, exps)
        assert ok == expected_ok(valid, exps)
        if ok:
            rows = None if F.shape[0] <= 16 else [0, F.shape[0] // 2, F.shape[0] - 1]
            assert G.check_factors(valid, used, F, rows)


def test_factor_check_catches_a_wrong_factor(oracle):
    """A single wrong factor, in a valid-input column or a recovery column, or
    two rows swapped, fails the repair identity."""
    import gf16_math as G
    valid = np.ones(30, dtype=bool)
    valid[[2, 9, 17]] = False
    ok, used, F = oracle.gf16_rec_matrix(valid, [4, 11, 40])
    assert ok and G.check_factors(valid, used, F)
    for r, c in [(0, 0), (1, 26), (2, 29)]:
        bad = F.copy()
        bad[r, c] ^= 0x100
        assert not G.check_factors(valid, used, bad)
    assert not G.check_factors(valid, used, F[[1, 0, 2]])
    assert not G.check_factors(valid, used[::-1], F)


def test_reference_random_singular_at_scale(oracle):
    """A random case of 1234 inputs with 1121 missing whose exponents leave a
    zero leading pivot: the field check and the oracle both refuse it."""
    cases = random_cases(1234 * 7 + 3, [1234], lambda n: [0.1, 0.5, 0.9], lambda n: 2)
    valid, exps = cases[5]
    assert int((~valid).sum()) == 1121
    assert not expected_ok(valid, exps)
    ok, used, _ = oracle.gf16_rec_matrix(valid, exps)
    assert not ok
