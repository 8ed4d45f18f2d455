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
- sorting of the exponents and their resize to numRec.
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
