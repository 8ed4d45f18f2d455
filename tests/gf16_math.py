"""GF(2^16) arithmetic for the PAR2 checks, written from the PAR2 specification.

Independent of the oracle and of the product: field GF(2)[x] / 0x1100B with
generator 2; input i's constant is 2^l_i where l_0 < l_1 < ... are the
integers >= 1 coprime to 65535 = 3 * 5 * 17 * 257; recovery slice e of input i
carries coefficient (2^l_i)^e.  Everything is numpy over uint16 arrays.
"""
import math

import numpy as np

POLY = 0x1100B
ORDER = 65535

_exp = np.zeros(2 * ORDER, dtype=np.uint16)
_log = np.zeros(65536, dtype=np.int64)
_v = 1
for _e in range(ORDER):
    _exp[_e] = _v
    _log[_v] = _e
    _v <<= 1
    if _v & 0x10000:
        _v ^= POLY
_exp[ORDER:] = _exp[:ORDER]
INPUT_LOGS = np.array([n for n in range(1, ORDER) if math.gcd(n, ORDER) == 1][:32768],
                      dtype=np.int64)


def mul(a, b):
    a = np.asarray(a, dtype=np.uint16)
    b = np.asarray(b, dtype=np.uint16)
    r = _exp[(_log[a] + _log[b]) % ORDER]
    return np.where((a == 0) | (b == 0), np.uint16(0), r).astype(np.uint16)


def inv(a):
    a = np.asarray(a, dtype=np.uint16)
    assert np.all(a != 0)
    return _exp[(ORDER - _log[a]) % ORDER]


def coeff_matrix(inputs, exponents):
    """coeff[r, c] = (2^l_inputs[c])^exponents[r]."""
    li = INPUT_LOGS[np.asarray(inputs, dtype=np.int64)]
    e = np.asarray(exponents, dtype=np.int64)
    return _exp[(li[None, :] * e[:, None]) % ORDER].astype(np.uint16)


def matmul(A, B):
    """A @ B over the field (rows of A at a time)."""
    A = np.asarray(A, dtype=np.uint16)
    B = np.asarray(B, dtype=np.uint16)
    out = np.zeros((A.shape[0], B.shape[1]), dtype=np.uint16)
    for r in range(A.shape[0]):
        out[r] = np.bitwise_xor.reduce(mul(A[r][:, None], B), axis=0)
    return out


def leading_pivots_nonzero(M):
    """True when every leading principal minor of square M is nonzero, i.e.
    LU without row exchanges runs to the end.  PAR2 repair (par2cmdline's
    elimination and the reference's Galois16RecMatrix) exchanges no rows, so
    this is when an exact-count repair matrix inverts."""
    M = np.array(M, dtype=np.uint16)
    n = M.shape[0]
    for k in range(n):
        if M[k, k] == 0:
            return False
        f = mul(M[k + 1:, k], inv(M[k, k]))
        M[k + 1:, k:] ^= mul(f[:, None], M[k, k:][None, :])
    return True


def rank(M):
    """Rank with row exchanges (plain Gaussian elimination)."""
    M = np.array(M, dtype=np.uint16)
    rows, cols = M.shape
    r = 0
    for c in range(cols):
        nz = np.nonzero(M[r:, c])[0]
        if nz.size == 0:
            continue
        p = r + nz[0]
        M[[r, p]] = M[[p, r]]
        f = mul(M[r + 1:, c], inv(M[r, c]))
        M[r + 1:, :] ^= mul(f[:, None], M[r][None, :])
        r += 1
        if r == rows:
            break
    return r


def check_factors(valid, used, F, rows=None):
    """The repair factors solve PAR2 repair: with G = F[:, validCount:] (the
    factors of the used recovery slices) and C the coefficients of the used
    exponents, G @ C must be the identity on the missing inputs' columns and
    must equal F's valid-input factors on the valid ones (each input in input
    order).  The inverse is unique, so this fixes every factor.  `rows`
    restricts the check to some rows of F."""
    valid = np.asarray(valid, dtype=bool)
    n = valid.size
    vidx = np.nonzero(valid)[0]
    midx = np.nonzero(~valid)[0]
    vc = vidx.size
    C = coeff_matrix(np.arange(n), used)
    sel = np.arange(F.shape[0]) if rows is None else np.asarray(rows)
    P = matmul(F[sel][:, vc:], C)
    eye = np.zeros((sel.size, midx.size), dtype=np.uint16)
    eye[np.arange(sel.size), sel] = 1
    return np.array_equal(P[:, midx], eye) and np.array_equal(P[:, vidx], F[sel][:, :vc])
