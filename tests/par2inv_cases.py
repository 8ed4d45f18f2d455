"""PAR2 repair-matrix cases shared by the oracle and device tests.

The fixed cases of the reference's inversion test (test/gf16/test-inv.cpp:87-147)
and random cases shaped like its random section (test/gf16/test-inv.cpp:151-196).
"""
import numpy as np

import gf16_math as G

# (input validity, recovery exponents).  The reference test takes par2cmdline's
# elimination as truth; here the expected success is the field fact it
# computes (every leading pivot nonzero, no row exchanges), from gf16_math.
FLAW = [False] + [True] * 6553 + [False]        # inputs 0 and 6554 missing
REFERENCE_FIXED = [
    ([False], [0]),
    ([False], [1]),
    ([False], [65534]),
    ([False, True], [0]),                           # first block bad
    ([False, False, True, False], [0, 1, 2]),       # just enough recovery
    ([False] * 4, [0, 1, 5]),                       # insufficient recovery
    ([False] * 4, [1, 5, 8, 100]),
    (FLAW, [0, 5]),                                 # the PAR2 flaw
    (FLAW, [0, 6]),
    ([False] * 5, [0, 3, 5, 17, 65534]),
    ([False] * 6, [0, 1, 2, 3, 32768, 65534]),
    ([False] * 7, [0, 1, 2, 3, 4, 5, 6]),
    ([True] + [False] * 8, [0, 1, 2, 3, 5, 6, 7, 8]),
]
# What the reference test's comments state for the cases above.
REFERENCE_FIXED_OK = [True, True, True, True, True, False, True, False, True, True, True, True,
                      True]


def expected_ok(valid, exps):
    """Whether PAR2 repair with exactly the first numRec sorted exponents
    inverts: every leading pivot of their coefficients on the missing inputs is
    nonzero."""
    valid = np.asarray(valid, dtype=bool)
    midx = np.nonzero(~valid)[0]
    if len(exps) < midx.size:
        return False
    M = G.coeff_matrix(midx, sorted(exps)[:midx.size])
    return G.leading_pivots_nonzero(M)


def random_cases(seed, sizes, probs, rounds):
    """The reference's random section (test/gf16/test-inv.cpp:151-196) in shape:
    input counts, probabilities of an input being missing (an input is valid
    when a uniform draw exceeds it, as there), one recovery exponent
    per missing input drawn without replacement from 0..65534 and sorted.  The
    draws come from numpy, so the cases are not the reference's own."""
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        for _ in range(rounds(n)):
            for p in probs(n):
                valid = rng.random(n) > p
                missing = int((~valid).sum())
                if missing < 1:
                    continue
                exps = np.sort(rng.permutation(65535)[:missing]).astype(np.uint16)
                out.append((valid, exps))
    return out
