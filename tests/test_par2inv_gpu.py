"""PAR2 repair matrix (Galois16RecMatrix::Compute) on the device vs the oracle.

Bit-exact factors and the same exponents kept, including the non-invertible
case where a recovery row is dropped; then a full PAR2 repair on the device
(create with the PAR2 kernels, drop inputs, rebuild them with the factors)
returning the dropped slices exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,missing,extra,seed", [
    (1, 1, 0, 1), (5, 2, 0, 2), (12, 3, 2, 3), (101, 1, 0, 4), (257, 16, 4, 5),
    (2000, 64, 3, 6), (999, 33, 0, 7)])
def test_matches_oracle(gpu, oracle, n, missing, extra, seed):
    from parparpar_amd import par2
    rng = np.random.default_rng(seed)
    valid = np.ones(n, dtype=bool)
    valid[rng.choice(n, size=missing, replace=False)] = False
    exps = rng.choice(np.arange(0, 1000), size=missing + extra, replace=False).astype(np.uint16)
    ok, used, F = par2.rec_matrix_compute(valid, exps)
    ok_o, used_o, F_o = oracle.gf16_rec_matrix(valid, exps)
    assert ok and ok_o
    assert np.array_equal(used, used_o)
    assert np.array_equal(F, F_o)


def test_non_invertible_row_dropped(gpu, oracle):
    from parparpar_amd import par2
    n = 1928
    valid = np.ones(n, dtype=bool)
    valid[[0, 1927]] = False   # bases whose ratio has order 17
    ok, used, F = par2.rec_matrix_compute(valid, [30, 22, 5])
    assert ok and list(used) == [5, 30]
    _, _, F_o = oracle.gf16_rec_matrix(valid, [30, 22, 5])
    assert np.array_equal(F, F_o)
    ok2, used2, _ = par2.rec_matrix_compute(valid, [5, 22])
    assert not ok2 and len(used2) == 1


def test_full_width(gpu, oracle):
    """32768 inputs (the PAR2 maximum), 48 missing."""
    from parparpar_amd import par2
    rng = np.random.default_rng(32768)
    n = 32768
    valid = np.ones(n, dtype=bool)
    valid[rng.choice(n, size=48, replace=False)] = False
    exps = np.arange(0, 50, dtype=np.uint16)
    ok, used, F = par2.rec_matrix_compute(valid, exps)
    ok_o, used_o, F_o = oracle.gf16_rec_matrix(valid, exps)
    assert ok and ok_o and np.array_equal(used, used_o) and np.array_equal(F, F_o)


def test_device_repair_round_trip(gpu):
    import torch
    from parparpar_amd import device, par2
    n, missing, words = 300, 12, 2048
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    data = torch.randint(0, 65536, (n, words), dtype=torch.int32, device="cuda",
                         generator=g).to(torch.int16)
    exps = np.arange(0, 20, dtype=np.uint16)
    coeff = device.gf16_coeffs(np.arange(n, dtype=np.uint16), exps)
    rec = device.gf16_muladd(data, coeff)                     # 20 recovery slices
    rng = np.random.default_rng(9)
    lost = np.sort(rng.choice(n, size=missing, replace=False))
    valid = np.ones(n, dtype=bool)
    valid[lost] = False
    ok, used, F = par2.rec_matrix_compute(valid, exps[::-1])  # unsorted on purpose
    assert ok and len(used) == missing
    src = torch.cat([data[torch.as_tensor(np.nonzero(valid)[0], device="cuda")],
                     rec[torch.as_tensor(used.astype(np.int64), device="cuda")]])
    fac = torch.from_numpy(F.view(np.int16)).cuda()
    out = device.gf16_muladd(src, fac)
    torch.cuda.synchronize()
    assert torch.equal(out, data[torch.as_tensor(lost, device="cuda")])


def test_arguments(gpu):
    import parparpar_amd as P
    from parparpar_amd import par2
    with pytest.raises(P.ParhipError):
        par2.rec_matrix_compute(np.ones(4, dtype=bool), [1, 2])       # nothing missing
    ok, used, _ = par2.rec_matrix_compute(np.array([1, 0, 0], dtype=bool), [7])
    assert not ok


@pytest.mark.parametrize("case", range(13))
def test_reference_fixed_cases(gpu, oracle, case):
    """The fixed cases of test/gf16/test-inv.cpp:87-147 on the device: same
    success, exponents and factors as the oracle."""
    from parparpar_amd import par2
    from par2inv_cases import REFERENCE_FIXED, REFERENCE_FIXED_OK
    valid, exps = REFERENCE_FIXED[case]
    ok, used, F = par2.rec_matrix_compute(valid, exps)
    ok_o, used_o, F_o = oracle.gf16_rec_matrix(valid, exps)
    assert ok == ok_o == REFERENCE_FIXED_OK[case]
    assert np.array_equal(used, used_o)
    if ok:
        assert np.array_equal(F, F_o)


def test_reference_flaw_drops_an_exponent(gpu, oracle):
    """test/gf16/test-inv.cpp:103-129: exponents 0, 5, 6 for the flawed pair."""
    from parparpar_amd import par2
    from par2inv_cases import FLAW
    ok, used, F = par2.rec_matrix_compute(FLAW, [6, 0, 5])
    assert ok and len(used) == 2 and used[0] in (0, 5) and used[1] == 6
    ok_o, used_o, F_o = oracle.gf16_rec_matrix(FLAW, [6, 0, 5])
    assert ok_o and np.array_equal(used, used_o) and np.array_equal(F, F_o)


@pytest.mark.parametrize("sizes,probs,rounds", [
    ([2, 100], [0.1, 0.5, 0.9], 10),
    ([1234], [0.1, 0.5, 0.9], 2),
    ([32768], [0.01], 1),
])
def test_reference_random(gpu, oracle, sizes, probs, rounds):
    """Random cases shaped like test/gf16/test-inv.cpp:151-196 (input counts 2,
    100, 1234 and 32768, one exponent per missing input): device and oracle
    agree bit for bit."""
    from parparpar_amd import par2
    from par2inv_cases import random_cases
    cases = random_cases(sizes[0] * 7 + len(probs), sizes, lambda n: probs, lambda n: rounds)
    assert cases
    for valid, exps in cases:
        print(f"n={valid.size} missing={int((~valid).sum())}", flush=True)
        ok, used, F = par2.rec_matrix_compute(valid, exps)
        ok_o, used_o, F_o = oracle.gf16_rec_matrix(valid, exps)
        assert ok == ok_o
        assert np.array_equal(used, used_o)
        if ok:
            assert np.array_equal(F, F_o)
