"""PAR2 GF(2^16) contraction as a GF(2) bit-matrix on MFMA (fp4 and int8), against
the oracle and the split-table kernel.

The same PAR2 muladd as k_gf16_tab (gf16/controller_cpu.cpp:463-537), written
as an integer matrix product over bits whose parity is the result.  Bit-exact
on ragged small shapes (odd input counts, words that do not fill a tile,
outputs that do not fill an output tile), with accumulate, through split-K
partials, and on sampled columns of the 1 GiB x 512 PAR2 create.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

METHODS = ["fp4", "i8"]


@pytest.fixture
def use_method(gpu):
    import parparpar_amd as pp

    def set_(name):
        pp.set_method(pp.METHOD_MFMA if name == "fp4" else pp.METHOD_MFMA_I8)
    yield set_
    pp.check(pp.lib().parhip_tune_override(0, 0, 0, 0))
    pp.set_method(pp.METHOD_AUTO)


def _case(torch, n, w, r, seed):
    from parparpar_amd import device
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randint(-32768, 32767, (n, w), dtype=torch.int16, device="cuda", generator=g)
    rng = np.random.default_rng(seed)
    inums = rng.choice(32768, n, replace=False).astype(np.uint16)
    exps = rng.choice(65535, r, replace=False).astype(np.uint16)
    co = device.gf16_coeffs(inums, exps)
    return x, inums, exps, co


@pytest.mark.parametrize("m", METHODS)
@pytest.mark.parametrize("n,w,r,acc", [(1, 1, 1, False), (2, 3, 2, True), (5, 33, 7, False),
                                       (16, 1000, 9, True), (37, 2049, 17, False), (64, 512, 64, True),
                                       (130, 700, 40, False), (37, 256, 17, False), (203, 1024, 45, True)])
def test_small_shapes(use_method, oracle, m, n, w, r, acc):
    import torch
    from parparpar_amd import device
    x, inums, exps, co = _case(torch, n, w, r, n * 1000 + w + r)
    init = torch.randint(-32768, 32767, (r, w), dtype=torch.int16, device="cuda")
    out = init.clone()
    use_method(m)
    device.gf16_muladd(x, co, out=out, accumulate=acc)
    torch.cuda.synchronize()
    samp = np.ascontiguousarray(x.cpu().numpy().view(np.uint8))
    ref = oracle.par2_columns(samp, 2 * w, inums, exps, np.arange(w, dtype=np.uint64))
    if acc:
        ref = ref ^ init.cpu().numpy().view(np.uint16)
    assert np.array_equal(out.cpu().numpy().view(np.uint16), ref)


@pytest.mark.parametrize("m", METHODS)
def test_split_k_and_tile_shapes(use_method, m):
    """Forced split-K group counts and every output-tile / column-block variant
    give the split-table kernel's bytes."""
    import torch
    import parparpar_amd as pp
    from parparpar_amd import device
    x, inums, exps, co = _case(torch, 203, 3000, 45, 7)
    ref = device.gf16_muladd(x, co)  # split tables (AUTO)
    torch.cuda.synchronize()
    use_method(m)
    for (ot, nb, groups) in [(0, 0, 1), (0, 0, 3), (0, 0, 17), (4, 4, 2), (4, 2, 0), (8, 2, 5)]:
        pp.check(pp.lib().parhip_tune_override(ot, nb, 0, groups))
        out = device.gf16_muladd(x, co)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (m, ot, nb, groups)
    pp.check(pp.lib().parhip_tune_override(0, 0, 0, 0))


def test_fp4_full_tile_kernel_shapes(use_method):
    """The full-tile fp4 kernel (slices of a multiple of 256 words): stages of
    8 and 16 inputs, split-K rows, a ragged last stage (203 inputs), output
    tiles past R (45), against the split-table kernel."""
    import torch
    import parparpar_amd as pp
    from parparpar_amd import device
    x, inums, exps, co = _case(torch, 203, 2048, 45, 11)
    ref = device.gf16_muladd(x, co)  # split tables (AUTO)
    torch.cuda.synchronize()
    use_method("fp4")
    for (ib, groups) in [(0, 1), (0, 3), (4, 1), (4, 5), (2, 17)]:
        pp.check(pp.lib().parhip_tune_override(0, 0, ib, groups))
        out = device.gf16_muladd(x, co)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (ib, groups)
    pp.check(pp.lib().parhip_tune_override(0, 0, 0, 0))


@pytest.mark.parametrize("m", METHODS)
def test_full_size_columns(use_method, oracle, m):
    """1 GiB = 32768 x 32 KiB, exponents 0..511: sampled columns against the oracle."""
    import torch
    from parparpar_amd import device
    n, words, r = 32768, 16384, 512
    print("gf16 mfma %s N=%d words=%d R=%d" % (m, n, words, r), flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.randint(-32768, 32767, (n, words), dtype=torch.int16, device="cuda", generator=g)
    inums = np.arange(n, dtype=np.uint16)
    exps = np.arange(r, dtype=np.uint16)
    co = device.gf16_coeffs(inums, exps)
    use_method(m)
    out = device.gf16_muladd(x, co)
    torch.cuda.synchronize()
    cols = np.array([0, 1, 4095, 8192, words - 1], dtype=np.int64)
    tc = torch.as_tensor(cols, device="cuda")
    samp = np.ascontiguousarray(x[:, tc].cpu().numpy().view(np.uint8))
    ref = oracle.par2_columns(samp, 2 * len(cols), inums, exps, np.arange(len(cols), dtype=np.uint64))
    assert np.array_equal(out[:, tc].cpu().numpy().view(np.uint16), ref)
    del x, out
    torch.cuda.empty_cache()


def test_full_size_whole_output_tables_fp4_i8(gpu, oracle):
    """1 GiB = 32768 x 32 KiB at 512 recovery slices: the whole outputs of the
    split-table kernel and of the bit-matrix kernel on fp4 and on int8 MFMA
    agree byte for byte, and sampled columns equal the oracle."""
    import torch
    import parparpar_amd as pp
    from parparpar_amd import device
    n, words, r = 32768, 16384, 512
    print("gf16 whole output tables/fp4/i8 N=%d words=%d R=%d" % (n, words, r), flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.randint(-32768, 32767, (n, words), dtype=torch.int16, device="cuda", generator=g)
    inums = np.arange(n, dtype=np.uint16)
    exps = np.arange(r, dtype=np.uint16)
    co = device.gf16_coeffs(inums, exps)
    outs = {}
    try:
        for name, m in (("tab", pp.METHOD_TABLE), ("fp4", pp.METHOD_MFMA), ("i8", pp.METHOD_MFMA_I8)):
            pp.set_method(m)
            outs[name] = device.gf16_muladd(x, co)
    finally:
        pp.set_method(pp.METHOD_AUTO)
    torch.cuda.synchronize()
    assert torch.equal(outs["tab"], outs["fp4"])
    assert torch.equal(outs["tab"], outs["i8"])
    cols = np.array([0, 7, 8191, words - 1], dtype=np.int64)
    tc = torch.as_tensor(cols, device="cuda")
    samp = np.ascontiguousarray(x[:, tc].cpu().numpy().view(np.uint8))
    ref = oracle.par2_columns(samp, 2 * len(cols), inums, exps, np.arange(len(cols), dtype=np.uint64))
    assert np.array_equal(outs["tab"][:, tc].cpu().numpy().view(np.uint16), ref)
    del x, outs
    torch.cuda.empty_cache()
