"""The alternative contraction strategies against the oracle: GF(2) bit-matrix
on MFMA (fp4, int8) and 8-bit split tables.

The same contraction as the nibble-table kernel: as an integer matrix product
over bits whose parity is the GF(2^64) result, or with byte tables.  Bit-exact
against the oracle on small shapes and on sampled columns of a 1 GiB create,
and equal to the nibble-table strategy's output.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

METHODS = ["fp4", "i8", "tab8"]


def _method(name):
    import parparpar_amd as pp
    return {"fp4": pp.METHOD_MFMA, "i8": pp.METHOD_MFMA_I8, "tab8": pp.METHOD_TABLE8}[name]


@pytest.fixture
def use_method(gpu):
    import parparpar_amd as pp

    def set_(name):
        pp.set_method(_method(name))
    yield set_
    pp.set_method(pp.METHOD_AUTO)


def rand_words(rng, *shape):
    return rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64)


@pytest.mark.parametrize("m", METHODS)
def test_golden_vector(use_method, oracle, m):
    import os
    from parparpar_amd import binding
    use_method(m)
    n, b = 5, 1024
    inp = np.zeros((n, b // 8), dtype=np.uint64)
    for f in range(n):
        inp[f] = np.arange(b // 8, dtype=np.uint64) * 8 + f * b
    out = np.zeros((2, b // 8), dtype=np.uint64)
    binding.compute_recovery(inp, out, n, 2, b, 0, n, 0)
    golden = os.path.join(os.path.dirname(__file__), "golden", "par3-golden.bin")
    assert out.tobytes() == open(golden, "rb").read()


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
    (200, 64, 8, 0, 200, False),
    (1000, 32, 4, 0, 1000, True),
    (4, 8, 6, 2, 3, False),
]


@pytest.mark.parametrize("m", METHODS)
@pytest.mark.parametrize("n,w,r,fi,fr,acc", SHAPES)
def test_shapes(use_method, oracle, m, n, w, r, fi, fr, acc):
    from parparpar_amd import binding
    use_method(m)
    rng = np.random.default_rng(n * 7919 + w * 31 + r)
    inp = rand_words(rng, n, w)
    init = rand_words(rng, r, w)
    out = init.copy()
    binding.compute_recovery(inp, out, n, r, w * 8, fi, fr, 0, accumulate=acc)
    expect = init.copy() if acc else np.zeros((r, w), dtype=np.uint64)
    oracle.recovery_with_coeff(inp, oracle.cauchy(n, r, fi, fr), expect, acc)
    assert np.array_equal(out, expect)


@pytest.mark.parametrize("m", METHODS)
def test_config1_64x64k_r8(use_method, oracle, m):
    from parparpar_amd import binding
    use_method(m)
    rng = np.random.default_rng(0xC1)
    n, words, r = 64, 65536 // 8, 8
    inp = np.frombuffer(rng.bytes(n * 65536), dtype=np.uint64).reshape(n, words).copy()
    out = np.zeros((r, words), dtype=np.uint64)
    binding.compute_recovery(inp, out, n, r, 65536, 0, n)
    assert np.array_equal(out, oracle.recovery(inp, r, 0, n))


@pytest.mark.parametrize("m", METHODS)
def test_strided_view_and_ptr_fallback(use_method, oracle, m):
    import torch
    from parparpar_amd import device
    use_method(m)
    rng = np.random.default_rng(77)
    n, w, r = 13, 301, 5
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    base = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w + 3), dtype=torch.int64, device="cuda",
                         generator=g)
    view = base[:, 1:w + 1]
    out = torch.zeros((r, w), dtype=torch.int64, device="cuda")
    device.recovery(view, r, 0, n, out=out)
    host = np.ascontiguousarray(view.cpu().numpy().view(np.uint64))
    ref = oracle.recovery(host, r, 0, n)
    assert np.array_equal(out.cpu().numpy().view(np.uint64), ref)
    # pointer-array calls keep the table strategy under any method
    co = device.build_cauchy(n, r, 0, n)
    ins = [view[i] for i in range(n)]
    outs = [torch.zeros(w, dtype=torch.int64, device="cuda") for _ in range(r)]
    in_ptrs = torch.tensor([t.data_ptr() for t in ins], dtype=torch.int64, device="cuda")
    out_ptrs = torch.tensor([t.data_ptr() for t in outs], dtype=torch.int64, device="cuda")
    device.recovery_ptrs(in_ptrs, out_ptrs, w, co, accumulate=True)
    torch.cuda.synchronize()
    got = torch.stack(outs).cpu().numpy().view(np.uint64)
    assert np.array_equal(got, ref)
    del rng


@pytest.mark.parametrize("m", METHODS)
def test_full_size_create_columns(use_method, oracle, m):
    """1024 x 1 MiB at R=8: sampled columns exact, and equal to the table strategy."""
    import torch
    import parparpar_amd as pp
    from parparpar_amd import device
    n, w, r = 1024, 131072, 8
    print("mfma %s full-size create N=%d W=%d R=%d" % (m, n, w, r), flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(0xC2A)
    inp = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda",
                        generator=g)
    use_method(m)
    out = device.recovery(inp, r, 0, n)
    pp.set_method(pp.METHOD_TABLE)
    tab = device.recovery(inp, r, 0, n)
    torch.cuda.synchronize()
    assert torch.equal(out, tab)
    rng = np.random.default_rng(3)
    cols = np.unique(np.concatenate([[0, w - 1], rng.integers(0, w, size=10)])).astype(np.int64)
    idx = torch.as_tensor(cols, device="cuda")
    sample = np.ascontiguousarray(inp[:, idx].cpu().numpy().view(np.uint64))
    ref = oracle.recovery_with_coeff(sample, oracle.cauchy(n, r, 0, n))
    assert np.array_equal(out[:, idx].cpu().numpy().view(np.uint64), ref)
    del inp, out, tab
    torch.cuda.empty_cache()
