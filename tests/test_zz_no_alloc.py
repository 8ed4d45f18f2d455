"""parhip_d_gf64_recovery_static: the plain create with its scratch memory in a
fixed device array.  Its bytes must equal parhip_d_gf64_recovery's (the
allocating entry point) call after call, including a second, smaller call that
finds the first call's contents in the array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _words(seed, n, w):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)


def _both(n, w, r, fi, fr, seed, accumulate=False):
    """(allocating entry's output, static entry's output) for the same call."""
    import torch
    import parparpar_amd as P
    L = P.lib()
    x = _words(seed, n, w)
    init = _words(seed + 1, r, w)
    sp = torch.cuda.current_stream().cuda_stream
    ref = init.clone()
    P.check(L.parhip_d_gf64_recovery(x.data_ptr(), n, w, ref.data_ptr(), r, w, w, fi, fr,
                                     1 if accumulate else 0, sp))
    out = init.clone()
    print("static entry: N=%d W=%d R=%d fi=%d fr=%d acc=%d" % (n, w, r, fi, fr, accumulate), flush=True)
    P.check(L.parhip_d_gf64_recovery_static(x.data_ptr(), n, w, out.data_ptr(), r, w, w, fi, fr,
                                            1 if accumulate else 0, sp))
    torch.cuda.synchronize()
    return ref, out, x


def test_static_equals_allocating_entry(gpu):
    import torch
    # first call: 64 MiB of input, split-K partials in the array
    ref, out, _ = _both(512, 16384, 16, 0, 512, 11)
    assert torch.equal(out, ref)
    # second call, smaller: unaligned points, a partial output group,
    # accumulate; the array still holds the first call's coefficients
    ref, out, _ = _both(37, 1500, 5, 3, 40, 12, accumulate=True)
    assert torch.equal(out, ref)
    # below 4 outputs the allocating entry uses the direct tables
    ref, out, _ = _both(100, 512, 2, 0, 100, 13)
    assert torch.equal(out, ref)


def test_static_full_size_and_oracle_columns(gpu, oracle):
    """1 GiB at R = 8 (the 4 x 8 MiB partials fit), then a small call."""
    import torch
    ref, out, x = _both(1024, 131072, 8, 0, 1024, 21)
    assert torch.equal(out, ref)
    cols = torch.as_tensor([0, 77777, 131071], device="cuda")
    samp = np.ascontiguousarray(x[:, cols].cpu().numpy().view(np.uint64))
    exp = oracle.recovery_with_coeff(samp, oracle.cauchy(1024, 8, 0, 1024))
    assert np.array_equal(out[:, cols].cpu().numpy().view(np.uint64), exp)
    del x, ref, out
    torch.cuda.empty_cache()
    ref, out, _ = _both(9, 64, 4, 0, 9, 22)
    assert torch.equal(out, ref)


def test_static_rejects_shapes_beyond_the_array(gpu):
    """Coefficients above the 64 MiB array: PARHIP_ERR_RANGE from the host
    check, before any launch (the buffers are never touched)."""
    import torch
    import parparpar_amd as P
    L = P.lib()
    x = torch.zeros((8, 64), dtype=torch.int64, device="cuda")
    y = torch.zeros((8, 64), dtype=torch.int64, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream
    # 2^21 input blocks x 8 output groups x 64 B = 1 GiB of coefficients
    rc = L.parhip_d_gf64_recovery_static(x.data_ptr(), 1 << 24, 64, y.data_ptr(), 64, 64, 64, 0, 0,
                                         0, sp)
    assert rc == P.PARHIP_ERR_RANGE
    assert b"scratch array" in L.parhip_last_error()
    assert not y.any()
