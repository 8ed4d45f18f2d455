"""PAR2 GF(2^16) parity: PAR2ProcHIP and the device muladd against the oracle.

The controller cases follow the reference's own controller harness
(test/gf16/test-ctrl.cpp:257-330): seeded random 16-bit sources, slice sizes
{2, 19998, 20000}, last-slice sizes {1, 2, s-1, s}, inputs {1, 15, 16} x
outputs {1, 15, 16}, random input indices, random exponents with exponent 0 in
the first output (multi-add path); every output compared byte-for-byte and its
checksum required valid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ctrl_cases():
    cases = []
    for s in (2, 19998, 20000):
        lasts = [1, 2] + ([s - 1, s] if s > 2 else [])
        for last in lasts:
            for nout in (1, 15, 16):
                for nin in (1, 15, 16):
                    if nin == 1 and last != s:
                        continue
                    if last != s and last != 1 and (nin > 15 or nout > 2):
                        continue
                    cases.append((s, last, nin, nout))
    return cases


@pytest.mark.parametrize("s,last,nin,nout", ctrl_cases())
def test_controller_like_test_ctrl(gpu, oracle, s, last, nin, nout):
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(0x01020304 + s * 7 + last * 3 + nin * 11 + nout)
    srcs = [rng.integers(0, 256, size=s, dtype=np.uint8) for _ in range(nin)]
    srcs[-1] = srcs[-1][:last].copy()
    inums = rng.integers(0, 32768, size=nin).tolist()
    exps = rng.integers(0, 65535, size=nout).tolist()
    exps[0] = 0
    gp = GfProc(s, input_batch=4)
    gp.set_recovery_slices(exps)
    for i in range(nin):
        gp.add(inums[i], srcs[i])
    gp.end()
    ref = oracle.par2_create(srcs, inums, exps, s)
    for e in range(nout):
        valid, out = gp.get(e)
        assert valid, "checksum invalid for output %d" % e
        assert np.array_equal(out, ref[e]), "output %d differs" % e
    gp.close()


def test_controller_custom_coeffs_and_passes(gpu, oracle):
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(9)
    s, nin, nout = 4096, 21, 6
    srcs = [rng.integers(0, 256, size=s, dtype=np.uint8) for _ in range(nin)]
    coeffs = rng.integers(0, 65536, size=(nin, nout), dtype=np.uint16)
    gp = GfProc(s, input_batch=8)
    gp.set_recovery_slices(nout)
    for i in range(nin):
        gp.add_coeffs(coeffs[i], srcs[i])
    gp.end()
    L = oracle.lib()
    for e in range(nout):
        ref = np.zeros(s + 1, dtype=np.uint8)
        for i in range(nin):
            L.or_gf16_muladd_region(ref.ctypes.data_as(oracle._u8p), oracle.p8(srcs[i]), s,
                                    int(coeffs[i, e]))
        valid, out = gp.get(e)
        assert valid and np.array_equal(out, ref[:s])
    # second pass after discardOutput: first batch overwrites
    gp.discard_output()
    gp.set_recovery_slices([0, 3])
    gp.add(5, srcs[0])
    gp.end()
    ref = oracle.par2_create([srcs[0]], [5], [0, 3], s)
    for e in range(2):
        valid, out = gp.get(e)
        assert valid and np.array_equal(out, ref[e])
    gp.close()


def test_exponent0_is_xor_parity(gpu):
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(10)
    s = 1000
    srcs = [rng.integers(0, 256, size=s, dtype=np.uint8) for _ in range(9)]
    gp = GfProc(s)
    gp.set_recovery_slices([0])
    for i, b in enumerate(srcs):
        gp.add(100 + i, b)
    gp.end()
    x = np.zeros(s, dtype=np.uint8)
    for b in srcs:
        x ^= b
    valid, out = gp.get(0)
    assert valid and np.array_equal(out, x)


def test_device_coeffs(gpu, oracle):
    import torch
    from parparpar_amd import device
    inums = np.arange(0, 32768, 97, dtype=np.uint16)
    exps = np.array([0, 1, 2, 255, 511, 4097, 65534], dtype=np.uint16)
    c = device.gf16_coeffs(inums, exps).cpu().numpy().view(np.uint16)
    L = oracle.lib()
    for e in range(exps.size):
        for i in range(inums.size):
            assert c[e, i] == L.or_gf16_coeff(int(inums[i]), int(exps[e]))


@pytest.mark.parametrize("n,words,r", [(32768, 16384, 512), (300, 5000, 37)])
def test_device_muladd_columns(gpu, oracle, n, words, r):
    """1 GiB as 32768 x 32 KiB at 512 recovery slices (exponents 0..511), and a
    ragged small shape: exact check on sampled word columns."""
    import torch
    from parparpar_amd import device
    print("gf16 muladd N=%d words=%d R=%d" % (n, words, r), flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(n + r)
    x = torch.randint(-32768, 32767, (n, words), dtype=torch.int16, device="cuda", generator=g)
    inums = np.arange(n, dtype=np.uint16)
    exps = np.arange(r, dtype=np.uint16)
    co = device.gf16_coeffs(inums, exps)
    out = device.gf16_muladd(x, co)
    torch.cuda.synchronize()
    rng = np.random.default_rng(n)
    cols = np.unique(np.concatenate([[0, words - 1], rng.integers(0, words, size=6)])).astype(np.int64)
    tc = torch.as_tensor(cols, device="cuda")
    samp = np.ascontiguousarray(x[:, tc].cpu().numpy().view(np.uint8))  # (n, 2*len(cols)) bytes
    ref = oracle.par2_columns(samp, 2 * len(cols), inums, exps, np.arange(len(cols), dtype=np.uint64))
    got = out[:, tc].cpu().numpy().view(np.uint16)
    assert np.array_equal(got, ref)
    del x, out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("s,batch,nin", [(20000, 4, 11), (19999, 3, 6), (2, 2, 5)])
def test_fill_then_dummy_inputs(gpu, oracle, s, batch, nin):
    """The benchmarking pair of the backend contract (gf16/controller.h:303-305):
    fill_input fills one staging slot per call until every slot of both staging
    areas holds the buffer; dummy_input then adds inputs over that data, so
    recovery e = XOR_i coeff(i, e) * buffer, checksum valid."""
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(s + batch)
    buf = rng.integers(0, 256, size=s, dtype=np.uint8)
    exps = [0, 1, 77, 65534]
    inums = rng.integers(0, 32768, size=nin).tolist()
    gp = GfProc(s, input_batch=batch)
    gp.set_recovery_slices(exps)
    calls = 1
    while not gp.fill_input(buf):
        calls += 1
        assert calls <= 2 * batch
    assert calls == 2 * batch            # one slot per call, two staging areas
    for i in inums:
        gp.dummy_input(i)
    gp.end()
    ref = oracle.par2_create([buf] * nin, inums, exps, s)
    for e in range(len(exps)):
        valid, out = gp.get(e)
        assert valid, "checksum invalid for output %d" % e
        assert np.array_equal(out, ref[e]), e
    gp.close()


@pytest.mark.parametrize("s", [262144, 1048579])
def test_large_slices_staged_in_parallel(gpu, oracle, s):
    """Slices of 256 KiB and more are copied and checksummed on the host copy
    threads; outputs equal the oracle and every checksum verifies (odd size too)."""
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(s)
    nin, exps = 5, [0, 3, 40000]
    srcs = [rng.integers(0, 256, size=s, dtype=np.uint8) for _ in range(nin)]
    inums = [1, 7, 100, 2000, 32767]
    gp = GfProc(s, input_batch=2)
    gp.set_recovery_slices(exps)
    for i in range(nin):
        gp.add(inums[i], srcs[i])
    gp.end()
    ref = oracle.par2_create(srcs, inums, exps, s)
    for e in range(len(exps)):
        valid, out = gp.get(e)
        assert valid, "checksum invalid for output %d" % e
        assert np.array_equal(out, ref[e]), e
    gp.close()


def _hip():
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def _patch_words(hip, dptr, edits):
    """Overwrite 16-bit words of a device row: edits = {word index: new value}."""
    import ctypes as C
    for w, v in edits.items():
        buf = np.array([v], dtype=np.uint16)
        assert hip.hipMemcpy(dptr + 2 * w, buf.ctypes.data, 2, 1) == 0  # host to device


def _read_words(hip, dptr, n):
    buf = np.zeros(n, dtype=np.uint16)
    assert hip.hipMemcpy(buf.ctypes.data, dptr, 2 * n, 2) == 0  # device to host
    return buf


@pytest.mark.parametrize("s", [20000, 1 << 20])
def test_cksum_detects_corruption(gpu, oracle, s):
    """cksumValid in the negative direction (test/gf16/test-ctrl.cpp:71, :110-112):
    between endInput and getOutput, one 16-bit word of recovery slice 1 is
    flipped, two different words of slice 2 are swapped, and the checksum word
    of slice 3 is changed; those three read back invalid, the untouched slices
    valid and equal to the oracle."""
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(s)
    nin, nout = 9, 5
    srcs = [rng.integers(0, 256, size=s, dtype=np.uint8) for _ in range(nin)]
    inums = rng.choice(32768, nin, replace=False).tolist()
    exps = [0, 3, 17, 200, 4000]
    gp = GfProc(s, input_batch=4)
    gp.set_recovery_slices(exps)
    for i in range(nin):
        gp.add(inums[i], srcs[i])
    gp.end()
    ref = oracle.par2_create(srcs, inums, exps, s)
    hip = _hip()
    words = s // 2
    d1, _ = gp.output_device_ptr(1)
    w1 = int(rng.integers(0, words))
    old = _read_words(hip, d1 + 2 * w1, 1)[0]
    _patch_words(hip, d1, {w1: old ^ 0x0100})
    d2, _ = gp.output_device_ptr(2)
    row = _read_words(hip, d2, words)
    a = int(rng.integers(0, words))
    b = next(j for j in range(a + 1, a + words) if row[j % words] != row[a]) % words
    _patch_words(hip, d2, {a: row[b], b: row[a]})
    d3, _ = gp.output_device_ptr(3)
    ck = _read_words(hip, d3 + 2 * words, 1)[0]
    _patch_words(hip, d3, {words: ck ^ 1})
    for e in range(nout):
        valid, out = gp.get(e)
        if e in (1, 2, 3):
            assert not valid, "corrupted output %d reported valid" % e
        else:
            assert valid and np.array_equal(out, ref[e]), e
    gp.close()


def test_short_input_in_large_slice(gpu, oracle):
    """Inputs shorter than the slice (one of 100 KiB, one of 3 bytes) in a
    1 MiB slice: zero-padded staging, valid checksums, the oracle's bytes."""
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(77)
    s = 1 << 20
    srcs = [rng.integers(0, 256, size=n, dtype=np.uint8) for n in (s, 100 << 10, s, 3)]
    inums = [5, 900, 31000, 7]
    exps = [0, 1, 2, 511]
    gp = GfProc(s, input_batch=2)
    gp.set_recovery_slices(exps)
    for i, b in zip(inums, srcs):
        gp.add(i, b)
    gp.end()
    ref = oracle.par2_create(srcs, inums, exps, s)
    for e in range(len(exps)):
        valid, out = gp.get(e)
        assert valid and np.array_equal(out, ref[e]), e
    gp.close()


@pytest.mark.parametrize("areas", [2, 4, 8])
def test_async_adds_and_passes(gpu, oracle, areas):
    """Adds that return before their copies (the copy workers stage them), many
    batches over 2, 4 or 8 staging areas, then a second pass after
    discardOutput overwrites the first: every output equals the oracle of its
    pass."""
    from parparpar_amd.par2 import GfProc
    rng = np.random.default_rng(78 + areas)
    s, nout = 6000, 7
    exps = rng.choice(65535, nout, replace=False).tolist()
    gp = GfProc(s, input_batch=5, staging_areas=areas)
    gp.set_recovery_slices(exps)
    for npass, nin in enumerate((37, 12, 203)):
        srcs = [rng.integers(0, 256, size=s, dtype=np.uint8) for _ in range(nin)]
        inums = rng.choice(32768, nin, replace=False).tolist()
        if npass:
            gp.discard_output()
        for i in range(nin):
            gp.add_async(inums[i], srcs[i])
        gp.end()
        ref = oracle.par2_create(srcs, inums, exps, s)
        for e in range(nout):
            valid, out = gp.get(e)
            assert valid and np.array_equal(out, ref[e]), (npass, e)
    assert gp.batches_started() >= 8 + 3 + 41
    gp.close()


def test_mixed_coefficient_sources_in_one_batch(gpu, oracle):
    """Inputs with caller-given coefficients and inputs by number share staging
    areas (a batch with any caller-given column takes its coefficients from the
    host, the others from the device): outputs equal the oracle's PAR2 create
    over the same input numbers."""
    from parparpar_amd.par2 import GfProc, gf16_coeff
    rng = np.random.default_rng(91)
    s, nout, nin = 4002, 6, 29
    exps = rng.choice(65535, nout, replace=False).tolist()
    srcs = [rng.integers(0, 256, size=s, dtype=np.uint8) for _ in range(nin)]
    inums = rng.choice(32768, nin, replace=False).tolist()
    gp = GfProc(s, input_batch=4, staging_areas=3)
    gp.set_recovery_slices(exps)
    for i in range(nin):
        if i % 3 == 1:
            gp.add_coeffs([gf16_coeff(inums[i], e) for e in exps], srcs[i])
        else:
            gp.add_async(inums[i], srcs[i])
    gp.end()
    ref = oracle.par2_create(srcs, inums, exps, s)
    for e in range(nout):
        valid, out = gp.get(e)
        assert valid and np.array_equal(out, ref[e]), e
    gp.close()
