"""Host -> HBM streaming (pinned staging ring) and the streaming create entry.

Chunked streaming must give exactly the single-contraction bytes (linearity):
checked against the oracle with tiny chunks (many chunks, ragged last chunk,
chunks smaller than one slice), against the resident-device result at the
default 64 MiB chunk, and for par3_create_streaming against the oracle on the
zero-padded file contents (src/gf64_create_streaming.cc:141-171).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def chunk(gpu):
    from parparpar_amd import binding
    yield binding.set_stream_chunk
    binding.set_stream_chunk(0)


def rand_words(rng, *shape):
    return rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64)


@pytest.mark.parametrize("n,w,r,chunk_bytes,acc", [
    (37, 1024, 5, 64 << 10, False),   # 8 slices per chunk, ragged last chunk
    (37, 1024, 5, 64 << 10, True),
    (9, 2048, 3, 4096, False),        # chunk smaller than a slice: one slice per chunk
    (64, 512, 8, 12288, True),        # 3 slices per chunk
])
def test_streamed_recovery_matches_oracle(chunk, oracle, n, w, r, chunk_bytes, acc):
    from parparpar_amd import binding
    chunk(chunk_bytes)
    rng = np.random.default_rng(n * 31 + w + r)
    inp = rand_words(rng, n, w)
    init = rand_words(rng, r, w)
    out = init.copy()
    binding.compute_recovery(inp, out, n, r, w * 8, 0, n, 0, accumulate=acc)
    expect = init.copy() if acc else np.zeros((r, w), dtype=np.uint64)
    oracle.recovery_with_coeff(inp, oracle.cauchy(n, r, 0, n), expect, acc)
    assert np.array_equal(out, expect)


def test_streamed_with_coeff(chunk, oracle):
    from parparpar_amd import binding
    chunk(16 << 10)
    rng = np.random.default_rng(5)
    n, w, r = 21, 700, 4
    inp = rand_words(rng, n, w)
    co = rand_words(rng, r, n)
    out = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery_with_coeff(co, inp, out, n, r, w * 8)
    assert np.array_equal(out, oracle.recovery_with_coeff(inp, co))


def test_default_chunk_large_input_equals_device(gpu, oracle):
    """80 x 1 MiB host input (above one 64 MiB chunk) equals the resident-device result."""
    import torch
    from parparpar_amd import binding, device
    n, w, r = 80, 131072, 4
    rng = np.random.default_rng(80)
    inp = np.frombuffer(rng.bytes(n * w * 8), dtype=np.uint64).reshape(n, w)
    out = np.zeros((r, w), dtype=np.uint64)
    binding.compute_recovery(inp, out, n, r, w * 8, 0, n)
    d = device.recovery(torch.from_numpy(inp.view(np.int64).copy()).cuda(), r, 0, n)
    assert np.array_equal(out, d.cpu().numpy().view(np.uint64))
    cols = np.array([0, 1, 77777, w - 1])
    sample = np.ascontiguousarray(inp[:, cols])
    assert np.array_equal(out[:, cols], oracle.recovery_with_coeff(sample, oracle.cauchy(n, r, 0, n)))


@pytest.mark.parametrize("size,bs,chunk_bytes", [
    (37 * 8192 + 1000, 8192, 64 << 10),   # ragged tail zero-padded, several chunks
    (5 * 4096, 4096, 0),                 # exact multiple, one chunk
    (100, 1024, 0),                      # shorter than one block
])
def test_create_streaming_file(chunk, oracle, tmp_path, size, bs, chunk_bytes):
    from parparpar_amd import binding
    chunk(chunk_bytes)
    rng = np.random.default_rng(size)
    data = rng.bytes(size)
    path = tmp_path / "src.bin"
    path.write_bytes(data)
    n = (size + bs - 1) // bs
    r = 3
    res = binding.par3_create_streaming(str(path), {"recoverySlices": r, "blockSize": bs,
                                                    "firstInput": 0, "firstRecovery": n})
    assert res["recoveryBytes"] == r * bs
    padded = np.frombuffer(data + bytes(n * bs - size), dtype=np.uint64).reshape(n, bs // 8)
    expect = oracle.recovery(np.ascontiguousarray(padded), r, 0, n)
    assert res["recoveryBuffer"] == expect.tobytes()
    assert res["durationMs"] > 0


def test_create_streaming_errors(gpu, tmp_path):
    from parparpar_amd import ParhipError, binding
    with pytest.raises(ParhipError, match="ENOENT"):
        binding.par3_create_streaming(str(tmp_path / "missing"), {"recoverySlices": 2, "blockSize": 1024})
    empty = tmp_path / "empty"
    empty.write_bytes(b"")
    with pytest.raises(ParhipError, match="empty"):
        binding.par3_create_streaming(str(empty), {"recoverySlices": 2, "blockSize": 1024})
    f = tmp_path / "f"
    f.write_bytes(os.urandom(100))
    with pytest.raises(ParhipError, match="multiple of 8"):
        binding.par3_create_streaming(str(f), {"recoverySlices": 2, "blockSize": 1020})
    with pytest.raises(ParhipError, match="positive"):
        binding.par3_create_streaming(str(f), {"recoverySlices": 0, "blockSize": 1024})


@pytest.mark.parametrize("mmap", [False, True])
def test_create_streaming_short_file_leg5(chunk, oracle, tmp_path, mmap):
    """test/par3-create-streaming-parity.js:321-330 (leg 5): a 1000-byte source,
    bytes (7i + 3) mod 256, 4096-byte blocks, 4 recovery slices: the one block is
    zero-padded, read or mapped."""
    from parparpar_amd import binding
    chunk(0)
    data = bytes((7 * i + 3) & 0xFF for i in range(1000))
    path = tmp_path / "leg5.bin"
    path.write_bytes(data)
    res = binding.par3_create_streaming(str(path), {"recoverySlices": 4, "blockSize": 4096,
                                                    "firstInput": 0, "firstRecovery": 1,
                                                    "useMmap": mmap})
    padded = np.frombuffer(data + bytes(4096 - 1000), dtype=np.uint64).reshape(1, 512)
    assert res["recoveryBuffer"] == oracle.recovery(padded, 4, 0, 1).tobytes()


def _from_file(path, out, r, w, fi, fr):
    import ctypes
    from parparpar_amd import lib
    p = None if path is None else str(path).encode()
    return lib().parhip_gf64_compute_recovery_blocks_from_file(
        p, None if out is None else out.ctypes.data, r, w, fi, fr, 0)


@pytest.mark.parametrize("n,bs,r,chunk_bytes", [
    (37, 8192, 5, 64 << 10),     # several chunks
    (300, 1 << 20, 8, 0),        # 300 MiB: five default chunks
    (1, 64, 1, 0),
])
def test_from_file_into_caller_buffer(chunk, oracle, tmp_path, monkeypatch, n, bs, r, chunk_bytes):
    """ComputeRecoveryBlocksFromFile: the file's blocks are the inputs, the
    recovery blocks land in the caller's buffer; equal to the oracle (small)
    or to the device-resident create plus oracle columns (300 MiB)."""
    import torch
    from parparpar_amd import device
    chunk(chunk_bytes)
    monkeypatch.setenv("PAR3_GF64_USE_MMAP", "1")
    w = bs // 8
    rng = np.random.default_rng(n * 3 + r)
    inp = np.frombuffer(rng.bytes(n * bs), dtype=np.uint64).reshape(n, w)
    path = tmp_path / "src.bin"
    inp.tofile(path)
    fi, fr = 3, n + 11
    out = np.zeros((r, w), dtype=np.uint64)
    assert _from_file(path, out, r, w, fi, fr) == 0
    if n * bs <= (4 << 20):
        assert np.array_equal(out, oracle.recovery_with_coeff(inp, oracle.cauchy(n, r, fi, fr)))
    else:
        d = device.recovery(torch.from_numpy(inp.view(np.int64).copy()).cuda(), r, fi, fr)
        assert np.array_equal(out, d.cpu().numpy().view(np.uint64))
        cols = np.array([0, 5, w // 2, w - 1])
        ref = oracle.recovery_with_coeff(np.ascontiguousarray(inp[:, cols]), oracle.cauchy(n, r, fi, fr))
        assert np.array_equal(out[:, cols], ref)
