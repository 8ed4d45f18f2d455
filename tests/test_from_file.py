"""parhip_gf64_compute_recovery_blocks_from_file's failures, as the reference's
ComputeRecoveryBlocksFromFile reports them (src/par3_engine.cc:1445-1537): -1
for the environment gate off, a null argument, a file that cannot be opened and
a size that is not a positive multiple of the block size.  Each is detected
before any device work, so these run without a GPU."""
import numpy as np
import pytest


def _call(path, out, r=2, w=8):
    from parparpar_amd import lib
    p = None if path is None else str(path).encode()
    return lib().parhip_gf64_compute_recovery_blocks_from_file(
        p, None if out is None else out.ctypes.data, r, w, 0, 100, 0)


def _err():
    from parparpar_amd import lib
    return lib().parhip_last_error().decode()


def test_gate_off(tmp_path, monkeypatch):
    monkeypatch.delenv("PAR3_GF64_USE_MMAP", raising=False)
    f = tmp_path / "f"
    f.write_bytes(bytes(64))
    assert _call(f, np.zeros((2, 8), dtype=np.uint64)) == -1
    assert "PAR3_GF64_USE_MMAP" in _err()
    monkeypatch.setenv("PAR3_GF64_USE_MMAP", "0")
    assert _call(f, np.zeros((2, 8), dtype=np.uint64)) == -1


@pytest.mark.parametrize("size,what", [(0, "positive multiple"), (63, "positive multiple"),
                                       (65, "positive multiple")])
def test_size_not_a_multiple(tmp_path, monkeypatch, size, what):
    monkeypatch.setenv("PAR3_GF64_USE_MMAP", "1")
    f = tmp_path / "f"
    f.write_bytes(bytes(size))
    assert _call(f, np.zeros((2, 8), dtype=np.uint64)) == -1
    assert what in _err()


def test_missing_file_and_nulls(tmp_path, monkeypatch):
    monkeypatch.setenv("PAR3_GF64_USE_MMAP", "1")
    out = np.zeros((2, 8), dtype=np.uint64)
    assert _call(tmp_path / "missing", out) == -1
    assert "open" in _err()
    assert _call(None, out) == -1
    f = tmp_path / "f"
    f.write_bytes(bytes(64))
    assert _call(f, None) == -1
    assert "null" in _err()
