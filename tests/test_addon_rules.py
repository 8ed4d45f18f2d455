"""Argument rules of the addon mirror, on the CPU (no library call is made).

The error scenarios of the reference's test/par3-native-edge.js: zero counts
(1, 2), block sizes 7 and 0 (5, 6), numInputs = int32 max against a small
buffer (8), a short output (13); each raises before anything is computed and
leaves the output untouched, as the N-API layer does
(src/gf64_addon.cc:1383-1391, :1553-1578).
"""
import numpy as np
import pytest

from parparpar_amd import binding

ENTRIES = [binding.compute_recovery, binding.compute_recovery_full,
           binding.compute_recovery_barycentric]


@pytest.mark.parametrize("fn", ENTRIES)
@pytest.mark.parametrize("args", [
    (np.zeros(1, np.uint64), 0, 1, 8),                 # scenario 1: numInputs = 0
    (np.zeros(1, np.uint64), 1, 0, 8),                 # scenario 2: numRecovery = 0
    (np.zeros(2, np.uint64), 2, 1, 7),                 # scenario 5: blockSize 7
    (np.zeros(2, np.uint64), 2, 1, 0),                 # scenario 6: blockSize 0
    (np.zeros(16, np.uint64), 2147483647, 1, 8),       # scenario 8: int32 max inputs
    (np.zeros(2, np.uint64), -1, 1, 8),
])
def test_rejected_before_compute(fn, args):
    inp, n, r, bs = args
    out = np.full(4, 7, dtype=np.uint64)
    with pytest.raises(ValueError):
        fn(inp, out, n, r, bs, 0, max(n, 1))
    assert (out == 7).all()


@pytest.mark.parametrize("fn", ENTRIES)
def test_short_output(fn):
    """Scenario 13: an output smaller than numRecovery x blockSize."""
    out = np.zeros(1, np.uint64)
    with pytest.raises(ValueError):
        fn(np.zeros(4, np.uint64), out, 2, 2, 16, 0, 2)


def test_with_coeff_and_streaming_rules():
    out = np.zeros(4, np.uint64)
    with pytest.raises(ValueError):   # coefficient matrix shorter than R x N
        binding.compute_recovery_with_coeff(np.zeros(3, np.uint64), np.zeros(4, np.uint64), out, 2, 2,
                                            16)
    with pytest.raises(TypeError):    # batches must be a list of buffers
        binding.compute_recovery_streaming(np.zeros(4, np.uint64), out, 1, 1, 8, 0, 1)
    with pytest.raises(ValueError):   # a batch that is not whole slices
        binding.compute_recovery_streaming([np.zeros(3, np.uint8)], out, 1, 1, 8, 0, 1)
    with pytest.raises(ValueError):   # batches hold fewer slices than numInputs
        binding.compute_recovery_streaming([np.zeros(1, np.uint64)], out, 2, 1, 8, 0, 2)
    assert binding.compute_recovery_streaming([], out, 0, 1, 8, 0, 0) is out   # empty: no-op


def test_packet_and_region_rules():
    with pytest.raises(TypeError):    # InputSetID must be 8 bytes
        binding.assemble_recovery_packets(np.zeros(8, np.uint8), np.zeros(4, np.uint8), 0, 1, 8)
    with pytest.raises(ValueError):   # block buffer length != numRecovery x blockSize
        binding.assemble_recovery_packets(np.zeros(8, np.uint8), np.zeros(8, np.uint8), 0, 2, 8)
    with pytest.raises(ValueError):   # mul_arr with coefficients but a zero count
        binding.mul_arr(np.zeros(2, np.uint64), np.zeros(2, np.uint64), np.zeros(1, np.uint64), 2, 0)
    with pytest.raises(ValueError):   # more words than the output holds
        binding.mul_arr(np.zeros(2, np.uint64), np.zeros(4, np.uint64), np.zeros(1, np.uint64), 4, 1)
