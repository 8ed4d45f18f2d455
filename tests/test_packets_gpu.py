"""PAR3 REC packets assembled and BLAKE3-sealed on the device, against the oracle.

Grid of the reference's own packetizer parity test
(test/par3-packetizer-parity.js: numRecovery {1, 4, 8, 32}, blockSize {32,
4096, 65536, 1048576}, firstRecoveryIndex {0, 1, 100, 1 << 20}), odd block
sizes on the byte path, chunk-count edges (one chunk at blockSize 984, two at
985), the three REC packets of a reference fixture, unaligned device buffers,
and the argument errors of AssembleRecoveryPackets_NAPI.
"""
import os
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ISID = bytes.fromhex("8fd25983a81a8890")


@pytest.mark.parametrize("num", [1, 4, 8, 32])
@pytest.mark.parametrize("bs", [32, 4096, 65536, 1048576])
def test_reference_grid(gpu, oracle, num, bs):
    from parparpar_amd import binding
    rng = np.random.default_rng(num * 7 + bs)
    rec = rng.bytes(num * bs)
    for first in (0, 1, 100, 1 << 20):
        got = binding.assemble_recovery_packets(rec, ISID, first, num, bs)
        assert got == oracle.assemble_recovery_packets(rec, ISID, first, num, bs), (first,)


@pytest.mark.parametrize("bs", [1, 7, 8, 12, 983, 984, 985, 1023, 1024, 2047, 3000, 4097,
                                64 * 1024 - 40, 64 * 1024 - 39, 65 * 1024 - 40, 131032])
def test_odd_and_edge_block_sizes(gpu, oracle, bs):
    from parparpar_amd import binding
    rng = np.random.default_rng(bs)
    num = 5
    rec = rng.bytes(num * bs)
    got = binding.assemble_recovery_packets(rec, ISID, 77, num, bs)
    assert got == oracle.assemble_recovery_packets(rec, ISID, 77, num, bs)


def test_fixture_rec_packets(gpu):
    from parparpar_amd import binding
    path = os.path.join(os.path.dirname(__file__), "golden", "par3-golden-parpar.par3")
    data = open(path, "rb").read()
    off, rec = 0, []
    while off < len(data):
        n = struct.unpack("<Q", data[off + 24:off + 32])[0]
        if data[off + 40:off + 48] == b"PAR REC\x00":
            rec.append(data[off:off + n])
        off += n
    bs = len(rec[0]) - 64
    first = struct.unpack("<Q", rec[0][48:56])[0]
    got = binding.assemble_recovery_packets(b"".join(p[64:] for p in rec), rec[0][32:40], first,
                                            len(rec), bs)
    assert got == b"".join(rec)


@pytest.mark.parametrize("shift", [0, 1, 4])
def test_device_entry_unaligned(gpu, oracle, shift):
    import torch
    from parparpar_amd import device
    num, bs = 3, 5000
    rng = np.random.default_rng(shift)
    rec_h = rng.bytes(num * bs)
    rec = torch.frombuffer(bytearray(rec_h), dtype=torch.uint8).reshape(num, bs).cuda()
    base = torch.zeros(num * (64 + bs) + 8, dtype=torch.uint8, device="cuda")
    out = base[shift:shift + num * (64 + bs)].view(num, 64 + bs)
    device.assemble_recovery_packets(rec, ISID, 9, out=out)
    torch.cuda.synchronize()
    assert bytes(out.cpu().numpy().tobytes()) == oracle.assemble_recovery_packets(rec_h, ISID, 9, num, bs)


def test_after_create_on_device(gpu, oracle):
    """Recovery blocks straight from the device contraction into packets."""
    import torch
    from parparpar_amd import device
    n, w, r = 16, 4096, 4
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    inp = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
    rec = device.recovery(inp, r, 0, n)
    pk = device.assemble_recovery_packets(rec.view(torch.uint8).view(r, w * 8), ISID, n)
    torch.cuda.synchronize()
    host = np.ascontiguousarray(inp.cpu().numpy().view(np.uint64))
    blocks = oracle.recovery(host, r, 0, n).tobytes()
    assert pk.cpu().numpy().tobytes() == oracle.assemble_recovery_packets(blocks, ISID, n, r, w * 8)


def test_argument_errors(gpu):
    import parparpar_amd as P
    from parparpar_amd import binding
    with pytest.raises(TypeError):
        binding.assemble_recovery_packets(b"\0" * 8, b"1234567", 0, 1, 8)
    with pytest.raises(ValueError):
        binding.assemble_recovery_packets(b"", ISID, 0, 0, 8)
    with pytest.raises(ValueError):
        binding.assemble_recovery_packets(b"\0" * 9, ISID, 0, 1, 8)
    out = bytearray(72)
    assert P.lib().parhip_assemble_recovery_packets(b"\0" * 8, ISID, 0, 0, 8, bytes(out)) == P.PARHIP_ERR_RANGE
    assert P.lib().parhip_assemble_recovery_packets(b"\0" * 8, ISID, 0, 1, -8, bytes(out)) == P.PARHIP_ERR_RANGE
    assert P.lib().parhip_assemble_recovery_packets(b"\0" * 8, ISID, 0, 1 << 40, 2, bytes(out)) == P.PARHIP_ERR_RANGE


def test_block_above_64_mib(gpu, oracle):
    """More than 1024 groups of 64 chunks: the parent tree takes two passes."""
    from parparpar_amd import binding
    bs = 65 * 1024 * 1024 + 8
    rng = np.random.default_rng(65)
    rec = rng.bytes(2 * bs)
    got = binding.assemble_recovery_packets(rec, ISID, 5, 2, bs)
    assert got == oracle.assemble_recovery_packets(rec, ISID, 5, 2, bs)
