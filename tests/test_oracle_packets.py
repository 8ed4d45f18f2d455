"""Pins of the oracle's PAR3 packet assembly and BLAKE3 before they are trusted.

- BLAKE3 of the empty message equals the value the BLAKE3 specification
  publishes (af1349b9...).
- Every packet in two reference fixtures verifies: its bytes [8, 24) are the
  16-byte BLAKE3 of bytes [24, end).  The fixtures come from two writers (this
  project's JS packetizer, which seals with the npm blake3 package, and
  par3cmdline, an independent C implementation), and the messages span 56 to
  4184 bytes, so single-chunk and multi-chunk (five 1 KiB chunks, a
  non-power-of-two tree) hashing are both covered.
- The three REC packets of par3-golden-parpar.par3 (block size 4096, indices
  8..10) are reproduced byte for byte by the oracle's assembler from their own
  block data, input set id and first index (layout of
  src/gf64_addon.cc:2890-2938).
- Where oracle/_ref is built, the oracle equals the reference's own
  hasher/hasher_blake3.c over lengths that cross every block and chunk edge.
"""
import os
import struct

import numpy as np
import pytest

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = ["par3-golden-parpar.par3", "par3-golden-par3cmdline.vol0+1.par3"]


def packets(data):
    off = 0
    while off + 48 <= len(data):
        assert data[off:off + 8] == b"PAR3\x00PKT"
        n = struct.unpack("<Q", data[off + 24:off + 32])[0]
        yield data[off:off + n]
        off += n
    assert off == len(data)


def test_blake3_empty_message(oracle):
    assert oracle.blake3(b"").hex() == \
        "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"


@pytest.mark.parametrize("fixture", FIXTURES)
def test_fixture_packet_checksums(oracle, fixture):
    data = open(os.path.join(GOLDEN_DIR, fixture), "rb").read()
    pk = list(packets(data))
    assert len(pk) >= 9
    lengths = set()
    for p in pk:
        assert oracle.blake3(p[24:])[:16] == p[8:24]
        lengths.add(len(p) - 24)
    assert max(lengths) > 4096 and min(lengths) < 1024


def test_fixture_rec_packets_reassembled(oracle):
    data = open(os.path.join(GOLDEN_DIR, "par3-golden-parpar.par3"), "rb").read()
    rec = [p for p in packets(data) if p[40:48] == b"PAR REC\x00"]
    assert len(rec) == 3
    bs = len(rec[0]) - 64
    first = struct.unpack("<Q", rec[0][48:56])[0]
    blocks = b"".join(p[64:] for p in rec)
    got = oracle.assemble_recovery_packets(blocks, rec[0][32:40], first, 3, bs)
    assert got == b"".join(rec)
    # negative trap: a different first index changes every packet's bytes
    other = oracle.assemble_recovery_packets(blocks, rec[0][32:40], first + 1, 3, bs)
    assert all(other[i * (64 + bs):(i + 1) * (64 + bs)] != rec[i] for i in range(3))


def test_blake3_tree_shapes_vs_reference(oracle):
    if not oracle.have_ref():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(3)
    for n in [0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 2049, 3072, 3073, 5 * 1024 + 7,
              8 * 1024, 9 * 1024 + 1, 31 * 1024, 65536 + 40, (1 << 20) + 40]:
        m = rng.bytes(n)
        cut = n // 3
        assert oracle.blake3(m) == oracle.ref_blake3(m[:cut], m[cut:]), n
