"""The reference's PAR2 controller (PAR2Proc, gf16/controller.cpp) driving the
PAR2ProcHIP backend, futures and USE_LIBUV flavours, through the driver
programs built by `make -C oracle backend` (oracle/par2_backend_driver.cc,
modelled on the reference's controller test, test/gf16/test-ctrl.cpp:140-215).
Every recovery slice must equal the oracle's PAR2 create and report
cksumValid.  Cases: the controller test's slice shapes (odd last slice), many
batches over 2 or 3 staging areas, each slice split over two backends, a
discarded first pass, and inputs added with explicit coefficients."""
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")

SPLIT, TWO_PASSES, COEFFS = 1, 2, 4
# (slice size, last slice size, inputs, outputs, inputs per batch, staging areas, flags)
CASES = [
    (20000, 19999, 15, 2, 0, 2, 0),
    (20000, 20000, 16, 16, 4, 2, 0),
    (19998, 1, 15, 15, 3, 3, SPLIT),
    (6000, 6000, 37, 7, 5, 2, TWO_PASSES),
    (4002, 4001, 29, 6, 4, 3, COEFFS),
    (8190, 77, 41, 5, 6, 2, SPLIT | TWO_PASSES | COEFFS),
    (2, 1, 16, 15, 0, 2, 0),
]


def _case(oracle, s, last, nin, nout, batch, areas, flags, seed):
    rng = np.random.default_rng(seed)
    exps = rng.choice(np.arange(1, 65535), nout, replace=False).astype(np.uint16)
    if not flags & COEFFS:
        exps[0] = 0  # exponent 0, as the reference's controller test sets it
    while True:
        inums = rng.choice(32768, nin, replace=False).astype(np.uint16)
        # the libuv controller keys pending adds by input number, or by the
        # first coefficient for explicit-coefficient adds: keep them distinct
        refs = [int(inums[i]) if not (flags & COEFFS and i & 1) else
                oracle.gf16_coeff(int(inums[i]), int(exps[0])) for i in range(nin)]
        if len(set(refs)) == nin:
            break
    data = rng.integers(0, 256, size=(nin, s), dtype=np.uint8)
    blob = struct.pack("<QQIIIII", s, last, nin, nout, batch, areas, flags)
    blob += inums.tobytes() + exps.tobytes() + data.tobytes()
    srcs = [data[i, :(last if i == nin - 1 else s)] for i in range(nin)]
    return blob, oracle.par2_create(srcs, inums.tolist(), exps.tolist(), s)


@pytest.mark.parametrize("flavour", ["futures", "libuv"])
@pytest.mark.parametrize("case", CASES, ids=[
    "s%d-last%d-%dx%d-b%d-a%d-f%d" % c for c in CASES])
def test_reference_controller_over_backend(gpu, oracle, tmp_path, flavour, case):
    exe = os.path.join(REF_BIN, "par2_backend_" + flavour)
    if not os.access(exe, os.X_OK):
        pytest.skip("driver not built (make -C oracle backend needs the reference tree)")
    s, last, nin, nout = case[:4]
    blob, ref = _case(oracle, *case, seed=hash(case) & 0xFFFF)
    cf, rf = tmp_path / "case.bin", tmp_path / "result.bin"
    cf.write_bytes(blob)
    print("driver", flavour, case, flush=True)
    r = subprocess.run([exe, str(cf), str(rf)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    res = np.frombuffer(rf.read_bytes(), dtype=np.uint8)
    out = res[:nout * s].reshape(nout, s)
    valid = res[nout * s:]
    assert valid.size == nout and valid.all(), valid
    for e in range(nout):
        assert np.array_equal(out[e], ref[e]), e
