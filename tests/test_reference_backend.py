"""PAR2ProcHIP, the reference controller's backend over the engine
(parparpar_amd/reference_backend/): it compiles against the reference's own
gf16/controller.h in both flavours the reference builds its backends in
(futures, USE_LIBUV), and links with the reference's controller
(gf16/controller.cpp) and the engine library into the driver programs the GPU
tests run (oracle/Makefile, target `backend`)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
ORACLE = os.path.join(ROOT, "oracle")
BACKEND = os.path.join(ROOT, "parparpar_amd", "reference_backend")

pytestmark = pytest.mark.skipif(
    not os.path.isdir(os.path.join(REF, "gf16")) or
    not os.path.exists(os.path.join(ROOT, "parparpar_amd", "libparhip.so")),
    reason="needs the reference tree (for its controller headers) and the built engine library")


@pytest.mark.parametrize("libuv", [False, True])
def test_backend_compiles_against_reference_controller(libuv):
    """-Wall -Werror, C++17, against gf16/controller.h as the reference's
    backends include it: every pure virtual of IPAR2ProcBackend is overridden
    (an abstract class would not be instantiable in the driver)."""
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
           "-I" + os.path.join(REF, "gf16"), "-I" + os.path.join(ROOT, "include"),
           os.path.join(BACKEND, "controller_hip.cpp")]
    if libuv:
        cmd.insert(1, "-DUSE_LIBUV")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_driver_programs_build():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "backend"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    for flavour in ("futures", "libuv"):
        exe = os.path.join(ORACLE, "_ref", "par2_backend_" + flavour)
        assert os.access(exe, os.X_OK), exe
        # no arguments: the usage line, before any device is touched
        u = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        assert u.returncode == 1 and "usage" in u.stderr, (flavour, u.stderr)
