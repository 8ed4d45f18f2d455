"""The C-ABI library loads and exports every symbol include/*.h declares.

No compute calls here: this runs on machines without a GPU.
"""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "parparpar_amd", "libparhip.so")

DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(parhip_\w+)\s*\(", re.M)


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        names.update(DECL.findall(text))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "parhip_gf64_compute_recovery_blocks" in names
    assert "parhip_gf64_solve_and_reconstruct" in names
    assert len(names) >= 25


@pytest.fixture(scope="module")
def built_lib():
    if not os.path.exists(LIB):
        from parparpar_amd import build
        build.build()
    return LIB


def test_library_exports_every_declared_symbol(built_lib):
    out = subprocess.run(["nm", "-D", "--defined-only", built_lib], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_symbols() if n not in exported]
    assert not missing, "declared but not exported: %s" % missing


def test_library_loads_with_ctypes(built_lib):
    L = ctypes.CDLL(built_lib)
    for n in declared_symbols():
        assert getattr(L, n) is not None


def test_package_binds_every_declared_symbol(built_lib):
    import parparpar_amd as P
    P.lib()
    assert set(declared_symbols()) <= set(P.exported_symbols())


def test_version_without_gpu(built_lib):
    import parparpar_amd as P
    assert P.lib().parhip_version() >= 100


def test_library_links_no_oracle(built_lib):
    # the product must not depend on the oracle
    out = subprocess.run(["ldd", built_lib], stdout=subprocess.PIPE, text=True).stdout
    assert "oracle" not in out and "libref" not in out
    srcs = glob.glob(os.path.join(ROOT, "parparpar_amd", "**", "*.*"), recursive=True)
    for s in srcs:
        if s.endswith((".py", ".hip", ".cpp", ".h")):
            text = open(s, errors="ignore").read()
            assert "import oracle" not in text and "from oracle" not in text, s


@pytest.mark.parametrize("std,compiler", [("c99", "gcc"), ("c++11", "g++")])
def test_headers_compile_as_c_and_cxx(tmp_path, std, compiler):
    """Every public header is plain C (for cgo / N-API C callers) and C++."""
    src = tmp_path / "use_headers.c"
    lines = ['#include "%s"' % h for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))]
    lines.append("int main(void) { return PARHIP_OK; }")
    src.write_text("\n".join(lines) + "\n")
    cmd = [compiler, "-std=" + std, "-Wall", "-Wextra", "-Werror", "-fsyntax-only"]
    if compiler == "gcc":
        cmd.append("-pedantic")
    else:
        cmd += ["-x", "c++"]
    r = subprocess.run(cmd + [str(src)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_c_program_links_against_library(tmp_path, built_lib):
    """A C caller links libparhip.so by name and resolves the entry points it uses."""
    src = tmp_path / "caller.c"
    src.write_text(
        '#include "parhip.h"\n#include "parhip_par2.h"\n#include "parhip_batch.h"\n'
        '#include "parhip_repair.h"\n'
        "int main(void) {\n"
        "  void *fns[] = {(void *)parhip_gf64_compute_recovery_blocks,\n"
        "                 (void *)parhip_par3_create_streaming,\n"
        "                 (void *)parhip_assemble_recovery_packets,\n"
        "                 (void *)parhip_gf16_rec_matrix_compute,\n"
        "                 (void *)parhip_batch_create, (void *)parhip_repair_server_create};\n"
        "  return fns[0] == 0;\n}\n")
    exe = tmp_path / "caller"
    pkg = os.path.join(ROOT, "parparpar_amd")
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", pkg, "-lparhip", "-Wl,-rpath," + pkg, "-Wl,--allow-shlib-undefined"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
