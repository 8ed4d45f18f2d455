"""Build the gfx950 engine library in-tree: parparpar_amd/libparhip.so.

    python -m parparpar_amd.build

HIP sources are compiled with hipcc for gfx950 only; host-only C++ with g++.
Every HIP object is built with the packed-FP32 target feature disabled
(no v_pk_*_f32 generated).  Objects go to parparpar_amd/_build/ (ignored).
"""
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libparhip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I" + CSRC]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + CSRC,
              "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]


def sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    host = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, host


def _compile(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return cmd, r.returncode, r.stdout


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    hip, host = sources()
    cmds, objs = [], []
    for f in hip:
        o = os.path.join(BUILD, f + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
        objs.append(o)
    for f in host:
        o = os.path.join(BUILD, f + ".o")
        # host-only C++ (HIP runtime API calls, no kernels): plain g++
        cmds.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
        objs.append(o)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for cmd, rc, out in ex.map(_compile, cmds):
            if verbose or rc:
                print(" ".join(cmd))
                print(out)
            if rc:
                raise RuntimeError("compile failed: " + cmd[-3])
    tmp = LIB + ".tmp"
    link = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    cmd, rc, out = _compile(link)
    if rc:
        print(" ".join(cmd))
        print(out)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
