"""ctypes front of the oracle (CPU restatement) and of the reference engine
compiled in place under oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
cpu_baseline leg of bench.py, always to check results, never as the thing
measured or shipped.  The product package never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_par3.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64

_lib = None
_ref = None


def build():
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "ref"])
        # the reference controller over the PAR2ProcHIP backend (links the engine)
        if os.path.exists(os.path.join(os.path.dirname(HERE), "parparpar_amd", "libparhip.so")):
            subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "backend"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.or_gf64_mul.restype = _u64
        L.or_gf64_mul.argtypes = [_u64, _u64]
        L.or_gf64_mul_peasant.restype = _u64
        L.or_gf64_mul_peasant.argtypes = [_u64, _u64]
        L.or_gf64_inv.restype = _u64
        L.or_gf64_inv.argtypes = [_u64]
        L.or_gf64_inv_fermat.restype = _u64
        L.or_gf64_inv_fermat.argtypes = [_u64]
        L.or_gf64_build_cauchy.argtypes = [_u64p, _sz, _sz, _u64, _u64]
        L.or_gf64_region_mul.argtypes = [_u64p, _u64p, _sz, _u64]
        L.or_gf64_region_mul_arr.argtypes = [_u64p, _u64p, _u64p, _sz, _sz]
        L.or_gf64_region_muladd_arr.argtypes = [_u64p, _u64p, _u64p, _sz, _sz]
        L.or_gf64_region_coupled_muladd_arr.argtypes = [_u64p, ctypes.POINTER(_u64p), _u64p, _sz, _sz]
        L.or_gf64_region_fused_output_muladd_arr.argtypes = [ctypes.POINTER(_u64p), _u64p,
                                                             ctypes.POINTER(_u64p), _sz, _sz]
        L.or_gf64_region_2d_muladd_arr.argtypes = [ctypes.POINTER(_u64p), _sz, ctypes.POINTER(_u64p),
                                                   _sz, _u64p, _sz, _sz]
        L.or_gf64_compute_recovery_with_coeff.argtypes = [_u64p, _sz, _u64p, _sz, _sz, _u64p, ctypes.c_int]
        L.or_gf64_compute_recovery.argtypes = [_u64p, _sz, _u64p, _sz, _sz, _u64, _u64, ctypes.c_int]
        L.or_gf64_recovery_columns.argtypes = [_u64p, _sz, _sz, _u64p, _sz, _u64p, _sz, _u64p]
        L.or_gf64_solve_and_reconstruct.restype = ctypes.c_int
        L.or_gf64_solve_and_reconstruct.argtypes = [_u64p, _u64p, _sz, _sz]
        L.or_gf64_solve.restype = ctypes.c_int
        L.or_gf64_solve.argtypes = [_u64p, _u64p, _u64p, _sz]
        L.or_gf16_mul.restype = ctypes.c_uint16
        L.or_gf16_mul.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.or_gf16_mul_peasant.restype = ctypes.c_uint16
        L.or_gf16_mul_peasant.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.or_gf16_exp.restype = ctypes.c_uint16
        L.or_gf16_exp.argtypes = [ctypes.c_uint32]
        L.or_gf16_input_log.restype = ctypes.c_uint16
        L.or_gf16_input_log.argtypes = [ctypes.c_uint32]
        L.or_gf16_coeff.restype = ctypes.c_uint16
        L.or_gf16_coeff.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.or_gf16_muladd_region.argtypes = [_u8p, _u8p, _sz, ctypes.c_uint16]
        L.or_par2_recovery_columns.argtypes = [_u8p, _sz, _sz, _u16p, _u16p, _sz, _u64p, _sz, _u16p]
        L.or_crc32.restype = ctypes.c_uint32
        L.or_crc32.argtypes = [_u8p, _sz]
        L.or_blake3.argtypes = [_u8p, _sz, _u8p]
        L.or_gf16_rec_matrix_compute.restype = ctypes.c_int
        L.or_gf16_rec_matrix_compute.argtypes = [_u8p, _sz, _sz, _u16p, ctypes.POINTER(_sz), _u16p]
        L.or_assemble_recovery_packets.argtypes = [_u8p, _u8p, _u64, _sz, _sz, _u8p]
        L.bl_par2_create.restype = ctypes.c_int
        L.bl_par2_create.argtypes = [_u16p, _sz, _sz, _u16p, _u16p, _sz, _u16p, ctypes.c_int]
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(REF_PATH)


def ref():
    """The reference engine compiled from its own sources (oracle/_ref)."""
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError("oracle/_ref/libref_par3.so is not built")
        R = ctypes.CDLL(REF_PATH)
        R.ref_init.restype = ctypes.c_int
        R.ref_method.restype = ctypes.c_int
        R.ref_gf64_mul.restype = _u64
        R.ref_gf64_mul.argtypes = [_u64, _u64]
        R.ref_gf64_inv.restype = _u64
        R.ref_gf64_inv.argtypes = [_u64]
        R.ref_build_cauchy.argtypes = [_u64p, _sz, _sz, _u64, _u64]
        R.ref_compute_recovery.argtypes = [_u64p, _sz, _u64p, _sz, _sz, _u64, _u64, ctypes.c_int, ctypes.c_int]
        R.ref_compute_recovery_with_coeff.argtypes = [_u64p, _sz, _u64p, _sz, _sz, _u64p, ctypes.c_int,
                                                      ctypes.c_int]
        R.ref_blake3_2.argtypes = [_u8p, _sz, _u8p, _sz, _u8p]
        R.ref_solve_and_reconstruct.restype = ctypes.c_int
        R.ref_solve_and_reconstruct.argtypes = [_u64p, _u64p, _sz, _sz, ctypes.c_int]
        R.gfmat_init.restype = None
        R.gfmat_coeff.restype = ctypes.c_uint16
        R.gfmat_coeff.argtypes = [ctypes.c_uint, ctypes.c_uint]
        R.gfmat_input_log.restype = ctypes.c_uint16
        R.gfmat_input_log.argtypes = [ctypes.c_uint]
        R.ref_init()
        R.gfmat_init()
        _ref = R
    return _ref


def p64(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def p16(a):
    assert a.dtype == np.uint16 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u16p)


def p8(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)


# ----------------------------- GF(2^64) -----------------------------------

def gf64_mul(a, b):
    return lib().or_gf64_mul(a, b)


def gf64_inv(a):
    return lib().or_gf64_inv(a)


def cauchy(n, r, first_input, first_recovery):
    out = np.empty((r, n), dtype=np.uint64)
    lib().or_gf64_build_cauchy(p64(out), n, r, first_input, first_recovery)
    return out


def recovery_with_coeff(inputs, coeff, recovery=None, accumulate=False):
    """inputs: (N, W) uint64; coeff: (R, N) uint64 -> (R, W) uint64."""
    n, w = inputs.shape
    r = coeff.shape[0]
    if recovery is None:
        recovery = np.zeros((r, w), dtype=np.uint64)
    lib().or_gf64_compute_recovery_with_coeff(p64(inputs), n, p64(recovery), r, w, p64(coeff),
                                              1 if accumulate else 0)
    return recovery


def recovery(inputs, r, first_input, first_recovery):
    n, w = inputs.shape
    out = np.zeros((r, w), dtype=np.uint64)
    lib().or_gf64_compute_recovery(p64(inputs), n, p64(out), r, w, first_input, first_recovery, 0)
    return out


def recovery_columns(inputs, coeff, cols):
    n, w = inputs.shape
    r = coeff.shape[0]
    cols = np.ascontiguousarray(cols, dtype=np.uint64)
    out = np.empty((r, len(cols)), dtype=np.uint64)
    lib().or_gf64_recovery_columns(p64(inputs), n, w, p64(coeff), r, p64(cols), len(cols), p64(out))
    return out


def solve_and_reconstruct(A, rhs):
    A = np.ascontiguousarray(A, dtype=np.uint64).copy()
    rhs = np.ascontiguousarray(rhs, dtype=np.uint64).copy()
    n = A.shape[0]
    rc = lib().or_gf64_solve_and_reconstruct(p64(A), p64(rhs), n, rhs.shape[1] if rhs.ndim > 1 else 1)
    return rc, A, rhs


def crc32(buf):
    b = np.frombuffer(bytes(buf), dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf.view(np.uint8)
    b = np.ascontiguousarray(b)
    return lib().or_crc32(p8(b), b.size)


# ----------------------------- GF(2^16) -----------------------------------

def gf16_coeff(input_num, exponent):
    return lib().or_gf16_coeff(input_num, exponent)


def par2_columns(inputs_u8, slice_size, input_nums, exponents, cols):
    """inputs_u8: (N, slice_size) uint8 -> (R, len(cols)) uint16 at 16-bit word columns."""
    n = inputs_u8.shape[0]
    inums = np.ascontiguousarray(input_nums, dtype=np.uint16)
    exps = np.ascontiguousarray(exponents, dtype=np.uint16)
    cols = np.ascontiguousarray(cols, dtype=np.uint64)
    out = np.empty((len(exps), len(cols)), dtype=np.uint16)
    lib().or_par2_recovery_columns(p8(inputs_u8), n, slice_size, p16(inums), p16(exps), len(exps),
                                   p64(cols), len(cols), p16(out))
    return out


def par2_create(inputs, input_nums, exponents, slice_size):
    """inputs: list of uint8 arrays (each <= slice_size bytes) -> (R, slice_size) uint8."""
    L = lib()
    outs = np.zeros((len(exponents), slice_size + 1), dtype=np.uint8)  # +1: odd-tail scratch
    for o, e in enumerate(exponents):
        for i, buf in enumerate(inputs):
            b = np.ascontiguousarray(buf, dtype=np.uint8)
            row = outs[o]
            L.or_gf16_muladd_region(row.ctypes.data_as(_u8p), p8(b), b.size,
                                    L.or_gf16_coeff(int(input_nums[i]), int(e)))
    return np.ascontiguousarray(outs[:, :slice_size])


def _bytes_ptr(b):
    a = np.frombuffer(bytes(b), dtype=np.uint8) if len(b) else np.zeros(1, dtype=np.uint8)
    return a, a.ctypes.data_as(_u8p)


def blake3(data):
    """32-byte BLAKE3 digest of `data` (bytes)."""
    a, p = _bytes_ptr(data)
    out = np.zeros(32, dtype=np.uint8)
    lib().or_blake3(p, len(data), out.ctypes.data_as(_u8p))
    return out.tobytes()


def ref_blake3(a, b=b""):
    """The reference's BLAKE3 over a || b (oracle/_ref), as its packetizer calls it."""
    x, px = _bytes_ptr(a)
    y, py = _bytes_ptr(b)
    out = np.zeros(32, dtype=np.uint8)
    ref().ref_blake3_2(px, len(a), py, len(b), out.ctypes.data_as(_u8p))
    return out.tobytes()


def assemble_recovery_packets(recovery, input_set_id, first_index, num, block_size):
    """numRecovery PAR3 REC packets of 64 + block_size bytes each."""
    assert len(recovery) == num * block_size and len(input_set_id) == 8
    r, pr = _bytes_ptr(recovery)
    i, pi = _bytes_ptr(input_set_id)
    out = np.zeros(num * (64 + block_size), dtype=np.uint8)
    lib().or_assemble_recovery_packets(pr, pi, first_index, num, block_size, out.ctypes.data_as(_u8p))
    return out.tobytes()


def gf16_rec_matrix(input_valid, recovery):
    """Galois16RecMatrix::Compute: (ok, used exponents, factors[numRec, numInputs])."""
    valid = np.ascontiguousarray(np.asarray(input_valid, dtype=np.uint8))
    n = valid.size
    vc = int(valid.sum())
    rec = np.ascontiguousarray(np.asarray(recovery, dtype=np.uint16)).copy()
    cnt = ctypes.c_size_t(rec.size)
    out = np.zeros((n - vc, n), dtype=np.uint16)
    ok = lib().or_gf16_rec_matrix_compute(valid.ctypes.data_as(_u8p), n, vc, p16(rec),
                                          ctypes.byref(cnt), p16(out))
    return bool(ok), rec[:cnt.value].copy(), out


def baseline_par2_create(inputs, input_nums, exponents, threads):
    """CPU baseline of the PAR2 create (baseline_gf16.c): inputs (N, words)
    uint16; returns ((R, words) uint16, threads used).  bench.py times it."""
    x = np.ascontiguousarray(inputs, dtype=np.uint16)
    n, words = x.shape
    inums = np.ascontiguousarray(input_nums, dtype=np.uint16)
    exps = np.ascontiguousarray(exponents, dtype=np.uint16)
    out = np.zeros((exps.size, words), dtype=np.uint16)
    used = lib().bl_par2_create(p16(x), n, words, p16(inums), p16(exps), exps.size, p16(out),
                                int(threads))
    return out, used
