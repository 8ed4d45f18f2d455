"""parparpar_amd -- MI355X (gfx950) recovery-slice engine for PAR3 / PAR2.

The engine is the in-tree shared library ``parparpar_amd/libparhip.so`` (built
by ``python -m parparpar_amd.build``) behind the C ABI in ``include/parhip.h``.
This package only loads it and gives Python the same call shapes the
reference's N-API addon exposes (``parparpar_amd.binding``); there is no
Python or CPU fallback for any computation: if the library is missing every
call raises.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libparhip.so")

_lib = None

_u64 = ctypes.c_uint64
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u64pp = ctypes.POINTER(_u64p)
_sz = ctypes.c_size_t
_int = ctypes.c_int
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every C-ABI entry point the package binds.
_SIGS = [
    ("parhip_version", _int, []),
    ("parhip_last_error", ctypes.c_char_p, []),
    ("parhip_set_device", _int, [_int]),
    ("parhip_set_method", _int, [_int]),
    ("parhip_release_cache", _int, []),
    ("parhip_set_stream_chunk", _int, [_sz]),
    ("parhip_set_coeff_cache", _int, [_int]),
    ("parhip_gf16_rec_matrix_compute", _int,
     [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.POINTER(ctypes.c_uint32), _vp]),
    ("parhip_d_gf16_rec_matrix_compute", _int,
     [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.POINTER(ctypes.c_uint32), _vp, _vp]),
    ("parhip_assemble_recovery_packets", _int,
     [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp]),
    ("parhip_d_assemble_recovery_packets", _int,
     [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp, _vp]),
    ("parhip_gf64_compute_recovery_blocks_from_file", _int, [ctypes.c_char_p, _vp, _sz, _sz, _u64, _u64, _int]),
    ("parhip_par3_create_streaming", _int,
     [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, _u64, _u64, ctypes.c_int32, _int,
      ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)), ctypes.POINTER(_sz),
      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    ("parhip_free", None, [_vp]),
    ("parhip_gf64_mul", _u64, [_u64, _u64]),
    ("parhip_gf64_inverse", _u64, [_u64]),
    ("parhip_gf64_compute_recovery_blocks", _int, [_vp, _sz, _vp, _sz, _sz, _u64, _u64, _int, _int]),
    ("parhip_gf64_compute_recovery_blocks_full", _int, [_vp, _sz, _vp, _sz, _sz, _u64, _u64, _int, _int]),
    ("parhip_gf64_compute_recovery_blocks_barycentric", _int,
     [_vp, _sz, _vp, _sz, _sz, _u64, _u64, _int, _int]),
    ("parhip_gf64_compute_recovery_blocks_fenger", _int, [_vp, _sz, _vp, _sz, _sz, _u64, _u64, _int]),
    ("parhip_gf64_compute_recovery_batches", _int,
     [_vp, _vp, _sz, _vp, _sz, _sz, _sz, _u64, _u64, _int]),
    ("parhip_gf64_compute_recovery_blocks_with_coeff", _int, [_vp, _sz, _vp, _sz, _sz, _vp, _int, _int]),
    ("parhip_gf64_build_cauchy_matrix", _int, [_vp, _sz, _sz, _u64, _u64]),
    ("parhip_gf64_build_cauchy_matrix_alloc", _vp, [_sz, _sz, _u64, _u64]),
    ("parhip_gf64_multiply_accumulate", _int, [_vp, _sz, _vp, _sz, _vp, _sz]),
    ("parhip_gf64_solve_and_reconstruct", _int, [_vp, _vp, _sz, _sz, _int]),
    ("parhip_gf64_solve", _int, [_vp, _vp, _vp, _sz]),
    ("parhip_gf64_region_mul", _int, [_vp, _vp, _sz, _u64]),
    ("parhip_gf64_region_mul_arr", _int, [_vp, _vp, _vp, _sz, _sz]),
    ("parhip_gf64_region_muladd_arr", _int, [_vp, _vp, _vp, _sz, _sz]),
    ("parhip_gf64_region_coupled_muladd_arr", _int, [_vp, _vp, _vp, _sz, _sz]),
    ("parhip_gf64_region_fused_output_muladd_arr", _int, [_vp, _vp, _vp, _sz, _sz]),
    ("parhip_gf64_region_2d_muladd_arr", _int, [_vp, _sz, _vp, _sz, _vp, _sz, _sz]),
    ("parhip_d_gf64_build_cauchy", _int, [_vp, _sz, _sz, _u64, _u64, _vp]),
    ("parhip_d_gf64_recovery_with_coeff", _int, [_vp, _sz, _sz, _vp, _sz, _sz, _sz, _vp, _int, _vp]),
    ("parhip_d_gf64_recovery", _int, [_vp, _sz, _sz, _vp, _sz, _sz, _sz, _u64, _u64, _int, _vp]),
    ("parhip_d_gf64_recovery_static", _int, [_vp, _sz, _sz, _vp, _sz, _sz, _sz, _u64, _u64, _int, _vp]),
    ("parhip_d_gf64_recovery_ptrs", _int, [_vp, _sz, _vp, _sz, _sz, _vp, _sz, _int, _vp]),
    ("parhip_gf64_workspace_bytes", _sz, [_sz, _sz, _sz]),
    ("parhip_tune_override", _int, [_int, _int, _int, ctypes.c_int64]),
    ("parhip_gf16_coeff", ctypes.c_uint16, [ctypes.c_uint16, ctypes.c_uint16]),
    ("parhip_gf16_mul", ctypes.c_uint16, [ctypes.c_uint16, ctypes.c_uint16]),
    ("parhip_d_gf16_coeffs", _int, [_vp, _vp, _sz, _vp, _sz, _vp]),
    ("parhip_d_gf16_muladd", _int, [_vp, _sz, _sz, _vp, _sz, _sz, _sz, _vp, _int, _vp]),
    ("parhip_par2_create", _vp, [_int, _sz, ctypes.c_uint]),
    ("parhip_par2_create_areas", _vp, [_int, _sz, ctypes.c_uint, ctypes.c_uint]),
    ("parhip_par2_destroy", None, [_vp]),
    ("parhip_par2_set_recovery_slices", _int, [_vp, ctypes.c_uint, _vp]),
    ("parhip_par2_set_current_slice_size", _int, [_vp, _sz]),
    ("parhip_par2_can_add", _int, [_vp]),
    ("parhip_par2_add_input", _int, [_vp, _vp, _sz, ctypes.c_uint16, _int]),
    ("parhip_par2_add_input_coeffs", _int, [_vp, _vp, _sz, _vp, _int]),
    ("parhip_par2_fill_input", _int, [_vp, _vp]),
    ("parhip_par2_dummy_input", _int, [_vp, ctypes.c_uint16, _int]),
    ("parhip_par2_flush", _int, [_vp]),
    ("parhip_par2_end_input", _int, [_vp]),
    ("parhip_par2_get_output", _int, [_vp, ctypes.c_uint, _vp]),
    ("parhip_par2_discard_output", _int, [_vp]),
    ("parhip_par2_input_batch_size", _int, [_vp]),
    ("parhip_par2_set_slice_size", _int, [_vp, _sz]),
    ("parhip_par2_wait_for_add", _int, [_vp]),
    ("parhip_par2_add_input_async", _int, [_vp, _vp, _sz, ctypes.c_uint16, _int, _vp, _vp]),
    ("parhip_par2_add_input_coeffs_async", _int, [_vp, _vp, _sz, _vp, _int, _vp, _vp]),
    ("parhip_par2_set_batch_callbacks", _int, [_vp, _vp, _vp, _vp]),
    ("parhip_par2_get_output_async", _int, [_vp, ctypes.c_uint, _vp, _vp, _vp]),
    ("parhip_par2_set_min_input_batch_size", _int, [_vp, ctypes.c_uint]),
    ("parhip_par2_batches_started", _int, [_vp]),
    ("parhip_par2_staging_areas", _int, [_vp]),
    ("parhip_par2_free_processing_mem", _int, [_vp]),
    ("parhip_par2_output_device_ptr", _int, [_vp, ctypes.c_uint, ctypes.POINTER(_vp), ctypes.POINTER(_sz)]),
    ("parhip_batch_create", _vp, [_vp, _sz, ctypes.c_uint]),
    ("parhip_batch_run", _int, [_vp, _vp]),
    ("parhip_batch_info", _int, [_vp, _vp, _vp, _vp]),
    ("parhip_batch_job_config", _int, [_vp, _sz, _vp, _vp, _vp]),
    ("parhip_batch_job_method", _int, [_vp, _sz]),
    ("parhip_batch_destroy", None, [_vp]),
    ("parhip_autotune_lookup", _int, [_sz, _sz, _sz, _vp, _vp, _vp, _vp]),
    ("parhip_autotune_lookup_method", _int, [_sz, _sz, _sz]),
    ("parhip_autotune_save", _int, [ctypes.c_char_p]),
    ("parhip_autotune_load", _int, [ctypes.c_char_p]),
    ("parhip_autotune_clear", _int, []),
    ("parhip_d_gf64_repair", _int, [_vp, _sz, _sz, _vp, _sz, _sz, _sz, _u64, _u64, _vp, _sz, _vp, _vp, _sz, _vp]),
    ("parhip_gf64_repair_workspace_bytes", _sz, [_sz, _sz, _sz]),
    ("parhip_repair_server_create", _vp, [_sz, ctypes.c_uint, ctypes.c_uint]),
    ("parhip_repair_server_add_archive", _int, [_vp, _vp, _sz, _sz, _vp, _sz, _sz, _sz, _u64, _u64]),
    ("parhip_repair_submit", ctypes.c_int64, [_vp, _int, _vp, _sz]),
    ("parhip_repair_wait", _int, [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]),
    ("parhip_repair_result", _int, [_vp, ctypes.c_int64, ctypes.POINTER(_vp), ctypes.POINTER(_sz)]),
    ("parhip_repair_release", _int, [_vp, ctypes.c_int64]),
    ("parhip_repair_server_stats", _int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("parhip_repair_server_cache_stats", _int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("parhip_repair_server_destroy", None, [_vp]),
    ("parhip_prof_only", _int, [ctypes.c_char_p]),
    ("parhip_prof_enable", _int, [_int]),
    ("parhip_prof_reset", _int, []),
    ("parhip_prof_get", _int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_u64)]),
    ("parhip_measure_lds_gather", _int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
]

PARHIP_OK = 0
PARHIP_ERR_SINGULAR = -1
PARHIP_ERR_ARG = -2
PARHIP_ERR_NOMEM = -3
PARHIP_ERR_DEVICE = -4
PARHIP_ERR_RANGE = -5

METHOD_AUTO = 0
METHOD_TABLE = 1
METHOD_MFMA = 2
METHOD_MFMA_I8 = 3
METHOD_TABLE8 = 4
METHOD_DYADIC = 5
METHOD_DYADIC8 = 6
METHOD_DYADIC16 = 7
METHOD_DYADIC32 = 8


class ParhipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("parhip error %d: %s" % (code, msg))
        self.code = code


def exported_symbols():
    return [name for name, _, _ in _SIGS]


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("parparpar_amd/libparhip.so is not built; run "
                               "`python -m parparpar_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def set_method(method):
    """Contraction strategy for the calling thread (METHOD_*)."""
    check(lib().parhip_set_method(int(method)))


def prof_enable(on=True):
    check(lib().parhip_prof_enable(1 if on else 0))


def prof_only(kernel=None):
    """Time only the kernels named `kernel` (None: all)."""
    check(lib().parhip_prof_only(kernel.encode() if kernel else None))


def prof_reset():
    check(lib().parhip_prof_reset())


def prof_get(kernel):
    """(total_ms, launches) of a named kernel on this thread since the last reset."""
    ms = ctypes.c_double(0)
    n = _u64(0)
    check(lib().parhip_prof_get(kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def check(rc):
    if rc != PARHIP_OK:
        msg = lib().parhip_last_error()
        raise ParhipError(rc, msg.decode() if msg else "")
    return rc
