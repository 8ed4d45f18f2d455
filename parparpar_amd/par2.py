"""PAR2 recovery over the PAR2 backend engine (parhip_par2_*), in the call shapes of the
reference's GfProc (src/gf.cc:239-735, used by lib/par2.js:112-228).

    gp = GfProc(sliceSize)                  # new binding.GfProc(sliceSize, ...)
    gp.set_recovery_slices([0, 1, 2, ...])  # exponents (lib/par2.js:210-228)
    gp.add(inputNum, buf)                   # gf.add(idx, buf, cb)
    gp.end()                                # gf.end(cb)
    valid = gp.get(idx, out)                # gf.get(idx, buf, cb(idx, cksumValid, buf))

All work runs in libparhip.so; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import ParhipError, check, lib

# void (*)(void* user, int cksumValid)
_VALID_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int)


class GfProc:
    def __init__(self, slice_size, device=0, input_batch=0, staging_areas=2):
        self._h = lib().parhip_par2_create_areas(device, slice_size, input_batch, staging_areas)
        if not self._h:
            raise ParhipError(-4, lib().parhip_last_error().decode())
        self.slice_size = slice_size
        self.num_outputs = 0

    def close(self):
        if self._h:
            lib().parhip_par2_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_recovery_slices(self, exponents):
        """exponents: list of recovery exponents, or an int n for custom-coefficient mode."""
        if isinstance(exponents, int):
            check(lib().parhip_par2_set_recovery_slices(self._h, exponents, None))
            self.num_outputs = exponents
        else:
            e = np.ascontiguousarray(exponents, dtype=np.uint16)
            check(lib().parhip_par2_set_recovery_slices(self._h, e.size, e.ctypes.data))
            self.num_outputs = e.size

    def set_current_slice_size(self, size):
        check(lib().parhip_par2_set_current_slice_size(self._h, size))
        self.slice_size = size

    def can_add(self):
        return lib().parhip_par2_can_add(self._h)

    def add(self, input_num, buf, flush=False):
        b = np.ascontiguousarray(buf).view(np.uint8)
        check(lib().parhip_par2_add_input(self._h, b.ctypes.data, b.size, input_num, 1 if flush else 0))

    def add_async(self, input_num, buf, flush=False):
        """Claim a slot and queue the copy; `buf` must stay unchanged until end()
        (no completion callback)."""
        b = np.ascontiguousarray(buf).view(np.uint8)
        check(lib().parhip_par2_add_input_async(self._h, b.ctypes.data, b.size, input_num,
                                                1 if flush else 0, None, None))

    def batches_started(self):
        return lib().parhip_par2_batches_started(self._h)

    def output_device_ptr(self, idx):
        """(device address, row stride in bytes) of recovery slice idx; the row
        holds the slice's 16-bit words, then its checksum word."""
        p = ctypes.c_void_p()
        st = ctypes.c_size_t()
        check(lib().parhip_par2_output_device_ptr(self._h, idx, ctypes.byref(p), ctypes.byref(st)))
        return p.value, st.value

    def add_coeffs(self, coeffs, buf, flush=False):
        b = np.ascontiguousarray(buf).view(np.uint8)
        c = np.ascontiguousarray(coeffs, dtype=np.uint16)
        if c.size < self.num_outputs:
            raise ValueError("need one coefficient per recovery slice")
        check(lib().parhip_par2_add_input_coeffs(self._h, b.ctypes.data, b.size, c.ctypes.data,
                                                 1 if flush else 0))

    def fill_input(self, buf):
        """Benchmarking: copy `buf` (the current slice size) into the next staging
        slot; True once every staging slot holds it."""
        b = np.ascontiguousarray(buf).view(np.uint8)
        if b.size < self.slice_size:
            raise ValueError("fill_input needs the current slice size (%d bytes)" % self.slice_size)
        rc = lib().parhip_par2_fill_input(self._h, b.ctypes.data)
        if rc < 0:
            check(rc)
        return rc == 1

    def dummy_input(self, input_num, flush=False):
        """Benchmarking: add input `input_num` over the data its staging slot holds."""
        check(lib().parhip_par2_dummy_input(self._h, input_num, 1 if flush else 0))

    def flush(self):
        check(lib().parhip_par2_flush(self._h))

    def end(self):
        check(lib().parhip_par2_end_input(self._h))

    def get(self, idx, out=None):
        """Copy recovery slice idx into `out` (uint8, slice_size); returns (cksumValid, out)."""
        if out is None:
            out = np.zeros(self.slice_size, dtype=np.uint8)
        rc = lib().parhip_par2_get_output(self._h, idx, out.ctypes.data)
        if rc < 0:
            check(rc)
        return rc == 1, out

    def get_all(self, out):
        """Every recovery slice into out (R, slice_size) uint8, read back on the
        engine's copy workers in parallel (asynchronous getOutput); True when
        every checksum verified."""
        import threading
        r = self.num_outputs
        if out.shape[0] < r or out.shape[1] < self.slice_size or not out.flags["C_CONTIGUOUS"]:
            raise ValueError("out must be a contiguous (R, slice_size) uint8 array")
        state = {"left": r, "valid": True}
        done = threading.Event()
        lock = threading.Lock()

        def on_valid(user, valid):
            with lock:
                state["valid"] = state["valid"] and bool(valid)
                state["left"] -= 1
                if state["left"] == 0:
                    done.set()
        cb = _VALID_FN(on_valid)
        for e in range(r):
            check(lib().parhip_par2_get_output_async(self._h, e, out[e].ctypes.data, cb, None))
        done.wait()
        return state["valid"]

    def discard_output(self):
        check(lib().parhip_par2_discard_output(self._h))

    @property
    def input_batch_size(self):
        return lib().parhip_par2_input_batch_size(self._h)


def gf16_coeff(input_num, exponent):
    return lib().parhip_gf16_coeff(input_num, exponent)


def rec_matrix_compute(input_valid, recovery):
    """Galois16RecMatrix::Compute on the device.

    input_valid: booleans per input; recovery: available exponents.  Returns
    (ok, exponents used, factors) with factors[row, col] = GetFactor(col, row)
    (valid inputs first, then missing, each in input order); ok is False when
    too few exponents remain after dropping non-invertible rows."""
    valid = np.ascontiguousarray(np.asarray(input_valid, dtype=np.uint8))
    n = int(valid.size)
    vc = int(valid.sum())
    rec = np.ascontiguousarray(np.asarray(recovery, dtype=np.uint16)).copy()
    cnt = ctypes.c_uint32(rec.size)
    out = np.zeros((n - vc, n), dtype=np.uint16)
    rc = lib().parhip_gf16_rec_matrix_compute(valid.ctypes.data, n, vc, rec.ctypes.data,
                                               ctypes.byref(cnt), out.ctypes.data)
    if rc < 0:
        check(rc)
    return rc == 1, rec[:cnt.value].copy(), out
