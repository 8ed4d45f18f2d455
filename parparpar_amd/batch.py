"""Batched PAR3 create over device-resident jobs (include/parhip_batch.h)."""
import ctypes

import torch

from . import ParhipError, check, lib

GRAPH = 1
AUTOTUNE = 2
MFMA = 4


class _Job(ctypes.Structure):
    _fields_ = [("d_inputs", ctypes.c_void_p), ("numInputs", ctypes.c_size_t),
                ("inputStride64", ctypes.c_size_t), ("d_recovery", ctypes.c_void_p),
                ("numRecovery", ctypes.c_size_t), ("recoveryStride64", ctypes.c_size_t),
                ("blockSize64", ctypes.c_size_t), ("firstInput", ctypes.c_uint64),
                ("firstRecovery", ctypes.c_uint64)]


class Batch:
    """jobs: list of (inputs (N, W) int64 tensor, recovery (R, W) int64 tensor,
    firstInput, firstRecovery)."""

    def __init__(self, jobs, graph=True, autotune=True, mfma=False):
        arr = (_Job * len(jobs))()
        for k, (inp, rec, fi, fr) in enumerate(jobs):
            n, w = inp.shape
            r = rec.shape[0]
            if rec.shape[1] < w or inp.stride(1) != 1 or rec.stride(1) != 1:
                raise ValueError("job %d: bad tensor layout" % k)
            arr[k] = _Job(inp.data_ptr(), n, inp.stride(0), rec.data_ptr(), r, rec.stride(0), w, fi, fr)
        self._keep = jobs
        flags = (GRAPH if graph else 0) | (AUTOTUNE if autotune else 0) | (MFMA if mfma else 0)
        self._h = lib().parhip_batch_create(ctypes.addressof(arr), len(jobs), flags)
        if not self._h:
            raise ParhipError(-4, lib().parhip_last_error().decode())

    def run(self, stream=None):
        s = torch.cuda.current_stream() if stream is None else stream
        check(lib().parhip_batch_run(self._h, s.cuda_stream))

    def info(self):
        t, a, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().parhip_batch_info(self._h, ctypes.byref(t), ctypes.byref(a), ctypes.byref(c)))
        return {"tiles": t.value, "atomic_jobs": a.value, "coeff_bytes": c.value}

    def job_config(self, k):
        rt, wpt, ib = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().parhip_batch_job_config(self._h, k, ctypes.byref(rt), ctypes.byref(wpt), ctypes.byref(ib)))
        return rt.value, wpt.value, ib.value

    def job_method(self, k):
        """METHOD_TABLE or METHOD_MFMA: the strategy job k runs on."""
        rc = lib().parhip_batch_job_method(self._h, k)
        if rc < 0:
            check(rc)
        return rc

    def close(self):
        if self._h:
            lib().parhip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def autotune_save(path):
    check(lib().parhip_autotune_save(path.encode()))


def autotune_load(path):
    check(lib().parhip_autotune_load(path.encode()))
