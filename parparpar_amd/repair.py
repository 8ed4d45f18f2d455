"""PAR3 repair on device-resident archives, and the repair server
(include/parhip_repair.h)."""
import ctypes

import numpy as np
import torch

from . import ParhipError, check, lib


def _s(stream):
    return (torch.cuda.current_stream() if stream is None else stream).cuda_stream


def repair(data, recovery, missing, first_input=0, first_recovery=None, rows=None, out=None,
           stream=None):
    """Rebuild the input slices `missing` of an archive.
    data: (N, W) int64 device tensor (lost rows may hold anything: they never
    contribute); recovery: (R, W)."""
    n, w = data.shape
    r = recovery.shape[0]
    fr = n if first_recovery is None else first_recovery
    miss = np.ascontiguousarray(missing, dtype=np.uint64)
    k = miss.size
    if out is None:
        out = torch.empty((k, w), dtype=torch.int64, device=data.device)
    rws = None if rows is None else np.ascontiguousarray(rows, dtype=np.uint64)
    check(lib().parhip_d_gf64_repair(data.data_ptr(), n, data.stride(0), recovery.data_ptr(), r,
                                     recovery.stride(0), w, first_input, fr, miss.ctypes.data, k,
                                     None if rws is None else rws.ctypes.data, out.data_ptr(),
                                     out.stride(0), _s(stream)))
    return out


class RepairServer:
    def __init__(self, hbm_budget_bytes, max_inflight=4, streams=0):
        self._h = lib().parhip_repair_server_create(hbm_budget_bytes, max_inflight, streams)
        if not self._h:
            raise ParhipError(-4, lib().parhip_last_error().decode())
        self._archives = []

    def add_archive(self, data, recovery, first_input=0, first_recovery=None):
        n, w = data.shape
        fr = n if first_recovery is None else first_recovery
        aid = lib().parhip_repair_server_add_archive(self._h, data.data_ptr(), n, data.stride(0),
                                                     recovery.data_ptr(), recovery.shape[0],
                                                     recovery.stride(0), w, first_input, fr)
        if aid < 0:
            check(aid)
        self._archives.append((data, recovery))
        return aid

    def submit(self, archive, missing):
        miss = np.ascontiguousarray(missing, dtype=np.uint64)
        rid = lib().parhip_repair_submit(self._h, archive, miss.ctypes.data, miss.size)
        if rid < 0:
            check(int(rid))
        return rid

    def wait(self, rid):
        lat = ctypes.c_double(0)
        check(lib().parhip_repair_wait(self._h, rid, ctypes.byref(lat)))
        return lat.value

    def result(self, rid):
        """(k, W) int64 tensor copy of the repaired slices."""
        p = ctypes.c_void_p()
        k = ctypes.c_size_t()
        check(lib().parhip_repair_result(self._h, rid, ctypes.byref(p), ctypes.byref(k)))
        return p.value, k.value

    def release(self, rid):
        check(lib().parhip_repair_release(self._h, rid))

    def stats(self):
        c, p50, p99 = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
        pk, pb = ctypes.c_uint(), ctypes.c_uint64()
        check(lib().parhip_repair_server_stats(self._h, ctypes.byref(c), ctypes.byref(p50),
                                               ctypes.byref(p99), ctypes.byref(pk), ctypes.byref(pb)))
        return {"completed": c.value, "p50_ms": p50.value, "p99_ms": p99.value,
                "peak_inflight": pk.value, "peak_bytes": pb.value}

    def cache_stats(self):
        """The resident inverse cache: hits, misses, entries, bytes, evictions."""
        v = [ctypes.c_uint64() for _ in range(5)]
        check(lib().parhip_repair_server_cache_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("hits", "misses", "entries", "bytes", "evictions"), (x.value for x in v)))

    def close(self):
        if self._h:
            lib().parhip_repair_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def copy_result(server, rid, w, device="cuda"):
    """Copy a finished request's output into a new tensor (and keep it)."""
    ptr, k = server.result(rid)
    out = torch.empty((k, w), dtype=torch.int64, device=device)
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = hip.hipMemcpy(out.data_ptr(), ptr, k * w * 8, 3)  # device to device
    if rc != 0:
        raise RuntimeError("hipMemcpy failed: %d" % rc)
    return out
