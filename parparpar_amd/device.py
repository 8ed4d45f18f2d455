"""Device-resident entry points over torch tensors (HBM-resident slices).

torch is used here only for device memory and streams; all computation is in
libparhip.so.  Slices are int64 tensors of shape (count, words) holding raw
little-endian 64-bit words (the bit patterns, not numbers).
"""
import torch

from . import check, lib


def _stream(stream):
    s = torch.cuda.current_stream() if stream is None else stream
    return s.cuda_stream


def _words(t, name):
    if not t.is_cuda:
        raise TypeError("%s must be a device tensor" % name)
    if t.dtype not in (torch.int64, torch.uint64):
        raise TypeError("%s must be int64 words" % name)
    if t.dim() != 2 or t.stride(1) != 1:
        raise TypeError("%s must be a 2-D row-contiguous (count, words) tensor" % name)
    return t


def build_cauchy(num_inputs, num_recovery, first_input, first_recovery, device="cuda",
                 stream=None):
    out = torch.empty((num_recovery, num_inputs), dtype=torch.int64, device=device)
    check(lib().parhip_d_gf64_build_cauchy(out.data_ptr(), num_inputs, num_recovery, first_input,
                                           first_recovery, _stream(stream)))
    return out


def recovery(inputs, num_recovery, first_input, first_recovery, out=None, accumulate=False,
             stream=None):
    """R[r] (^)= XOR_i cauchy(r, i) * inputs[i]; Cauchy matrix cached on the device."""
    _words(inputs, "inputs")
    n, w = inputs.shape
    if out is None:
        out = torch.empty((num_recovery, w), dtype=torch.int64, device=inputs.device)
    _words(out, "out")
    check(lib().parhip_d_gf64_recovery(inputs.data_ptr(), n, inputs.stride(0), out.data_ptr(),
                                       num_recovery, out.stride(0), w, first_input, first_recovery,
                                       1 if accumulate else 0, _stream(stream)))
    return out


def recovery_with_coeff(inputs, coeff, out=None, accumulate=False, stream=None):
    _words(inputs, "inputs")
    n, w = inputs.shape
    r = coeff.shape[0]
    if coeff.shape[1] != n or not coeff.is_contiguous():
        raise ValueError("coeff must be a contiguous (R, N) tensor")
    if out is None:
        out = torch.empty((r, w), dtype=torch.int64, device=inputs.device)
    _words(out, "out")
    check(lib().parhip_d_gf64_recovery_with_coeff(inputs.data_ptr(), n, inputs.stride(0),
                                                  out.data_ptr(), r, out.stride(0), w,
                                                  coeff.data_ptr(), 1 if accumulate else 0,
                                                  _stream(stream)))
    return out


def recovery_ptrs(in_ptrs, out_ptrs, words, coeff, accumulate=True, stream=None):
    """Scattered slices: in_ptrs/out_ptrs are int64 device tensors of device addresses."""
    n = in_ptrs.numel()
    r = out_ptrs.numel()
    check(lib().parhip_d_gf64_recovery_ptrs(in_ptrs.data_ptr(), n, out_ptrs.data_ptr(), r, words,
                                            coeff.data_ptr(), coeff.stride(0),
                                            1 if accumulate else 0, _stream(stream)))


# ------------------------------------------------------------ PAR2 GF(2^16)

def gf16_coeffs(input_nums, exponents, device="cuda", stream=None):
    """(R, N) int16 device tensor of PAR2 coefficients (bit patterns)."""
    import numpy as np
    inums = np.ascontiguousarray(input_nums, dtype=np.uint16)
    exps = np.ascontiguousarray(exponents, dtype=np.uint16)
    out = torch.empty((exps.size, inums.size), dtype=torch.int16, device=device)
    check(lib().parhip_d_gf16_coeffs(out.data_ptr(), inums.ctypes.data, inums.size,
                                     exps.ctypes.data, exps.size, _stream(stream)))
    return out


def gf16_muladd(inputs, coeff, out=None, words=None, accumulate=False, stream=None):
    """out[e] (^)= XOR_i coeff[e, i] * inputs[i] over 16-bit words.
    inputs: (N, >=words) int16 device tensor; coeff: (R, N) int16 device tensor."""
    if inputs.dtype != torch.int16 or inputs.dim() != 2 or inputs.stride(1) != 1:
        raise TypeError("inputs must be a row-contiguous (N, words) int16 device tensor")
    n, w = inputs.shape
    words = w if words is None else words
    r = coeff.shape[0]
    if coeff.shape[1] != n or not coeff.is_contiguous() or coeff.dtype != torch.int16:
        raise ValueError("coeff must be a contiguous (R, N) int16 tensor")
    if out is None:
        out = torch.empty((r, words), dtype=torch.int16, device=inputs.device)
    check(lib().parhip_d_gf16_muladd(inputs.data_ptr(), n, inputs.stride(0), out.data_ptr(), r,
                                     out.stride(0), words, coeff.data_ptr(), 1 if accumulate else 0,
                                     _stream(stream)))
    return out


def assemble_recovery_packets(recovery, input_set_id, first_index, out=None, stream=None):
    """PAR3 REC packets from a (R, S) uint8 device tensor of recovery blocks:
    returns a (R, 64 + S) uint8 device tensor (enqueued on `stream`)."""
    r, bs = recovery.shape
    if recovery.dtype != torch.uint8 or not recovery.is_contiguous():
        raise TypeError("recovery must be a contiguous uint8 tensor")
    if out is None:
        out = torch.empty((r, 64 + bs), dtype=torch.uint8, device=recovery.device)
    isid = bytes(input_set_id)
    check(lib().parhip_d_assemble_recovery_packets(recovery.data_ptr(), isid, int(first_index), r, bs,
                                                   out.data_ptr(), _stream(stream)))
    return out
