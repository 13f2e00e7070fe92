"""PyTorch-ROCm front end: the device-pointer entry points of libs2let_amd.so
on torch tensors (complex128, on the current HIP device and stream), and a
``torch.autograd.Function`` whose backward is the hand-written adjoint
transform (s2let_hip_analysis_adjoint_wav2lm / s2let_hip_synthesis_adjoint_lm2wav).

The library computes; torch only provides device memory and the stream.  A
missing library raises (there is no CPU fallback on this path).
"""
from __future__ import annotations

import ctypes

import torch

from . import FP64, MIXED, Parameters, check, lib, n_scal, n_wav, parameters  # noqa: F401

__all__ = ["analysis", "synthesis", "analysis_adjoint", "synthesis_adjoint", "analysis_real", "synthesis_real",
           "mw_alm2map", "mw_map2alm",
           "batch", "autotune", "WaveletAnalysis", "wavelet_analysis", "FP64", "MIXED", "parameters"]


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t):
    if not (t.is_cuda and t.dtype == torch.complex128 and t.is_contiguous()):
        raise ValueError("expected a contiguous complex128 tensor on the GPU")
    return ctypes.c_void_p(t.data_ptr())


def _empty(n, like):
    return torch.zeros(int(n), dtype=torch.complex128, device=like.device)


def analysis(flm, p: Parameters, mode=FP64, out=None):
    """s2let_analysis_lm2wav on device: returns (f_wav, f_scal)."""
    fw, fs = out if out is not None else (_empty(n_wav(p), flm), _empty(n_scal(p), flm))
    check(lib().s2let_hip_analysis_lm2wav(_dev(fw), _dev(fs), _dev(flm), ctypes.byref(p), int(mode), _stream()))
    return fw, fs


def synthesis(f_wav, f_scal, p: Parameters, mode=FP64, out=None):
    """s2let_synthesis_wav2lm on device: returns flm."""
    flm = out if out is not None else _empty(p.L * p.L, f_wav)
    check(lib().s2let_hip_synthesis_wav2lm(_dev(flm), _dev(f_wav), _dev(f_scal), ctypes.byref(p), int(mode),
                                           _stream()))
    return flm


def analysis_adjoint(f_wav, f_scal, p: Parameters, mode=FP64, out=None):
    """s2let_analysis_adjoint_wav2lm on device: returns flm."""
    flm = out if out is not None else _empty(p.L * p.L, f_wav)
    check(lib().s2let_hip_analysis_adjoint_wav2lm(_dev(flm), _dev(f_wav), _dev(f_scal), ctypes.byref(p), int(mode),
                                                  _stream()))
    return flm


def synthesis_adjoint(flm, p: Parameters, mode=FP64, out=None):
    """s2let_synthesis_adjoint_lm2wav on device: returns (f_wav, f_scal)."""
    fw, fs = out if out is not None else (_empty(n_wav(p), flm), _empty(n_scal(p), flm))
    check(lib().s2let_hip_synthesis_adjoint_lm2wav(_dev(fw), _dev(fs), _dev(flm), ctypes.byref(p), int(mode),
                                                   _stream()))
    return fw, fs


def _devr(t):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 tensor on the GPU")
    return ctypes.c_void_p(t.data_ptr())


def analysis_real(flm, p: Parameters, mode=FP64, out=None):
    """s2let_analysis_lm2wav_real on device (flm Hermitian complex128): returns real (f_wav, f_scal)."""
    if out is None:
        out = (torch.zeros(int(n_wav(p)), dtype=torch.float64, device=flm.device),
               torch.zeros(int(n_scal(p)), dtype=torch.float64, device=flm.device))
    fw, fs = out
    check(lib().s2let_hip_analysis_lm2wav_real(_devr(fw), _devr(fs), _dev(flm), ctypes.byref(p), int(mode),
                                               _stream()))
    return fw, fs


def synthesis_real(f_wav, f_scal, p: Parameters, mode=FP64, out=None):
    """s2let_synthesis_wav2lm_real on device (real f_wav, f_scal): returns the Hermitian flm."""
    flm = out if out is not None else torch.zeros(p.L * p.L, dtype=torch.complex128, device=f_wav.device)
    check(lib().s2let_hip_synthesis_wav2lm_real(_dev(flm), _devr(f_wav), _devr(f_scal), ctypes.byref(p), int(mode),
                                                _stream()))
    return flm


def batch(op, outs_a, outs_b, ins_a, ins_b, params, mode=FP64, graph=True):
    """A batch of maps of uneven shape (s2let_hip_batch_graph / s2let_hip_batch):
    op is one of the OP_* codes; map k uses params[k] and the k-th tensors of
    each list (outs_b / ins_b may be None for ops with one output / input).
    graph=True replays the batch as one cached HIP graph (captured on the first
    call with the same shapes and tensors)."""
    n = len(params)
    P = (Parameters * n)(*params)

    def arr(ts):
        if ts is None:
            return None
        return (ctypes.c_void_p * n)(*[_dev(t).value for t in ts])

    fn = lib().s2let_hip_batch_graph if graph else lib().s2let_hip_batch
    check(fn(int(op), n, arr(outs_a), arr(outs_b), arr(ins_a), arr(ins_b), P, int(mode), _stream()))


def autotune(p: Parameters, mode=FP64):
    """s2let_hip_autotune on the current stream: picks launch shapes for p's band-limits."""
    check(lib().s2let_hip_autotune(ctypes.byref(p), int(mode), _stream()))


def mw_alm2map(flm, L, spin=0, mode=FP64):
    f = _empty(L * (2 * L - 1), flm)
    check(lib().s2let_hip_mw_alm2map(_dev(f), _dev(flm), int(L), int(spin), int(mode), _stream()))
    return f


def mw_map2alm(f, L, spin=0, mode=FP64):
    flm = _empty(L * L, f)
    check(lib().s2let_hip_mw_map2alm(_dev(flm), _dev(f), int(L), int(spin), int(mode), _stream()))
    return flm


class WaveletAnalysis(torch.autograd.Function):
    """W = A flm (s2let_analysis_lm2wav).  For a real loss, the gradient with
    respect to flm is A^H applied to the output gradients, computed by the
    hand-written analysis adjoint kernels."""

    @staticmethod
    def forward(ctx, flm, p, mode):
        ctx.p, ctx.mode = p, mode
        return analysis(flm.contiguous(), p, mode)

    @staticmethod
    def backward(ctx, g_wav, g_scal):
        g_wav = g_wav.contiguous() if g_wav is not None else None
        g_scal = g_scal.contiguous() if g_scal is not None else None
        if g_wav is None:
            g_wav = torch.zeros(n_wav(ctx.p), dtype=torch.complex128, device=g_scal.device)
        if g_scal is None:
            g_scal = torch.zeros(n_scal(ctx.p), dtype=torch.complex128, device=g_wav.device)
        return analysis_adjoint(g_wav, g_scal, ctx.p, ctx.mode), None, None


def wavelet_analysis(flm, p: Parameters, mode=FP64):
    return WaveletAnalysis.apply(flm, p, mode)
