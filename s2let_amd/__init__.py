"""MI355X-native S2LET: exact directional/spin scale-discretised wavelet
transform on MW-sampled spheres, computed by hand-written HIP kernels.

The compute lives in ``libs2let_amd.so`` (built by ``s2let_amd.build``), which
exports the reference's C API (``s2let_analysis_lm2wav`` ...; see
include/s2let_hip.h).  This module is the Python host side:

* :func:`lib` loads the library (and fails loudly when it is missing);
* the ``capi`` functions call the reference-named entry points on numpy
  arrays (host pointers, FP64, synchronous), mirroring the C signatures;
* :mod:`s2let_amd.torch_ops` exposes the device-pointer entry points on
  torch tensors, with autograd through the hand-written adjoints.

Layouts are the reference's: flm[l*l+l+m]; MW maps theta-major L x (2L-1);
f_wav concatenates per-scale SO(3) blocks (gamma slowest).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# S2LET_AMD_LIB selects another build of the library (e.g. to compare two builds in one process tree)
LIB_PATH = os.environ.get("S2LET_AMD_LIB", os.path.join(_HERE, "libs2let_amd.so"))
_lib = None

FP64 = 0
MIXED = 1


class Parameters(ctypes.Structure):
    """s2let_parameters_t (s2let_types.h:56-137): field order is the ABI."""

    _fields_ = [
        ("verbosity", ctypes.c_int),
        ("reality", ctypes.c_int),
        ("upsample", ctypes.c_int),
        ("B", ctypes.c_double),
        ("L", ctypes.c_int),
        ("J_min", ctypes.c_int),
        ("N", ctypes.c_int),
        ("spin", ctypes.c_int),
        ("original_spin", ctypes.c_int),
        ("sampling_scheme", ctypes.c_int),
        ("dl_method", ctypes.c_int),
    ]

    def key(self):
        return (self.L, self.N, self.B, self.J_min, self.spin, self.upsample, self.original_spin)


def parameters(L, N=1, B=2.0, J_min=0, spin=0, upsample=0, original_spin=0, reality=0, sampling_scheme=0,
               dl_method=0, verbosity=0):
    return Parameters(int(verbosity), int(reality), int(upsample), float(B), int(L), int(J_min), int(N), int(spin),
                      int(original_spin), int(sampling_scheme), int(dl_method))


_SIGS = {
    "s2let_j_max": (ctypes.c_int, [ctypes.c_void_p]),
    "s2let_bandlimit": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "s2let_L0": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "s2let_n_wav": (ctypes.c_int, [ctypes.c_void_p]),
    "s2let_n_scal": (ctypes.c_int, [ctypes.c_void_p]),
    "s2let_n_lmn_wav": (ctypes.c_int, [ctypes.c_void_p]),
    "s2let_n_lm_scal": (ctypes.c_int, [ctypes.c_void_p]),
    "s2let_n_wav_j": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "s2let_n_gamma": (ctypes.c_int, [ctypes.c_void_p]),
    "s2let_n_px": (ctypes.c_int, [ctypes.c_void_p]),
    "s2let_hip_last_error": (ctypes.c_int, []),
    "s2let_hip_last_error_string": (ctypes.c_char_p, []),
    "s2let_hip_version": (ctypes.c_char_p, []),
    "s2let_hip_prepare": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "s2let_hip_release": (None, []),
    "s2let_switch_wavtype": (None, [ctypes.c_int]),
    "s2let_tiling_axisym": (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "s2let_tiling_direction": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "s2let_tiling_wavelet": (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}
for _n in ("analysis_lm2wav", "synthesis_wav2lm", "analysis_px2wav", "synthesis_wav2px", "analysis_adjoint_wav2lm",
           "analysis_adjoint_wav2px", "synthesis_adjoint_lm2wav", "synthesis_adjoint_px2wav"):
    _SIGS["s2let_" + _n] = (None, [ctypes.c_void_p] * 4)
for _n in ("analysis_lm2lmn", "synthesis_lmn2lm", "analysis_adjoint_lmn2lm", "synthesis_adjoint_lm2lmn"):
    _SIGS["s2let_" + _n] = (None, [ctypes.c_void_p] * 6)
for _n in ("analysis_lm2lmn_real", "synthesis_lmn2lm_real", "analysis_adjoint_lmn2lm_real",
           "synthesis_adjoint_lm2lmn_real"):
    _SIGS["s2let_" + _n] = (None, [ctypes.c_void_p] * 6)
for _n in ("mw_alm2map", "mw_map2alm"):
    _SIGS["s2let_" + _n] = (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
for _n in ("analysis_lm2wav", "synthesis_wav2lm", "analysis_adjoint_wav2lm", "synthesis_adjoint_lm2wav"):
    _SIGS["s2let_hip_" + _n] = (ctypes.c_int, [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_void_p])
for _n in ("mw_alm2map", "mw_map2alm", "mw_alm2map_adjoint", "mw_map2alm_adjoint"):
    _SIGS["s2let_hip_" + _n] = (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_void_p])
for _n in ("analysis_lm2wav_real", "synthesis_wav2lm_real", "analysis_px2wav_real", "synthesis_wav2px_real",
           "analysis_adjoint_wav2lm_real", "analysis_adjoint_wav2px_real", "synthesis_adjoint_lm2wav_real",
           "synthesis_adjoint_px2wav_real"):
    _SIGS["s2let_" + _n] = (None, [ctypes.c_void_p] * 4)
for _n in ("mw_alm2map_real", "mw_map2alm_real"):
    _SIGS["s2let_" + _n] = (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int])
for _n in ("analysis_lm2wav_real", "synthesis_wav2lm_real", "analysis_adjoint_wav2lm_real",
           "synthesis_adjoint_lm2wav_real"):
    _SIGS["s2let_hip_" + _n] = (ctypes.c_int, [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_void_p])
for _n in ("analysis_lm2wav_manual", "synthesis_wav2lm_manual", "analysis_adjoint_wav2lm_manual",
           "synthesis_adjoint_lm2wav_manual"):
    _SIGS["s2let_" + _n] = (None, [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_void_p] + [ctypes.c_int] * 4)
_SIGS["s2let_allocate_f_wav_manual"] = (None, [ctypes.c_void_p] * 3 + [ctypes.c_int] * 3 + [ctypes.c_void_p])
for _n in ("lm_wav_analysis", "lm_wav_synthesis", "lm_wav_analysis_multires", "lm_wav_synthesis_multires"):
    _SIGS["s2let_transform_axisym_" + _n] = (None, [ctypes.c_void_p] * 6)
for _n in ("lm_allocate_f_wav", "lm_allocate_f_wav_multires", "lm_allocate_wav", "lm_wav", "allocate_mw_f_wav",
           "allocate_mw_f_wav_multires", "allocate_mw_f_wav_real", "allocate_mw_f_wav_multires_real",
           "wav_hardthreshold_real", "wav_hardthreshold_multires_real"):
    _SIGS["s2let_transform_axisym_" + _n] = (None, [ctypes.c_void_p] * 3)
for _k in ("analysis", "analysis_adjoint", "synthesis", "synthesis_adjoint"):
    for _s in ("", "_multires", "_real", "_multires_real"):
        _SIGS["s2let_transform_axisym_wav_%s_mw%s" % (_k, _s)] = (None, [ctypes.c_void_p] * 4)
_SIGS["s2let_fits_mw_read_bandlimit"] = (ctypes.c_int, [ctypes.c_char_p])
_SIGS["s2let_fits_hpx_read_nside"] = (ctypes.c_int, [ctypes.c_char_p])
for _n in ("mw", "mwss"):
    _SIGS["s2let_fits_%s_write_map" % _n] = (None, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int])
    _SIGS["s2let_fits_%s_read_map" % _n] = (None, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int])
    _SIGS["s2let_fits_%s_write_spin_maps" % _n] = (None, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_int])
    _SIGS["s2let_fits_%s_read_spin_maps" % _n] = (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                                       ctypes.c_int])
_SIGS["s2let_hip_selftest"] = (ctypes.c_int, [])
_SIGS["s2let_hip_timing"] = (None, [ctypes.c_int])
_SIGS["s2let_hip_timing_report"] = (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int])
_SIGS["s2let_hip_batch"] = (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5 +
                            [ctypes.c_int, ctypes.c_void_p])
_SIGS["s2let_hip_batch_graph"] = _SIGS["s2let_hip_batch"]
_SIGS["s2let_hip_autotune"] = (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p])
_SIGS["s2let_hip_analysis_lm2wav_noalloc"] = _SIGS["s2let_hip_analysis_lm2wav"]
_SIGS["s2let_hip_plan_info"] = (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int])
OP_ANALYSIS_LM2WAV, OP_SYNTHESIS_WAV2LM, OP_ANALYSIS_ADJOINT_WAV2LM, OP_SYNTHESIS_ADJOINT_LM2WAV = 0, 1, 2, 3


def exported_symbols():
    """Names declared in include/s2let_hip.h that the library must export."""
    return sorted(_SIGS)


def lib():
    """Loads libs2let_amd.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("s2let_amd: %s not built; run `python -m s2let_amd.build`" % LIB_PATH)
        # One HIP runtime per process: when PyTorch-ROCm is installed, load it first so
        # that the library binds to the libamdhip64 that torch uses (same SONAME).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    L = lib()
    return L.s2let_hip_last_error(), L.s2let_hip_last_error_string().decode()


def check(rc):
    if rc != 0:
        code, msg = last_error()
        raise RuntimeError("s2let_amd error %d: %s" % (code, msg))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _cin(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _ref(p):
    return ctypes.byref(p)


def _rin(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------------------- sizes
def j_max(p):
    return lib().s2let_j_max(_ref(p))


def bandlimit(j, p):
    return lib().s2let_bandlimit(int(j), _ref(p))


def L0(j, p):
    return lib().s2let_L0(int(j), _ref(p))


def n_wav(p):
    return lib().s2let_n_wav(_ref(p))


def n_scal(p):
    return lib().s2let_n_scal(_ref(p))


def n_lmn_wav(p):
    return lib().s2let_n_lmn_wav(_ref(p))


def n_px(p):
    """Samples of one map: L (2L - 1) on the MW grid, (L + 1) 2L on the MWSS grid."""
    return lib().s2let_n_px(_ref(p))


def n_lm_scal(p):
    return lib().s2let_n_lm_scal(_ref(p))


def switch_wavtype(t):
    lib().s2let_switch_wavtype(int(t))


def tiling_wavelet(p):
    J = j_max(p)
    psi = np.zeros((J + 1) * p.L * p.L, np.complex128)
    phi = np.zeros(p.L)
    lib().s2let_tiling_wavelet(_ptr(psi), _ptr(phi), _ref(p))
    return psi, phi


def n_wav_j(j, p):
    return lib().s2let_n_wav_j(int(j), _ref(p))


def tiling_direction(p):
    """s_lm of the directional tiling, L^2 entries (s2let_tiling.c:281-319)."""
    s = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_tiling_direction(_ptr(s), _ref(p))
    return s


def tiling_axisym(p):
    J = j_max(p)
    kappa = np.zeros((J + 1) * p.L)
    kappa0 = np.zeros(p.L)
    lib().s2let_tiling_axisym(_ptr(kappa), _ptr(kappa0), _ref(p))
    return kappa.reshape(J + 1, p.L), kappa0


# ----------------------------------------------------------------------------- host API (reference names)
def analysis_lm2wav(flm, p):
    flm = _cin(flm)
    fw, fs = np.zeros(n_wav(p), np.complex128), np.zeros(n_scal(p), np.complex128)
    lib().s2let_analysis_lm2wav(_ptr(fw), _ptr(fs), _ptr(flm), _ref(p))
    return fw, fs


def synthesis_wav2lm(f_wav, f_scal, p):
    a, b = _cin(f_wav), _cin(f_scal)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_synthesis_wav2lm(_ptr(flm), _ptr(a), _ptr(b), _ref(p))
    return flm


def analysis_px2wav(f, p):
    f = _cin(f)
    fw, fs = np.zeros(n_wav(p), np.complex128), np.zeros(n_scal(p), np.complex128)
    lib().s2let_analysis_px2wav(_ptr(fw), _ptr(fs), _ptr(f), _ref(p))
    return fw, fs


def synthesis_wav2px(f_wav, f_scal, p):
    a, b = _cin(f_wav), _cin(f_scal)
    f = np.zeros(n_px(p), np.complex128)
    lib().s2let_synthesis_wav2px(_ptr(f), _ptr(a), _ptr(b), _ref(p))
    return f


def analysis_adjoint_wav2lm(f_wav, f_scal, p):
    a, b = _cin(f_wav), _cin(f_scal)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_analysis_adjoint_wav2lm(_ptr(flm), _ptr(a), _ptr(b), _ref(p))
    return flm


def analysis_adjoint_wav2px(f_wav, f_scal, p):
    a, b = _cin(f_wav), _cin(f_scal)
    f = np.zeros(n_px(p), np.complex128)
    lib().s2let_analysis_adjoint_wav2px(_ptr(f), _ptr(a), _ptr(b), _ref(p))
    return f


def synthesis_adjoint_lm2wav(flm, p):
    flm = _cin(flm)
    fw, fs = np.zeros(n_wav(p), np.complex128), np.zeros(n_scal(p), np.complex128)
    lib().s2let_synthesis_adjoint_lm2wav(_ptr(fw), _ptr(fs), _ptr(flm), _ref(p))
    return fw, fs


def synthesis_adjoint_px2wav(f, p):
    f = _cin(f)
    fw, fs = np.zeros(n_wav(p), np.complex128), np.zeros(n_scal(p), np.complex128)
    lib().s2let_synthesis_adjoint_px2wav(_ptr(fw), _ptr(fs), _ptr(f), _ref(p))
    return fw, fs


def mw_alm2map(flm, L, spin=0):
    flm = _cin(flm)
    f = np.zeros(L * (2 * L - 1), np.complex128)
    lib().s2let_mw_alm2map(_ptr(f), _ptr(flm), int(L), int(spin))
    return f


def mw_map2alm(f, L, spin=0):
    f = _cin(f)
    flm = np.zeros(L * L, np.complex128)
    lib().s2let_mw_map2alm(_ptr(flm), _ptr(f), int(L), int(spin))
    return flm


# ----------------------------------------------------------------------------- real signals
# Real maps / wavelet coefficients (float64) and the harmonic coefficients of a
# real signal (s2let_analysis.c:371-523, s2let_synthesis.c:386-540,
# s2let_analysis_adjoint.c:402-556, s2let_synthesis_adjoint.c:380-532, s2let_mw.c:51-70).
def _real_wav(p):
    return np.zeros(n_wav(p), np.float64), np.zeros(n_scal(p), np.float64)


def analysis_lm2wav_real(flm, p):
    flm = _cin(flm)
    fw, fs = _real_wav(p)
    lib().s2let_analysis_lm2wav_real(_ptr(fw), _ptr(fs), _ptr(flm), _ref(p))
    return fw, fs


def synthesis_wav2lm_real(f_wav, f_scal, p):
    a, b = _rin(f_wav), _rin(f_scal)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_synthesis_wav2lm_real(_ptr(flm), _ptr(a), _ptr(b), _ref(p))
    return flm


def analysis_px2wav_real(f, p):
    f = _rin(f)
    fw, fs = _real_wav(p)
    lib().s2let_analysis_px2wav_real(_ptr(fw), _ptr(fs), _ptr(f), _ref(p))
    return fw, fs


def synthesis_wav2px_real(f_wav, f_scal, p):
    a, b = _rin(f_wav), _rin(f_scal)
    f = np.zeros(n_px(p), np.float64)
    lib().s2let_synthesis_wav2px_real(_ptr(f), _ptr(a), _ptr(b), _ref(p))
    return f


def analysis_adjoint_wav2lm_real(f_wav, f_scal, p):
    a, b = _rin(f_wav), _rin(f_scal)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_analysis_adjoint_wav2lm_real(_ptr(flm), _ptr(a), _ptr(b), _ref(p))
    return flm


def analysis_adjoint_wav2px_real(f_wav, f_scal, p):
    a, b = _rin(f_wav), _rin(f_scal)
    f = np.zeros(n_px(p), np.float64)
    lib().s2let_analysis_adjoint_wav2px_real(_ptr(f), _ptr(a), _ptr(b), _ref(p))
    return f


def synthesis_adjoint_lm2wav_real(flm, p):
    flm = _cin(flm)
    fw, fs = _real_wav(p)
    lib().s2let_synthesis_adjoint_lm2wav_real(_ptr(fw), _ptr(fs), _ptr(flm), _ref(p))
    return fw, fs


def synthesis_adjoint_px2wav_real(f, p):
    f = _rin(f)
    fw, fs = _real_wav(p)
    lib().s2let_synthesis_adjoint_px2wav_real(_ptr(fw), _ptr(fs), _ptr(f), _ref(p))
    return fw, fs


def mw_alm2map_real(flm, L):
    flm = _cin(flm)
    f = np.zeros(L * (2 * L - 1), np.float64)
    lib().s2let_mw_alm2map_real(_ptr(f), _ptr(flm), int(L))
    return f


def mw_map2alm_real(f, L):
    f = _rin(f)
    flm = np.zeros(L * L, np.complex128)
    lib().s2let_mw_map2alm_real(_ptr(flm), _ptr(f), int(L))
    return flm


def _reality(p, r):
    """Copy of p with its reality flag set to r (sizes the Wigner-coefficient arrays)."""
    q = Parameters.from_buffer_copy(p)
    q.reality = int(r)
    return q


def analysis_lm2lmn(flm, p):
    psi, phi = tiling_wavelet(p)
    flm = _cin(flm)
    wl, sl = np.zeros(n_lmn_wav(_reality(p, 0)), np.complex128), np.zeros(n_lm_scal(p), np.complex128)
    lib().s2let_analysis_lm2lmn(_ptr(wl), _ptr(sl), _ptr(flm), _ptr(psi), _ptr(phi), _ref(p))
    return wl, sl


def synthesis_lmn2lm(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    a, b = _cin(f_wav_lmn), _cin(f_scal_lm)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_synthesis_lmn2lm(_ptr(flm), _ptr(a), _ptr(b), _ptr(psi), _ptr(phi), _ref(p))
    return flm


def analysis_adjoint_lmn2lm(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    a, b = _cin(f_wav_lmn), _cin(f_scal_lm)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_analysis_adjoint_lmn2lm(_ptr(flm), _ptr(a), _ptr(b), _ptr(psi), _ptr(phi), _ref(p))
    return flm


def synthesis_adjoint_lm2lmn(flm, p):
    psi, phi = tiling_wavelet(p)
    flm = _cin(flm)
    wl, sl = np.zeros(n_lmn_wav(_reality(p, 0)), np.complex128), np.zeros(n_lm_scal(p), np.complex128)
    lib().s2let_synthesis_adjoint_lm2lmn(_ptr(wl), _ptr(sl), _ptr(flm), _ptr(psi), _ptr(phi), _ref(p))
    return wl, sl


def analysis_lm2lmn_real(flm, p):
    """Real-signal Wigner coefficients, n >= 0 only (s2let_analysis.c:102-160)."""
    psi, phi = tiling_wavelet(p)
    flm = _cin(flm)
    wl, sl = np.zeros(n_lmn_wav(_reality(p, 1)), np.complex128), np.zeros(n_lm_scal(p), np.complex128)
    lib().s2let_analysis_lm2lmn_real(_ptr(wl), _ptr(sl), _ptr(flm), _ptr(psi), _ptr(phi), _ref(p))
    return wl, sl


def synthesis_lmn2lm_real(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    a, b = _cin(f_wav_lmn), _cin(f_scal_lm)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_synthesis_lmn2lm_real(_ptr(flm), _ptr(a), _ptr(b), _ptr(psi), _ptr(phi), _ref(p))
    return flm


def analysis_adjoint_lmn2lm_real(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    a, b = _cin(f_wav_lmn), _cin(f_scal_lm)
    flm = np.zeros(p.L * p.L, np.complex128)
    lib().s2let_analysis_adjoint_lmn2lm_real(_ptr(flm), _ptr(a), _ptr(b), _ptr(psi), _ptr(phi), _ref(p))
    return flm


def synthesis_adjoint_lm2lmn_real(flm, p):
    psi, phi = tiling_wavelet(p)
    flm = _cin(flm)
    wl, sl = np.zeros(n_lmn_wav(_reality(p, 1)), np.complex128), np.zeros(n_lm_scal(p), np.complex128)
    lib().s2let_synthesis_adjoint_lm2lmn_real(_ptr(wl), _ptr(sl), _ptr(flm), _ptr(psi), _ptr(phi), _ref(p))
    return wl, sl


# ----------------------------------------------------------------------------- manual tiling
class ManualTiling:
    """User wavelets wav_lm ((J+1) L^2), scaling weights scal_l (L) and band-limits for
    the *_manual entry points (s2let_analysis.c:188-276 and siblings)."""

    def __init__(self, wav_lm, scal_l, scal_bandlimit, wav_bandlimits, L, spin, N):
        self.wav_lm = np.ascontiguousarray(wav_lm, dtype=np.complex128)
        self.scal_l = np.ascontiguousarray(scal_l, dtype=np.float64)
        self.wav_bl = np.ascontiguousarray(wav_bandlimits, dtype=np.int32)
        self.scal_bl, self.L, self.spin, self.N = int(scal_bandlimit), int(L), int(spin), int(N)
        self.J = self.wav_bl.size - 1
        assert self.wav_lm.size == (self.J + 1) * self.L * self.L and self.scal_l.size >= self.L

    def _args(self):
        return (_ptr(self.scal_l), _ptr(self.wav_lm), self.scal_bl, _ptr(self.wav_bl), self.J, self.L, self.spin,
                self.N)

    def n_wav(self):
        tot = 0
        for b in self.wav_bl:
            Lj = min(int(b), self.L)
            Nj = min(self.N, Lj)
            Nj += (Nj + self.N) % 2
            tot += Lj * (2 * Lj - 1) * (2 * Nj - 1)
        return tot

    def n_scal(self):
        b = min(self.scal_bl, self.L)
        return b * (2 * b - 1)


def analysis_lm2wav_manual(flm, t):
    flm = _cin(flm)
    fw, fs = np.zeros(t.n_wav(), np.complex128), np.zeros(t.n_scal(), np.complex128)
    lib().s2let_analysis_lm2wav_manual(_ptr(fw), _ptr(fs), _ptr(flm), *t._args())
    return fw, fs


def synthesis_wav2lm_manual(f_wav, f_scal, t, flm=None):
    """Adds into flm (a copy of it is returned; zeros when None), as the reference does."""
    a, b = _cin(f_wav), _cin(f_scal)
    out = np.zeros(t.L * t.L, np.complex128) if flm is None else np.array(flm, dtype=np.complex128)
    lib().s2let_synthesis_wav2lm_manual(_ptr(out), _ptr(a), _ptr(b), *t._args())
    return out


def analysis_adjoint_wav2lm_manual(f_wav, f_scal, t, flm=None):
    a, b = _cin(f_wav), _cin(f_scal)
    out = np.zeros(t.L * t.L, np.complex128) if flm is None else np.array(flm, dtype=np.complex128)
    lib().s2let_analysis_adjoint_wav2lm_manual(_ptr(out), _ptr(a), _ptr(b), *t._args())
    return out


def synthesis_adjoint_lm2wav_manual(flm, t):
    flm = _cin(flm)
    fw, fs = np.zeros(t.n_wav(), np.complex128), np.zeros(t.n_scal(), np.complex128)
    lib().s2let_synthesis_adjoint_lm2wav_manual(_ptr(fw), _ptr(fs), _ptr(flm), *t._args())
    return fw, fs


# ----------------------------------------------------------------------------- axisymmetric wavelets
# s2let_transform_axisym_lm.c, s2let_transform_axisym_mw.c (spin 0, MW sampling)
def _axisym_sizes(p, multires):
    bl = [min(bandlimit(j, p), p.L) if multires else p.L for j in range(p.J_min, j_max(p) + 1)]
    bs = min(bandlimit(p.J_min - 1, p), p.L) if multires else p.L
    return bl, bs


def axisym_lm_wav(p):
    """Kernels (wav_lm ((J+1) L), scal_lm (L)) of s2let_transform_axisym_lm_wav."""
    wav, scal = np.zeros((j_max(p) + 1) * p.L), np.zeros(p.L)
    lib().s2let_transform_axisym_lm_wav(_ptr(wav), _ptr(scal), _ref(p))
    return wav, scal


def axisym_lm_wav_analysis(flm, wav_lm, scal_lm, p, multires=False):
    bl, bs = _axisym_sizes(p, multires)
    fw, fs = np.zeros(sum(b * b for b in bl), np.complex128), np.zeros(bs * bs, np.complex128)
    fn = "s2let_transform_axisym_lm_wav_analysis" + ("_multires" if multires else "")
    getattr(lib(), fn)(_ptr(fw), _ptr(fs), _ptr(_cin(flm)), _ptr(_rin(wav_lm)), _ptr(_rin(scal_lm)), _ref(p))
    return fw, fs


def axisym_lm_wav_synthesis(f_wav_lm, f_scal_lm, wav_lm, scal_lm, p, multires=False, flm=None):
    """Adds into flm (a copy is returned; zeros when None), as the reference does."""
    out = np.zeros(p.L * p.L, np.complex128) if flm is None else np.array(flm, dtype=np.complex128)
    fn = "s2let_transform_axisym_lm_wav_synthesis" + ("_multires" if multires else "")
    getattr(lib(), fn)(_ptr(out), _ptr(_cin(f_wav_lm)), _ptr(_cin(f_scal_lm)), _ptr(_rin(wav_lm)),
                       _ptr(_rin(scal_lm)), _ref(p))
    return out


def _axisym_fn(kind, multires, real):
    return "s2let_transform_axisym_wav_%s_mw%s%s" % (kind, "_multires" if multires else "", "_real" if real else "")


def axisym_wav_analysis_mw(f, p, multires=False, real=False):
    bl, bs = _axisym_sizes(p, multires)
    dt, cv = (np.float64, _rin) if real else (np.complex128, _cin)
    fw, fs = np.zeros(sum(b * (2 * b - 1) for b in bl), dt), np.zeros(bs * (2 * bs - 1), dt)
    getattr(lib(), _axisym_fn("analysis", multires, real))(_ptr(fw), _ptr(fs), _ptr(cv(f)), _ref(p))
    return fw, fs


def axisym_wav_synthesis_adjoint_mw(f, p, multires=False, real=False):
    bl, bs = _axisym_sizes(p, multires)
    dt, cv = (np.float64, _rin) if real else (np.complex128, _cin)
    fw, fs = np.zeros(sum(b * (2 * b - 1) for b in bl), dt), np.zeros(bs * (2 * bs - 1), dt)
    getattr(lib(), _axisym_fn("synthesis_adjoint", multires, real))(_ptr(fw), _ptr(fs), _ptr(cv(f)), _ref(p))
    return fw, fs


def axisym_wav_synthesis_mw(f_wav, f_scal, p, multires=False, real=False):
    dt, cv = (np.float64, _rin) if real else (np.complex128, _cin)
    f = np.zeros(p.L * (2 * p.L - 1), dt)
    getattr(lib(), _axisym_fn("synthesis", multires, real))(_ptr(f), _ptr(cv(f_wav)), _ptr(cv(f_scal)), _ref(p))
    return f


def axisym_wav_analysis_adjoint_mw(f_wav, f_scal, p, multires=False, real=False):
    dt, cv = (np.float64, _rin) if real else (np.complex128, _cin)
    f = np.zeros(p.L * (2 * p.L - 1), dt)
    getattr(lib(), _axisym_fn("analysis_adjoint", multires, real))(_ptr(f), _ptr(cv(f_wav)), _ptr(cv(f_scal)),
                                                                   _ref(p))
    return f


def axisym_wav_hardthreshold_real(g_wav, threshold, p, multires=False):
    """In place on a copy: samples with |trunc(x)| < threshold[j - J_min] become 0."""
    g = np.array(g_wav, dtype=np.float64)
    fn = "s2let_transform_axisym_wav_hardthreshold%s_real" % ("_multires" if multires else "")
    getattr(lib(), fn)(_ptr(g), _ptr(_rin(threshold)), _ref(p))
    return g


# ----------------------------------------------------------------------------- FITS maps (s2let_fits.c)
def _npx(L, mwss):
    return (L + 1) * 2 * L if mwss else L * (2 * L - 1)


def fits_write_map(filename, f, L, mwss=False):
    getattr(lib(), "s2let_fits_%s_write_map" % ("mwss" if mwss else "mw"))(str(filename).encode(), _ptr(_rin(f)), int(L))


def fits_read_map(filename, L, mwss=False):
    f = np.zeros(_npx(L, mwss))
    getattr(lib(), "s2let_fits_%s_read_map" % ("mwss" if mwss else "mw"))(_ptr(f), str(filename).encode(), int(L))
    return f


def fits_write_spin_maps(filename, fQ, fU, L, mwss=False):
    getattr(lib(), "s2let_fits_%s_write_spin_maps" % ("mwss" if mwss else "mw"))(str(filename).encode(), _ptr(_rin(fQ)),
                                                                       _ptr(_rin(fU)), int(L))


def fits_read_spin_maps(filename, L, mwss=False):
    q, u = np.zeros(_npx(L, mwss)), np.zeros(_npx(L, mwss))
    getattr(lib(), "s2let_fits_%s_read_spin_maps" % ("mwss" if mwss else "mw"))(_ptr(q), _ptr(u), str(filename).encode(),
                                                                      int(L))
    return q, u


def fits_read_bandlimit(filename):
    return lib().s2let_fits_mw_read_bandlimit(str(filename).encode())


def fits_read_nside(filename):
    return lib().s2let_fits_hpx_read_nside(str(filename).encode())


def timing(enable=True):
    """Starts (clears) or stops per-stage HIP-event timing inside the library."""
    lib().s2let_hip_timing(1 if enable else 0)


def timing_report():
    """{stage: (launches, total_ms)} of the calls made since timing(True)."""
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    n = lib().s2let_hip_timing_report(buf, len(buf))
    if n < 0:
        check(n)
    return {k: (v[0], v[1]) for k, v in json.loads(buf.value.decode()).items()}


def plan_info(p, mode=FP64):
    """The plan's layout and launch choices (s2let_hip_plan_info) as a dict."""
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    n = lib().s2let_hip_plan_info(_ref(p), int(mode), buf, len(buf))
    if n < 0:
        check(n)
    return json.loads(buf.value.decode())


def prepare(p, mode=FP64):
    check(lib().s2let_hip_prepare(_ref(p), int(mode)))


def release():
    lib().s2let_hip_release()
