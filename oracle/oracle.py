"""ctypes front end of the CPU FP64 oracle (oracle/s2let_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Every function
takes and returns numpy arrays (complex128 / float64) in the reference's own
layouts (flm[l*l+l+m], MW maps theta-major, SO(3) samples gamma slowest).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libs2let_oracle.so")
_lib = None


class Params(ctypes.Structure):
    """Same layout as s2let_parameters_t (s2let_types.h:56-137)."""

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


def params(L, N, B=2.0, J_min=0, spin=0, upsample=0, original_spin=0, sampling=0):
    """sampling: 0 MW, 1 MWSS (S2LET_SAMPLING_MW_SS)."""
    return Params(0, 0, int(upsample), float(B), int(L), int(J_min), int(N), int(spin),
                  int(original_spin), int(sampling), 0)


def map_size(L, ss=0):
    return (L + 1) * 2 * L if ss else L * (2 * L - 1)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(n):
    return np.zeros(int(n), dtype=np.complex128)


def _in(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def set_threads(n):
    lib().oracle_set_num_threads(int(n))


def num_threads():
    return lib().oracle_num_threads()


def switch_wavtype(t):
    lib().oracle_switch_wavtype(int(t))


def j_max(p):
    return lib().oracle_j_max(ctypes.byref(p))


def bandlimit(j, p):
    return lib().oracle_bandlimit(int(j), ctypes.byref(p))


def L0(j, p):
    return lib().oracle_L0(int(j), ctypes.byref(p))


def binomial(n, k, exact=1):
    f = lib().oracle_binomial_coefficient
    f.restype = ctypes.c_ulong
    return f(int(n), int(k), int(exact))


def n_wav(p):
    return lib().oracle_n_wav(ctypes.byref(p))


def n_scal(p):
    return lib().oracle_n_scal(ctypes.byref(p))


def n_lmn_wav(p):
    return lib().oracle_n_lmn_wav(ctypes.byref(p))


def n_lm_scal(p):
    return lib().oracle_n_lm_scal(ctypes.byref(p))


def flmn_size(L, N):
    return lib().oracle_flmn_size(int(L), int(N))


def elmn2ind(el, m, n, L, N):
    return lib().oracle_elmn2ind(int(el), int(m), int(n), int(L), int(N))


def tiling_axisym(p):
    J = j_max(p)
    kappa = np.zeros((J + 1) * p.L)
    kappa0 = np.zeros(p.L)
    lib().oracle_tiling_axisym(_p(kappa), _p(kappa0), ctypes.byref(p))
    return kappa.reshape(J + 1, p.L), kappa0


def tiling_direction(p):
    s = _c(p.L * p.L)
    lib().oracle_tiling_direction(_p(s), ctypes.byref(p))
    return s


def tiling_wavelet(p):
    J = j_max(p)
    psi = _c((J + 1) * p.L * p.L)
    phi = np.zeros(p.L)
    lib().oracle_tiling_wavelet(_p(psi), _p(phi), ctypes.byref(p))
    return psi, phi


def delta_plane(l):
    out = np.zeros((2 * l + 1) * (2 * l + 1))
    lib().oracle_delta_plane(_p(out), int(l))
    return out.reshape(2 * l + 1, 2 * l + 1)


def dl_beta_full(L, beta):
    """list of d^l(beta) planes, plane[l][m+l, m'+l] = d^l_{m m'}(beta)."""
    tot = sum((2 * l + 1) ** 2 for l in range(L))
    out = np.zeros(tot)
    lib().oracle_dl_beta_full(_p(out), int(L), ctypes.c_double(beta))
    planes, off = [], 0
    for l in range(L):
        w = 2 * l + 1
        planes.append(out[off:off + w * w].reshape(w, w))
        off += w * w
    return planes


def ssht_inverse(flm, L, spin=0, ss=0):
    f = _c(map_size(L, ss))
    flm = _in(flm)
    lib().oracle_ssht_inverse_s(_p(f), _p(flm), int(L), int(spin), int(ss))
    return f


def ssht_forward(f, L, spin=0, ss=0):
    flm = _c(L * L)
    f = _in(f)
    lib().oracle_ssht_forward_s(_p(flm), _p(f), int(L), int(spin), int(ss))
    return flm


def ssht_adjoint_inverse(f, L, spin=0, ss=0):
    flm = _c(L * L)
    f = _in(f)
    lib().oracle_ssht_adjoint_inverse_s(_p(flm), _p(f), int(L), int(spin), int(ss))
    return flm


def ssht_adjoint_forward(flm, L, spin=0, ss=0):
    f = _c(map_size(L, ss))
    flm = _in(flm)
    lib().oracle_ssht_adjoint_forward_s(_p(f), _p(flm), int(L), int(spin), int(ss))
    return f


def so3_inverse(flmn, L, N, L0=0, nmode=0, ss=0):
    f = _c(map_size(L, ss) * (2 * N - 1))
    flmn = _in(flmn)
    lib().oracle_so3_inverse_s(_p(f), _p(flmn), int(L), int(N), int(L0), int(nmode), int(ss))
    return f


def so3_forward(f, L, N, L0=0, nmode=0, ss=0):
    flmn = _c(flmn_size(L, N))
    f = _in(f)
    lib().oracle_so3_forward_s(_p(flmn), _p(f), int(L), int(N), int(L0), int(nmode), int(ss))
    return flmn


def so3_adjoint_inverse(f, L, N, L0=0, nmode=0, ss=0):
    flmn = _c(flmn_size(L, N))
    f = _in(f)
    lib().oracle_so3_adjoint_inverse_s(_p(flmn), _p(f), int(L), int(N), int(L0), int(nmode), int(ss))
    return flmn


def so3_adjoint_forward(flmn, L, N, L0=0, nmode=0, ss=0):
    f = _c(map_size(L, ss) * (2 * N - 1))
    flmn = _in(flmn)
    lib().oracle_so3_adjoint_forward_s(_p(f), _p(flmn), int(L), int(N), int(L0), int(nmode), int(ss))
    return f


def analysis_lm2lmn(flm, p):
    psi, phi = tiling_wavelet(p)
    wl, sl = _c(n_lmn_wav(p)), _c(n_lm_scal(p))
    flm = _in(flm)
    lib().oracle_analysis_lm2lmn(_p(wl), _p(sl), _p(flm), _p(psi), _p(phi), ctypes.byref(p))
    return wl, sl


def synthesis_lmn2lm(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    flm = _c(p.L * p.L)
    a, b = _in(f_wav_lmn), _in(f_scal_lm)
    lib().oracle_synthesis_lmn2lm(_p(flm), _p(a), _p(b), _p(psi), _p(phi), ctypes.byref(p))
    return flm


def analysis_adjoint_lmn2lm(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    flm = _c(p.L * p.L)
    a, b = _in(f_wav_lmn), _in(f_scal_lm)
    lib().oracle_analysis_adjoint_lmn2lm(_p(flm), _p(a), _p(b), _p(psi), _p(phi), ctypes.byref(p))
    return flm


def synthesis_adjoint_lm2lmn(flm, p):
    psi, phi = tiling_wavelet(p)
    wl, sl = _c(n_lmn_wav(p)), _c(n_lm_scal(p))
    flm = _in(flm)
    lib().oracle_synthesis_adjoint_lm2lmn(_p(wl), _p(sl), _p(flm), _p(psi), _p(phi), ctypes.byref(p))
    return wl, sl


def flmn_size_real(L, N):
    return lib().oracle_flmn_size_real(int(L), int(N))


def elmn2ind_real(el, m, n, L):
    return lib().oracle_elmn2ind_real(int(el), int(m), int(n), int(L))


def n_lmn_wav_real(p):
    return lib().oracle_n_lmn_wav_real(ctypes.byref(p))


def analysis_lm2lmn_real(flm, p):
    psi, phi = tiling_wavelet(p)
    wl, sl = _c(n_lmn_wav_real(p)), _c(n_lm_scal(p))
    flm = _in(flm)
    lib().oracle_analysis_lm2lmn_real(_p(wl), _p(sl), _p(flm), _p(psi), _p(phi), ctypes.byref(p))
    return wl, sl


def synthesis_adjoint_lm2lmn_real(flm, p):
    psi, phi = tiling_wavelet(p)
    wl, sl = _c(n_lmn_wav_real(p)), _c(n_lm_scal(p))
    flm = _in(flm)
    lib().oracle_synthesis_adjoint_lm2lmn_real(_p(wl), _p(sl), _p(flm), _p(psi), _p(phi), ctypes.byref(p))
    return wl, sl


def synthesis_lmn2lm_real(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    flm = _c(p.L * p.L)
    a, b = _in(f_wav_lmn), _in(f_scal_lm)
    lib().oracle_synthesis_lmn2lm_real(_p(flm), _p(a), _p(b), _p(psi), _p(phi), ctypes.byref(p))
    return flm


def analysis_adjoint_lmn2lm_real(f_wav_lmn, f_scal_lm, p):
    psi, phi = tiling_wavelet(p)
    flm = _c(p.L * p.L)
    a, b = _in(f_wav_lmn), _in(f_scal_lm)
    lib().oracle_analysis_adjoint_lmn2lm_real(_p(flm), _p(a), _p(b), _p(psi), _p(phi), ctypes.byref(p))
    return flm


class Manual:
    """A manual tiling: wavelets wav_lm ((J+1) L^2), scal_l (L), band-limits."""

    def __init__(self, wav_lm, scal_l, scal_bl, wav_bl, L, spin, N):
        self.wav_lm = np.ascontiguousarray(wav_lm, dtype=np.complex128)
        self.scal_l = np.ascontiguousarray(scal_l, dtype=np.float64)
        self.wav_bl = np.ascontiguousarray(wav_bl, dtype=np.int32)
        self.scal_bl, self.L, self.spin, self.N = int(scal_bl), int(L), int(spin), int(N)
        self.J = len(self.wav_bl) - 1

    def args(self):
        return (_p(self.scal_l), _p(self.wav_lm), self.scal_bl, _p(self.wav_bl), self.J, self.L, self.spin, self.N)

    def n_wav(self):
        return lib().oracle_n_wav_manual(_p(self.wav_bl), self.J, self.L, self.N)

    def n_scal(self):
        b = min(self.scal_bl, self.L)
        return b * (2 * b - 1)


def analysis_lm2wav_manual(flm, m):
    fw, fs = _c(m.n_wav()), _c(m.n_scal())
    flm = _in(flm)
    lib().oracle_analysis_lm2wav_manual(_p(fw), _p(fs), _p(flm), *m.args())
    return fw, fs


def synthesis_adjoint_lm2wav_manual(flm, m):
    fw, fs = _c(m.n_wav()), _c(m.n_scal())
    flm = _in(flm)
    lib().oracle_synthesis_adjoint_lm2wav_manual(_p(fw), _p(fs), _p(flm), *m.args())
    return fw, fs


def synthesis_wav2lm_manual(f_wav, f_scal, m, flm=None):
    out = _c(m.L * m.L) if flm is None else np.array(flm, dtype=np.complex128)
    a, b = _in(f_wav), _in(f_scal)
    lib().oracle_synthesis_wav2lm_manual(_p(out), _p(a), _p(b), *m.args())
    return out


def analysis_adjoint_wav2lm_manual(f_wav, f_scal, m, flm=None):
    out = _c(m.L * m.L) if flm is None else np.array(flm, dtype=np.complex128)
    a, b = _in(f_wav), _in(f_scal)
    lib().oracle_analysis_adjoint_wav2lm_manual(_p(out), _p(a), _p(b), *m.args())
    return out


def _wav_pair(p):
    return _c(n_wav(p)), _c(n_scal(p))


def analysis_lm2wav(flm, p):
    fw, fs = _wav_pair(p)
    flm = _in(flm)
    lib().oracle_analysis_lm2wav(_p(fw), _p(fs), _p(flm), ctypes.byref(p))
    return fw, fs


def synthesis_wav2lm(f_wav, f_scal, p):
    flm = _c(p.L * p.L)
    a, b = _in(f_wav), _in(f_scal)
    lib().oracle_synthesis_wav2lm(_p(flm), _p(a), _p(b), ctypes.byref(p))
    return flm


def analysis_px2wav(f, p):
    fw, fs = _wav_pair(p)
    f = _in(f)
    lib().oracle_analysis_px2wav(_p(fw), _p(fs), _p(f), ctypes.byref(p))
    return fw, fs


def synthesis_wav2px(f_wav, f_scal, p):
    f = _c(map_size(p.L, p.sampling_scheme))
    a, b = _in(f_wav), _in(f_scal)
    lib().oracle_synthesis_wav2px(_p(f), _p(a), _p(b), ctypes.byref(p))
    return f


def analysis_adjoint_wav2lm(f_wav, f_scal, p):
    flm = _c(p.L * p.L)
    a, b = _in(f_wav), _in(f_scal)
    lib().oracle_analysis_adjoint_wav2lm(_p(flm), _p(a), _p(b), ctypes.byref(p))
    return flm


def analysis_adjoint_wav2px(f_wav, f_scal, p):
    f = _c(map_size(p.L, p.sampling_scheme))
    a, b = _in(f_wav), _in(f_scal)
    lib().oracle_analysis_adjoint_wav2px(_p(f), _p(a), _p(b), ctypes.byref(p))
    return f


def synthesis_adjoint_lm2wav(flm, p):
    fw, fs = _wav_pair(p)
    flm = _in(flm)
    lib().oracle_synthesis_adjoint_lm2wav(_p(fw), _p(fs), _p(flm), ctypes.byref(p))
    return fw, fs


def synthesis_adjoint_px2wav(f, p):
    fw, fs = _wav_pair(p)
    f = _in(f)
    lib().oracle_synthesis_adjoint_px2wav(_p(fw), _p(fs), _p(f), ctypes.byref(p))
    return fw, fs


def axisym_lm_wav(p):
    J = j_max(p)
    wav, scal = np.zeros((J + 1) * p.L), np.zeros(p.L)
    lib().oracle_axisym_lm_wav(_p(wav), _p(scal), ctypes.byref(p))
    return wav, scal


def _axisym_sizes(p, multires):
    bl = [min(bandlimit(j, p), p.L) if multires else p.L for j in range(p.J_min, j_max(p) + 1)]
    bs = min(bandlimit(p.J_min - 1, p), p.L) if multires else p.L
    return bl, bs


def axisym_lm_wav_analysis(flm, wav_lm, scal_lm, p, multires=0):
    bl, bs = _axisym_sizes(p, multires)
    fw, fs = _c(sum(b * b for b in bl)), _c(bs * bs)
    lib().oracle_axisym_lm_wav_analysis(_p(fw), _p(fs), _p(_in(flm)), _p(np.ascontiguousarray(wav_lm, float)),
                                        _p(np.ascontiguousarray(scal_lm, float)), ctypes.byref(p), int(multires))
    return fw, fs


def axisym_lm_wav_synthesis(f_wav_lm, f_scal_lm, wav_lm, scal_lm, p, multires=0, flm=None):
    out = _c(p.L * p.L) if flm is None else np.array(flm, dtype=np.complex128)
    lib().oracle_axisym_lm_wav_synthesis(_p(out), _p(_in(f_wav_lm)), _p(_in(f_scal_lm)),
                                         _p(np.ascontiguousarray(wav_lm, float)),
                                         _p(np.ascontiguousarray(scal_lm, float)), ctypes.byref(p), int(multires))
    return out


def axisym_mw(kind, multires, x, y, p):
    """kind 0 analysis (x = f), 1 analysis adjoint (x, y = f_wav, f_scal), 2 synthesis
    (x, y = f_wav, f_scal), 3 synthesis adjoint (x = f)."""
    bl, bs = _axisym_sizes(p, multires)
    nw, ns, nf = sum(b * (2 * b - 1) for b in bl), bs * (2 * bs - 1), p.L * (2 * p.L - 1)
    if kind in (0, 3):
        fw, fs = _c(nw), _c(ns)
        lib().oracle_axisym_mw(int(kind), int(multires), _p(fw), _p(fs), _p(_in(x)), ctypes.byref(p))
        return fw, fs
    f = _c(nf)
    lib().oracle_axisym_mw(int(kind), int(multires), _p(f), _p(_in(x)), _p(_in(y)), ctypes.byref(p))
    return f


def axisym_wav_analysis_mw_multires(f, p):
    J = j_max(p)
    tot = sum(min(bandlimit(j, p), p.L) * (2 * min(bandlimit(j, p), p.L) - 1)
              for j in range(p.J_min, J + 1))
    bl = min(bandlimit(p.J_min - 1, p), p.L)
    fw, fs = _c(tot), _c(bl * (2 * bl - 1))
    f = _in(f)
    lib().oracle_axisym_wav_analysis_mw_multires(_p(fw), _p(fs), _p(f), ctypes.byref(p))
    return fw, fs
