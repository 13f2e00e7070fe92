"""pys2let-compatible Python API (src/pys2let/pys2let.pyx:251-1177) on the MI355X library.

Same function names, argument order, defaults, array shapes and index conventions
as the reference's Cython module, so that `from s2let_amd.pys2let import ...` can
replace `from pys2let import ...`.  Harmonic inputs and outputs of the
*_lm_wav / lm2wav / wav2lm functions use the healpy ordering of the reference
(`lm_hp2lm` / `lm2lm_hp`: m >= 0 only, the m < 0 half rebuilt as
(-1)^m conj(f_{l,|m|}), i.e. real signals); the lmn functions and alm2map_mw /
map2alm_mw take the full L^2 ordering, as there.  Every transform runs on the
GPU through the C ABI (libs2let_amd.so); `ssht_dl_beta_risbo` (a wrapper of an
SSHT routine, not of S2LET) is not provided.
"""
from __future__ import annotations

import numpy as np

from . import (Parameters, analysis_adjoint_lmn2lm as _aa_lmn2lm, analysis_adjoint_wav2px as _aa_wav2px,
               analysis_lm2lmn as _a_lm2lmn, analysis_lm2wav as _a_lm2wav, analysis_px2wav as _a_px2wav,
               axisym_lm_wav as _ax_lm_wav, axisym_lm_wav_analysis as _ax_lm_ana,
               axisym_lm_wav_synthesis as _ax_lm_syn, axisym_wav_analysis_adjoint_mw as _ax_ana_adj,
               axisym_wav_analysis_mw as _ax_ana, axisym_wav_synthesis_adjoint_mw as _ax_syn_adj,
               axisym_wav_synthesis_mw as _ax_syn, ManualTiling, analysis_lm2wav_manual as _a_manual,
               synthesis_wav2lm_manual as _s_manual, lib, mw_alm2map as _alm2map, mw_map2alm as _map2alm,
               n_wav_j as _n_wav_j, parameters as _params, synthesis_adjoint_lm2lmn as _sa_lm2lmn,
               synthesis_adjoint_px2wav as _sa_px2wav, synthesis_lmn2lm as _s_lmn2lm,
               synthesis_wav2lm as _s_wav2lm, synthesis_wav2px as _s_wav2px, tiling_axisym as _tiling_axisym,
               tiling_direction as _tiling_direction, tiling_wavelet as _tiling_wavelet)
from . import bandlimit as _bandlimit, j_max as _j_max


def _p(B, L, J_min, N=0, spin=0, upsample=0, original_spin=0):
    return _params(int(L), int(N), float(B), int(J_min), int(spin), int(upsample), int(original_spin))


# ---------------------------------------------------------------------------- indices and sizes
def pys2let_j_max(B, L, J_min):
    return _j_max(_p(B, L, J_min))


def mw_lm(el, em):
    """Index of (l, m) in the full L^2 ordering, as the reference returns it (l^2 + l m)."""
    return el * el + el * em


def healpy_lm(el, em, L):
    return int(em * (2 * L - 1 - em) / 2 + el)


def mw_size(L):
    return L * (2 * L - 1)


def _hp_index(L):
    """(healpy index, l, m) for 0 <= m <= l < L."""
    el, em = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    keep = em <= el
    el, em = el[keep], em[keep]
    return (em * (2 * L - 1 - em) // 2 + el), el, em


def lm2lm_hp(f_lm, L):
    hp, el, em = _hp_index(L)
    out = np.zeros(L * (L + 1) // 2, dtype=complex)
    out[hp] = np.asarray(f_lm)[el * el + el + em]
    return out


def lm_hp2lm(flm_hp, L):
    hp, el, em = _hp_index(L)
    flm_hp = np.asarray(flm_hp)
    out = np.zeros(L * L, dtype=complex)
    out[el * el + el - em] = (-1.0) ** em * np.conj(flm_hp[hp])
    out[el * el + el + em] = flm_hp[hp]
    return out


def wav_ind(j, n, B, L, N, J_min, upsample):
    """(offset, band-limit, samples per direction, samples of scale j) of scale j,
    direction n in f_wav (pys2let.pyx:962-990)."""
    p = _p(B, L, J_min, N, 0, upsample)
    offset = sum(_n_wav_j(jp, p) for jp in range(J_min, j))
    bl = L if upsample else min(_bandlimit(j, p), L)
    offset += n * mw_size(bl)
    return offset, bl, mw_size(bl), _n_wav_j(j, p)


def mw_sampling(L):
    """MW sample positions: theta_t = pi (2t+1)/(2L-1), t < L; phi_p = 2 pi p/(2L-1)."""
    t = np.arange(L)
    p = np.arange(2 * L - 1)
    return np.pi * (2 * t + 1) / (2 * L - 1), 2 * np.pi * p / (2 * L - 1)


# ---------------------------------------------------------------------------- tilings
def wavelet_tiling(B, L, N, J_min, spin, original_spin=0):
    """(scal_l (L,), wav_l (L^2, J - J_min + 1)) of the directional tiling; original_spin is
    not forwarded, as in pys2let.pyx:1034-1067."""
    p = _p(B, L, J_min, N, spin, 0, 0)
    J = _j_max(p)
    psi, phi = _tiling_wavelet(p)
    psi = np.asarray(psi).reshape(J + 1, L * L)
    wav_l = np.zeros((L * L, J - J_min + 1), dtype=complex)
    s2 = spin * spin
    wav_l[s2:, :] = psi[J_min:J + 1, s2:].T
    return np.asarray(phi).astype(complex), wav_l


def axisym_wav_l(B, L, J_min):
    """(kappa_0 (L,), kappa (L, J - J_min + 1)) of the axisymmetric tiling."""
    p = _p(B, L, J_min)
    kappa, kappa0 = _tiling_axisym(p)
    return np.array(kappa0, dtype=float), np.array(kappa, dtype=float)[J_min:].T.copy()


def verify_tiling(L, scal_l, wav_l, scal_bandlimit, wav_bandlimits):
    """True when the tiling (scal_l (L,), wav_l ((J+1) L,)) is admissible and each function
    vanishes above its band-limit (pys2let.pyx:442-480)."""
    scal_l = np.asarray(scal_l, dtype=float)
    wav_l = np.asarray(wav_l, dtype=float)
    J = len(wav_bandlimits) - 1

    def measured(f):
        c = np.cumsum(f)
        c = c / c[-1]
        return c, int(np.where(c == 1)[0][0]) + 1

    total = scal_l ** 2
    c, m = measured(scal_l)
    if c[int(scal_bandlimit) - 1] < 1:
        print("Wrong band-limit given for scaling function (", scal_bandlimit, ")")
        print("Measured band-limit is ", m)
        return False
    print("Measured / given band-limits for scaling fct are", m, "/", scal_bandlimit)
    for j in range(J + 1):
        w = wav_l[j * L:(j + 1) * L]
        total = total + w ** 2
        c, m = measured(w)
        if c[int(wav_bandlimits[j]) - 1] < 1:
            print("Wrong band-limit given for wavelet", j + 1, "on", J + 1, "(", wav_bandlimits[j], ")")
            print("Measured band-limit is ", m)
            print("Wavelet:", w)
            return False
        print("Measured / given band-limits for wavelet", j, "are", m, "/", wav_bandlimits[j])
    if np.allclose(total, 1.0):
        return True
    print("Admissibility condition not satisfied:")
    print(total)
    return False


def construct_hybrid_tiling(L, J_min0, Bs, L_transitions):
    """A tiling that switches from B = Bs[k] to Bs[k+1] at the band-limit L_transitions[k]
    (pys2let.pyx:1132-1177): returns (scal_l, wav_l (L, J+1), scal_bandlimit,
    wav_bandlimits, J, L_bounds)."""
    Bs = np.asarray(Bs)
    nb = Bs.size
    jmins = np.zeros(nb, dtype=int)
    jmins[0] = J_min0
    jmaxs = [pys2let_j_max(Bs[k], L, jmins[k]) for k in range(nb)]
    bounds = np.zeros(nb + 1, dtype=np.int32)
    left = np.zeros(nb, dtype=np.int32)
    right = np.zeros(nb, dtype=np.int32)
    J = 0
    for k in range(nb):
        left[k] = jmins[k] if k == 0 else int(np.log(bounds[k]) / np.log(Bs[k])) + 1
        right[k] = jmaxs[k] + 1 if k == nb - 1 else int(np.log(L_transitions[k]) / np.log(Bs[k])) + 1
        if k < nb - 1:
            bounds[k + 1] = np.rint(Bs[k] ** (right[k] - 1))
        J += int(right[k] - left[k] - 1)
    bounds[-1] = L
    wav_out = np.zeros((L, J + 1))
    bl_out = np.zeros(J + 1, dtype=np.int32)
    scal_out, scal_bl = None, None
    col = 0
    for k in range(nb):
        kappa0, kappa = axisym_wav_l(Bs[k], L, jmins[k])
        js = np.arange(jmins[k], jmaxs[k] + 1)
        bls = np.array([np.rint(min(x, L)) for x in Bs[k] ** (js + 1)]).astype(np.int32)
        if k == 0:
            scal_bl = Bs[k] ** jmins[k]
            scal_out = kappa0.copy()
        lo, hi = bounds[k], bounds[k + 1]
        for jj in range(left[k] - jmins[k], right[k] - jmins[k]):
            bl_out[col] = bls[jj]
            wav_out[lo:hi, col] = kappa[lo:hi, jj]
            if jj == left[k] - jmins[k] and kappa[lo, jj] < 1.0:
                wav_out[lo:hi, col] = np.sqrt(wav_out[lo:hi, col] ** 2 + kappa[lo:hi, jj - 1] ** 2)
            if k < nb - 1 and jj == right[k] - 1 - jmins[k] and kappa[lo, jj] < 1.0:
                wav_out[lo:hi, col] = np.sqrt(wav_out[lo:hi, col] ** 2 + kappa[lo:hi, jj + 1] ** 2)
            if jj < right[k] - 1 - jmins[k]:
                col += 1
    return scal_out, wav_out, scal_bl, bl_out, J, bounds


# ---------------------------------------------------------------------------- axisymmetric
def _axisym_params(B, L, J_min):
    return _p(B, L, J_min)


def analysis_axisym_lm_wav(flm_hp, B, L, J_min, spin_lowered=False):
    p = _axisym_params(B, L, J_min)
    J = _j_max(p)
    wav, scal = _ax_lm_wav(p)
    fw, fs = _ax_lm_ana(lm_hp2lm(flm_hp, L), wav, scal, p)
    hp, el, em = _hp_index(L)
    f_scal_hp = np.zeros(L * (L + 1) // 2, dtype=complex)
    f_scal_hp[hp] = fs[el * el + el + em]
    f_wav_hp = np.zeros((L * (L + 1) // 2, J - J_min + 1), dtype=complex)
    for j in range(J - J_min + 1):
        f_wav_hp[hp, j] = fw[L * L * j + el * el + el + em]
    return f_wav_hp, f_scal_hp


def synthesis_axisym_lm_wav(f_wav_lm_hp, f_scal_lm_hp, B, L, J_min, spin_lowered=False):
    p = _axisym_params(B, L, J_min)
    J = _j_max(p)
    wav, scal = _ax_lm_wav(p)
    f_wav_lm_hp = np.asarray(f_wav_lm_hp)
    fw = np.concatenate([lm_hp2lm(np.ascontiguousarray(f_wav_lm_hp[:, j]), L) for j in range(J - J_min + 1)])
    flm = _ax_lm_syn(fw, lm_hp2lm(f_scal_lm_hp, L), wav, scal, p)
    return lm2lm_hp(flm, L)


def analysis_axisym_wav_mw(f, B, L, J_min, spin_lowered=False):
    return _ax_ana(f, _axisym_params(B, L, J_min))


def analysis_adjoint_axisym_wav_mw(f_wav, f_scal, B, L, J_min, spin_lowered=False):
    return _ax_ana_adj(f_wav, f_scal, _axisym_params(B, L, J_min))


def synthesis_axisym_wav_mw(f_wav, f_scal, B, L, J_min, spin_lowered=False):
    return _ax_syn(f_wav, f_scal, _axisym_params(B, L, J_min))


def synthesis_adjoint_axisym_wav_mw(f, B, L, J_min, spin_lowered=False):
    return _ax_syn_adj(f, _axisym_params(B, L, J_min))


# ---------------------------------------------------------------------------- directional
def analysis_lm2wav(flm_hp, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _a_lm2wav(lm_hp2lm(flm_hp, L), _p(B, L, J_min, N, spin, upsample, original_spin))


def synthesis_wav2lm(f_wav, f_scal, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return lm2lm_hp(_s_wav2lm(f_wav, f_scal, _p(B, L, J_min, N, spin, upsample, original_spin)), L)


def analysis_px2wav(f, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _a_px2wav(f, _p(B, L, J_min, N, spin, upsample, original_spin))


def analysis_adjoint_wav2px(f_wav, f_scal, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _aa_wav2px(f_wav, f_scal, _p(B, L, J_min, N, spin, upsample, original_spin))


def synthesis_wav2px(f_wav, f_scal, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _s_wav2px(f_wav, f_scal, _p(B, L, J_min, N, spin, upsample, original_spin))


def synthesis_adjoint_px2wav(f, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _sa_px2wav(f, _p(B, L, J_min, N, spin, upsample, original_spin))


def analysis_lm2lmn(flm, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _a_lm2lmn(flm, _p(B, L, J_min, N, spin, upsample, original_spin))


def analysis_adjoint_lmn2lm(f_wav_lmn, f_scal_lm, B, L, J_min, N, spin, upsample, spin_lowered=False,
                            original_spin=0):
    return _aa_lmn2lm(f_wav_lmn, f_scal_lm, _p(B, L, J_min, N, spin, upsample, original_spin))


def synthesis_lmn2lm(f_wav_lmn, f_scal_lm, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _s_lmn2lm(f_wav_lmn, f_scal_lm, _p(B, L, J_min, N, spin, upsample, original_spin))


def synthesis_adjoint_lm2lmn(flm, B, L, J_min, N, spin, upsample, spin_lowered=False, original_spin=0):
    return _sa_lm2lmn(flm, _p(B, L, J_min, N, spin, upsample, original_spin))


def alm2map_mw(f_lm, L, spin):
    return _alm2map(f_lm, L, spin)


def map2alm_mw(f, L, spin):
    return _map2alm(f, L, spin)


# ---------------------------------------------------------------------------- manual tilings
def _manual_wavelets(L, N, spin, kappa_wav_l, J):
    """wav_lm[j] = sqrt((2l+1)/(8 pi^2)) kappa_j(l) s_lm with the directional s_lm of N
    (pys2let.pyx:500-513)."""
    s = np.asarray(_tiling_direction(_p(2.0, L, 0, N, spin)))
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    k = np.asarray(kappa_wav_l, dtype=float).reshape(J + 1, L)
    return (np.sqrt((2 * ell + 1) / (8.0 * np.pi * np.pi))[None, :] * k[:, ell] * s[None, :]).ravel()


def analysis_lm2wav_manualtiling(flm_hp, L, N, spin, kappa_scal_l, kappa_wav_l, scal_bandlimit, wav_bandlimits):
    J = len(wav_bandlimits) - 1
    ell = np.arange(L)
    scal_l = np.sqrt((2 * ell + 1) / (4.0 * np.pi)) * np.asarray(kappa_scal_l, dtype=float)[:L]
    t = ManualTiling(_manual_wavelets(L, N, spin, kappa_wav_l, J), scal_l, int(scal_bandlimit),
                     np.asarray(wav_bandlimits, dtype=np.int32), L, spin, N)
    f_wav, f_scal = _a_manual(lm_hp2lm(flm_hp, L), t)
    b = int(scal_bandlimit)
    out_scal = np.zeros(b * (2 * b - 1), dtype=complex)  # sized as pys2let.pyx:515-518
    out_scal[:f_scal.size] = f_scal
    return f_wav, out_scal


def synthesis_wav2lm_manualtiling(f_wav, f_scal, L, N, spin, kappa_scal_l, kappa_wav_l, scal_bandlimit,
                                  wav_bandlimits):
    J = len(wav_bandlimits) - 1
    ell = np.arange(L)
    scal_l = np.sqrt((2 * ell + 1) / (4.0 * np.pi)) * np.asarray(kappa_scal_l, dtype=float)[:L]
    t = ManualTiling(_manual_wavelets(L, N, spin, kappa_wav_l, J), scal_l, int(scal_bandlimit),
                     np.asarray(wav_bandlimits, dtype=np.int32), L, spin, N)
    return lm2lm_hp(_s_manual(f_wav, np.asarray(f_scal)[:t.n_scal()], t), L)


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("annotations", "np", "Parameters", "lib",
                                                                    "ManualTiling")]
