"""Pins the oracle's SSHT (MW) and SO(3) transforms and the S2LET drivers.

SSHT: spin 0 against scipy's sph_harm_y (Condon-Shortley) at the MW nodes;
spin s against the direct sum sY_lm = (-1)^s sqrt((2l+1)/4pi) e^{im phi}
d^l_{m,-s}(theta) with Wigner's explicit d; forward by exact round trip.
SO(3): direct O(L^4 N) sum of f = sum (2l+1)/(8pi^2) f^l_mn e^{im a} d^l_mn(b) e^{in g}.
Adjoints: dot-product identities (test/c/test_adjoint.c style).
S2LET: the reference's round-trip, adjoint and N=1 axisym suites
(test/c/test_back_and_forth.c, test/c/test_adjoint.c,
test/c/test_axisym_vs_directional.c, test/python/test_directional_harmonic.py).
"""
from math import pi, sqrt

import numpy as np
import pytest
from scipy.special import sph_harm_y

from tests.test_oracle_wigner import wigner_d


def crandn(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def mw_theta(L):
    return pi * (2 * np.arange(L) + 1) / (2 * L - 1)


def mw_phi(L):
    return 2 * pi * np.arange(2 * L - 1) / (2 * L - 1)


def test_ssht_spin0_vs_scipy(oracle):
    rng = np.random.default_rng(11)
    L = 12
    flm = crandn(rng, L * L)
    f = oracle.ssht_inverse(flm, L, 0).reshape(L, 2 * L - 1)
    T, P = np.meshgrid(mw_theta(L), mw_phi(L), indexing="ij")
    ref = np.zeros_like(T, dtype=complex)
    for l in range(L):
        for m in range(-l, l + 1):
            ref += flm[l * l + l + m] * sph_harm_y(l, m, T, P)
    assert np.abs(f - ref).max() < 1e-12


@pytest.mark.parametrize("spin", [-2, -1, 1, 2, 3])
def test_ssht_spin_vs_direct(oracle, spin):
    rng = np.random.default_rng(12 + spin)
    L = 7
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    f = oracle.ssht_inverse(flm, L, spin).reshape(L, 2 * L - 1)
    th, ph = mw_theta(L), mw_phi(L)
    ref = np.zeros((L, 2 * L - 1), complex)
    for t in range(L):
        for l in range(abs(spin), L):
            for m in range(-l, l + 1):
                ref[t] += flm[l * l + l + m] * (-1) ** spin * sqrt((2 * l + 1) / (4 * pi)) * np.exp(
                    1j * m * ph) * wigner_d(l, m, -spin, th[t])
    assert np.abs(f - ref).max() < 1e-12


@pytest.mark.parametrize("L,spin", [(2, 0), (9, 0), (16, 2), (33, -1), (64, 0)])
def test_ssht_roundtrip_and_adjoints(oracle, L, spin):
    rng = np.random.default_rng(L + spin)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    f = oracle.ssht_inverse(flm, L, spin)
    assert np.abs(oracle.ssht_forward(f, L, spin) - flm).max() < 1e-12
    g = crandn(rng, L * (2 * L - 1))
    glm = crandn(rng, L * L)
    lhs = np.vdot(g, oracle.ssht_inverse(glm, L, spin))
    rhs = np.vdot(oracle.ssht_adjoint_inverse(g, L, spin), glm)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))
    lhs = np.vdot(glm, oracle.ssht_forward(g, L, spin))
    rhs = np.vdot(oracle.ssht_adjoint_forward(glm, L, spin), g)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))


def _so3_direct(oracle, flmn, L, N, nmode):
    al = 2 * pi * np.arange(2 * L - 1) / (2 * L - 1)
    ga = 2 * pi * np.arange(2 * N - 1) / (2 * N - 1)
    th = mw_theta(L)
    f = np.zeros((2 * N - 1, L, 2 * L - 1), complex)
    for n in range(-N + 1, N):
        if nmode == 1 and n % 2:
            continue
        if nmode == 2 and n % 2 == 0:
            continue
        for l in range(abs(n), L):
            for m in range(-l, l + 1):
                c = flmn[oracle.elmn2ind(l, m, n, L, N)] * (2 * l + 1) / (8 * pi * pi)
                d = np.array([wigner_d(l, m, n, b) for b in th])
                f += c * np.exp(1j * n * ga)[:, None, None] * d[None, :, None] * np.exp(1j * m * al)[None, None, :]
    return f.ravel()


@pytest.mark.parametrize("L,N,nmode", [(5, 3, 0), (6, 4, 2), (6, 3, 1)])
def test_so3_vs_direct(oracle, L, N, nmode):
    rng = np.random.default_rng(L * 10 + N)
    flmn = crandn(rng, oracle.flmn_size(L, N))
    f = oracle.so3_inverse(flmn, L, N, 0, nmode)
    assert np.abs(f - _so3_direct(oracle, flmn, L, N, nmode)).max() < 1e-12


@pytest.mark.parametrize("L,N,L0", [(8, 3, 0), (16, 4, 4), (12, 5, 2)])
def test_so3_roundtrip_adjoints(oracle, L, N, L0):
    rng = np.random.default_rng(L + N)
    flmn = crandn(rng, oracle.flmn_size(L, N))
    for n in range(-N + 1, N):
        for l in range(abs(n), min(L0, L)):
            for m in range(-l, l + 1):
                flmn[oracle.elmn2ind(l, m, n, L, N)] = 0
    f = oracle.so3_inverse(flmn, L, N, L0)
    assert np.abs(oracle.so3_forward(f, L, N, L0) - flmn).max() < 1e-12
    g = crandn(rng, f.size)
    lhs = np.vdot(g, oracle.so3_inverse(flmn, L, N, L0))
    rhs = np.vdot(oracle.so3_adjoint_inverse(g, L, N, L0), flmn)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))
    lhs = np.vdot(flmn, oracle.so3_forward(g, L, N, L0))
    rhs = np.vdot(oracle.so3_adjoint_forward(flmn, L, N, L0), g)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))


REF_STATES = [(16, 4, 2, s, u) for s in (0, 2) for u in (0, 1)]  # test_back_and_forth.c:577-611


def _uniform_flm(rng, L, spin):
    flm = rng.uniform(-1, 1, L * L) + 1j * rng.uniform(-1, 1, L * L)
    flm[:spin * spin] = 0
    return flm


@pytest.mark.parametrize("L,N,J_min,spin,up", REF_STATES + [(64, 4, 0, 0, 0), (32, 3, 1, 0, 0), (32, 1, 1, 0, 0)])
def test_s2let_back_and_forth(oracle, L, N, J_min, spin, up):
    p = oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(1)
    flm = _uniform_flm(rng, L, spin)
    wl, sl = oracle.analysis_lm2lmn(flm, p)
    assert np.abs(oracle.synthesis_lmn2lm(wl, sl, p) - flm).max() < 1e-12
    fw, fs = oracle.analysis_lm2wav(flm, p)
    assert np.abs(oracle.synthesis_wav2lm(fw, fs, p) - flm).max() < 1e-12
    f = oracle.ssht_inverse(flm, L, spin)
    fw, fs = oracle.analysis_px2wav(f, p)
    assert np.abs(oracle.synthesis_wav2px(fw, fs, p) - f).max() < 1e-12


@pytest.mark.parametrize("L,N,J_min,spin,up", REF_STATES)
def test_s2let_adjoints(oracle, L, N, J_min, spin, up):
    p = oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(2)
    x = _uniform_flm(rng, L, spin)
    yw, ys = crandn(rng, oracle.n_wav(p)), crandn(rng, oracle.n_scal(p))
    # analysis lm2wav vs its adjoint
    aw, as_ = oracle.analysis_lm2wav(x, p)
    lhs = np.vdot(yw, aw) + np.vdot(ys, as_)
    rhs = np.vdot(oracle.analysis_adjoint_wav2lm(yw, ys, p), x)
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))
    # synthesis wav2lm vs its adjoint
    y = crandn(rng, L * L)
    lhs = np.vdot(y, oracle.synthesis_wav2lm(yw, ys, p))
    bw, bs = oracle.synthesis_adjoint_lm2wav(y, p)
    rhs = np.vdot(bw, yw) + np.vdot(bs, ys)
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))
    # pixel-space pairs
    g = crandn(rng, L * (2 * L - 1))
    aw, as_ = oracle.analysis_px2wav(g, p)
    lhs = np.vdot(yw, aw) + np.vdot(ys, as_)
    rhs = np.vdot(oracle.analysis_adjoint_wav2px(yw, ys, p), g)
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))
    lhs = np.vdot(g, oracle.synthesis_wav2px(yw, ys, p))
    bw, bs = oracle.synthesis_adjoint_px2wav(g, p)
    rhs = np.vdot(bw, yw) + np.vdot(bs, ys)
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))
    # harmonic-space pairs (test_adjoint.c:28-120)
    wl, sl = oracle.analysis_lm2lmn(x, p)
    zl, zs = crandn(rng, wl.size), crandn(rng, sl.size)
    lhs = np.vdot(zl, wl) + np.vdot(zs, sl)
    rhs = np.vdot(oracle.analysis_adjoint_lmn2lm(zl, zs, p), x)
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))
    lhs = np.vdot(y, oracle.synthesis_lmn2lm(zl, zs, p))
    bl, bs2 = oracle.synthesis_adjoint_lm2lmn(y, p)
    rhs = np.vdot(bl, zl) + np.vdot(bs2, zs)
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))


def test_axisym_vs_directional_n1(oracle):
    # test/c/test_axisym_vs_directional.c:55-100 (multiresolution case)
    p = oracle.params(16, 1, 2.0, 2, 0, 0)
    rng = np.random.default_rng(3)
    flm = _uniform_flm(rng, 16, 0)
    f = oracle.ssht_inverse(flm, 16, 0)
    aw, as_ = oracle.axisym_wav_analysis_mw_multires(f, p)
    dw, ds = oracle.analysis_px2wav(f, p)
    assert np.abs(aw - dw[:aw.size]).max() < 1e-12
    assert np.abs(as_ - ds[:as_.size]).max() < 1e-12


def test_zero_tail_of_f_wav(oracle):
    # Allocation uses (2N-1) gamma slices per scale, transforms pack (2N_j-1):
    # c1 (L=64, N=4, J_min=0) leaves a 24-entry tail untouched.
    p = oracle.params(64, 4, 2.0, 0)
    rng = np.random.default_rng(4)
    fw, fs = oracle.analysis_lm2wav(crandn(rng, 64 * 64), p)
    assert fw.size == 132454
    assert np.all(fw[132430:] == 0)
