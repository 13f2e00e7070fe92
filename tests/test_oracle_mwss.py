"""Pins the oracle's MWSS sampling (S2LET_SAMPLING_MW_SS, s2let_types.h:26-32):
theta_t = pi t / L (t = 0..L, both poles), phi_p = pi p / L (p < 2L), gamma as MW.

SSHT: spin 0 against scipy's sph_harm_y at the MWSS nodes, spin s against the
direct sum with Wigner's explicit d; forward by exact round trip on band-limited
signals; adjoints by dot products.  SO(3): direct sum of the definition at the
MWSS nodes, round trip, adjoints.  S2LET: the reference's MWSS round trips
(test/c/test_back_and_forth.c:445-526) and sizes (s2let_helper.c:5-20,65-95).
The forward quadrature of maps that are not band-limited (the torus Nyquist
term) has nothing to pin it here: parity unpinned there.
"""
from math import pi, sqrt

import numpy as np
import pytest
from scipy.special import sph_harm_y

from tests.test_oracle_wigner import wigner_d


def crandn(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def ss_theta(L):
    return pi * np.arange(L + 1) / L


def ss_phi(L):
    return pi * np.arange(2 * L) / L


def test_ssht_mwss_spin0_vs_scipy(oracle):
    rng = np.random.default_rng(21)
    L = 10
    flm = crandn(rng, L * L)
    f = oracle.ssht_inverse(flm, L, 0, ss=1).reshape(L + 1, 2 * L)
    T, P = np.meshgrid(ss_theta(L), ss_phi(L), indexing="ij")
    ref = np.zeros_like(T, dtype=complex)
    for l in range(L):
        for m in range(-l, l + 1):
            ref += flm[l * l + l + m] * sph_harm_y(l, m, T, P)
    assert np.abs(f - ref).max() < 1e-12


@pytest.mark.parametrize("spin", [-2, 1, 3])
def test_ssht_mwss_spin_vs_direct(oracle, spin):
    rng = np.random.default_rng(30 + spin)
    L = 7
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    f = oracle.ssht_inverse(flm, L, spin, ss=1).reshape(L + 1, 2 * L)
    th, ph = ss_theta(L), ss_phi(L)
    ref = np.zeros((L + 1, 2 * L), complex)
    for t in range(L + 1):
        for l in range(abs(spin), L):
            for m in range(-l, l + 1):
                ref[t] += flm[l * l + l + m] * (-1) ** spin * sqrt((2 * l + 1) / (4 * pi)) * np.exp(
                    1j * m * ph) * wigner_d(l, m, -spin, th[t])
    assert np.abs(f - ref).max() < 1e-12


@pytest.mark.parametrize("L,spin", [(2, 0), (8, 0), (15, 2), (16, -1), (32, 0)])
def test_ssht_mwss_roundtrip_and_adjoints(oracle, L, spin):
    rng = np.random.default_rng(L + 3 * spin)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    f = oracle.ssht_inverse(flm, L, spin, ss=1)
    assert f.size == (L + 1) * 2 * L
    assert np.abs(oracle.ssht_forward(f, L, spin, ss=1) - flm).max() < 1e-12
    g = crandn(rng, f.size)
    glm = crandn(rng, L * L)
    lhs = np.vdot(g, oracle.ssht_inverse(glm, L, spin, ss=1))
    rhs = np.vdot(oracle.ssht_adjoint_inverse(g, L, spin, ss=1), glm)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))
    lhs = np.vdot(glm, oracle.ssht_forward(g, L, spin, ss=1))
    rhs = np.vdot(oracle.ssht_adjoint_forward(glm, L, spin, ss=1), g)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))


def _so3_direct_ss(oracle, flmn, L, N, nmode):
    al = pi * np.arange(2 * L) / L
    ga = 2 * pi * np.arange(2 * N - 1) / (2 * N - 1)
    th = ss_theta(L)
    f = np.zeros((2 * N - 1, L + 1, 2 * L), complex)
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


@pytest.mark.parametrize("L,N,nmode", [(5, 3, 0), (6, 4, 2)])
def test_so3_mwss_vs_direct(oracle, L, N, nmode):
    rng = np.random.default_rng(L * 7 + N)
    flmn = crandn(rng, oracle.flmn_size(L, N))
    f = oracle.so3_inverse(flmn, L, N, 0, nmode, ss=1)
    assert np.abs(f - _so3_direct_ss(oracle, flmn, L, N, nmode)).max() < 1e-12


@pytest.mark.parametrize("L,N,L0", [(8, 3, 0), (12, 4, 3)])
def test_so3_mwss_roundtrip_adjoints(oracle, L, N, L0):
    rng = np.random.default_rng(L + N + 1)
    flmn = crandn(rng, oracle.flmn_size(L, N))
    for n in range(-N + 1, N):
        for l in range(abs(n), min(L0, L)):
            for m in range(-l, l + 1):
                flmn[oracle.elmn2ind(l, m, n, L, N)] = 0
    f = oracle.so3_inverse(flmn, L, N, L0, ss=1)
    assert np.abs(oracle.so3_forward(f, L, N, L0, ss=1) - flmn).max() < 1e-12
    g = crandn(rng, f.size)
    lhs = np.vdot(g, oracle.so3_inverse(flmn, L, N, L0, ss=1))
    rhs = np.vdot(oracle.so3_adjoint_inverse(g, L, N, L0, ss=1), flmn)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))
    lhs = np.vdot(flmn, oracle.so3_forward(g, L, N, L0, ss=1))
    rhs = np.vdot(oracle.so3_adjoint_forward(flmn, L, N, L0, ss=1), g)
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))


REF_STATES = [(16, 4, 2, s, u) for s in (0, 2) for u in (0, 1)]  # test_back_and_forth.c:577-611


@pytest.mark.parametrize("L,N,J_min,spin,up", REF_STATES)
def test_s2let_mwss_sizes_and_round_trips(oracle, L, N, J_min, spin, up):
    p = oracle.params(L, N, 2.0, J_min, spin, up, sampling=1)
    # sizes: (L_j + 1) 2 L_j (2N - 1) per scale, 2 Ls (Ls + 1) for the scaling map
    tot = 0
    for j in range(J_min, oracle.j_max(p) + 1):
        b = L if up else min(oracle.bandlimit(j, p), L)
        tot += (b + 1) * 2 * b * (2 * N - 1)
    assert oracle.n_wav(p) == tot
    bs = L if up else min(oracle.bandlimit(J_min - 1, p), L)
    assert oracle.n_scal(p) == 2 * bs * (bs + 1)
    rng = np.random.default_rng(L + spin + up)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    fw, fs = oracle.analysis_lm2wav(flm, p)
    assert np.abs(oracle.synthesis_wav2lm(fw, fs, p) - flm).max() < 1e-12
    # test_back_and_forth.c:445-490: map -> wavelets -> map -> flm
    f = oracle.ssht_inverse(flm, L, spin, ss=1)
    fw, fs = oracle.analysis_px2wav(f, p)
    g = oracle.synthesis_wav2px(fw, fs, p)
    assert np.abs(oracle.ssht_forward(g, L, spin, ss=1) - flm).max() < 1e-12
    # adjoint identities of the drivers
    yw, ys = crandn(rng, fw.size), crandn(rng, fs.size)
    lhs = np.vdot(yw, fw) + np.vdot(ys, fs)
    rhs = np.vdot(oracle.analysis_adjoint_wav2px(yw, ys, p), f)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    h = crandn(rng, f.size)
    bw, bs_ = oracle.synthesis_adjoint_px2wav(h, p)
    lhs = np.vdot(h, oracle.synthesis_wav2px(yw, ys, p))
    rhs = np.vdot(bw, yw) + np.vdot(bs_, ys)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
