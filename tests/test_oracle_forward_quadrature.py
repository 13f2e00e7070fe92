"""Pins the oracle's forward MW / MWSS spherical-harmonic transform and forward
SO(3) transform on sample arrays that are NOT band-limited.

On band-limited maps the forward transform is fixed by the round trip; off that
subspace it is fixed by the sampling theorem's algorithm (McEwen & Wiaux 2011,
cited at README.md:74-85; the transforms are called at s2let_analysis.c:471,
s2let_synthesis.c:335,361): the samples of each phi-Fourier coefficient are
extended to the whole theta torus with F_m(2 pi - theta) = (-1)^{m+s} F_m(theta),
the unique trigonometric polynomial of degree L-1 in theta and in phi through
the extended samples is formed, and its harmonic coefficients are its exact
integrals against sY*_lm sin(theta).  Here that definition is restated
independently: numpy FFTs build the interpolant, Gauss-Legendre quadrature of
high order integrates it in theta (the integrand is a trigonometric polynomial
of degree <= 2L - 1, so 160 nodes are exact to rounding at these L) and the phi
integral is the exact 2 pi delta; sY comes from scipy (spin 0) or Wigner's
explicit d (spin s).  SO(3): the gamma samples are interpolated by their
(2N-1)-point DFT and each n is integrated like a spin -n map against
D^l_mn = e^{-im alpha} d^l_mn(beta) e^{-in gamma}.

MWSS (theta_t = pi t / L, t <= L; phi_p = pi p / L, p < 2L): 2L samples per
torus direction leave one Nyquist frequency; the definition used here (and by
the oracle) keeps frequencies |k| <= L - 1 only.  How the reference's SSHT
treats that Nyquist term on maps that are not band-limited is not known here
(SSHT is absent), so for MWSS this test pins the oracle to its stated
definition; for MW the definition is the sampling theorem's own.

A least-squares pseudo-inverse of the inverse transform agrees with the
quadrature on band-limited maps but not on these: the last test shows that it
fails the same bound, so the pin tells the two apart."""
from math import factorial, pi, sqrt

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss
from scipy.special import sph_harm_y

NGL = 160


def crandn(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def wigner_d_vec(l, mp, m, beta):
    """d^l_{mp,m}(beta) by Wigner's explicit sum, vectorised over beta."""
    beta = np.asarray(beta, float)
    c, s = np.cos(beta / 2), np.sin(beta / 2)
    out = np.zeros_like(beta)
    for k in range(0, 2 * l + 1):
        a, b, cc, d = l + m - k, k, l - k - mp, k - m + mp
        if min(a, b, cc, d) < 0:
            continue
        out += (-1) ** (k - m + mp) * c ** (2 * l - 2 * k + m - mp) * s ** (2 * k - m + mp) / (
            factorial(a) * factorial(b) * factorial(cc) * factorial(d))
    return out * sqrt(factorial(l + m) * factorial(l - m) * factorial(l + mp) * factorial(l - mp))


def gl_theta():
    x, w = leggauss(NGL)
    return pi * (x + 1) / 2, w * pi / 2


def sy(l, m, s, theta):
    """sY_lm(theta, 0) = (-1)^s sqrt((2l+1)/4pi) d^l_{m,-s}(theta) (scipy for spin 0)."""
    if s == 0:
        return sph_harm_y(l, m, theta, 0.0)
    return (-1) ** s * sqrt((2 * l + 1) / (4 * pi)) * wigner_d_vec(l, m, -s, theta)


def torus_interpolant(f, L, s, ss):
    """Coefficients G[k, m] (k, m = -(L-1)..L-1, stored at k + L - 1, m + L - 1) of the
    trigonometric polynomial sum_{k,m} G[k, m] e^{i k theta} e^{i m phi} through the
    theta-extended samples f (rows theta, columns phi); for MWSS the Nyquist terms are
    left out.  `s` is the extension sign exponent: F_m(2pi - theta) = (-1)^{m+s} F_m(theta)."""
    ms = np.arange(-(L - 1), L)
    if ss:
        P = 2 * L
        Fm = np.fft.fft(f, axis=1) / P                 # [t, m mod 2L], t = 0..L
        Fm = Fm[:, ms % P]
        T = 2 * L
        ext = np.zeros((T, 2 * L - 1), complex)
        ext[:L + 1] = Fm
        for t in range(L + 1, T):
            ext[t] = (-1.0) ** (ms + s) * Fm[T - t]
        G = np.fft.fft(ext, axis=0) / T                 # theta_t = 2 pi t / T
        return G[ms % T]
    M = 2 * L - 1
    Fm = np.fft.fft(f, axis=1) / M                      # [t, m mod M], t = 0..L-1
    Fm = Fm[:, ms % M]
    ext = np.zeros((M, M), complex)
    ext[:L] = Fm
    for t in range(L - 1):                              # theta_{M-1-t} = 2 pi - theta_t
        ext[M - 1 - t] = (-1.0) ** (ms + s) * Fm[t]
    G = np.fft.fft(ext, axis=0) / M                     # theta_t = 2 pi t / M + pi / M
    k = ms
    return G[k % M] * np.exp(-1j * k * pi / M)[:, None]


def interp_theta(G, L, theta):
    """F~_m(theta) = sum_k G[k, m] e^{i k theta}: [len(theta), 2L-1]."""
    k = np.arange(-(L - 1), L)
    return np.exp(1j * np.outer(theta, k)) @ G


def ssht_forward_by_quadrature(f, L, s, ss):
    th, w = gl_theta()
    Ft = interp_theta(torus_interpolant(f, L, s, ss), L, th)   # [node, m]
    flm = np.zeros(L * L, complex)
    for l in range(abs(s), L):
        for m in range(-l, l + 1):
            flm[l * l + l + m] = 2 * pi * np.sum(w * np.sin(th) * Ft[:, m + L - 1] * np.conj(sy(l, m, s, th)))
    return flm


def samples(L, ss):
    return (L + 1, 2 * L) if ss else (L, 2 * L - 1)


@pytest.mark.parametrize("ss", [0, 1])
@pytest.mark.parametrize("L,s", [(2, 0), (5, 0), (8, 0), (6, 2), (7, -2), (8, 1)])
def test_ssht_forward_is_interpolant_integral(oracle, L, s, ss):
    rng = np.random.default_rng(100 * L + 10 * s + ss)
    f = crandn(rng, np.prod(samples(L, ss)))
    ref = ssht_forward_by_quadrature(f.reshape(samples(L, ss)), L, s, ss)
    got = oracle.ssht_forward(f, L, s, ss)
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def so3_forward_by_quadrature(oracle, f, L, N, L0, ss):
    """f[g, t, p] (gamma slowest): f^l_mn = int f e^{-im a} d^l_mn(b) e^{-in g} sin b, by the
    (2N-1)-point gamma DFT and, per n, the theta/phi interpolant of a spin -n map."""
    G = 2 * N - 1
    nt, npx = samples(L, ss)
    fg = np.fft.fft(f.reshape(G, nt, npx), axis=0) / G   # [n mod G, t, p]: coefficient of e^{in gamma}
    th, w = gl_theta()
    out = np.zeros(oracle.flmn_size(L, N), complex)
    for n in range(-N + 1, N):
        # the n-slice carries e^{i m a} d^l_mn(b): extension sign (-1)^{m+n} (spin -n map)
        Ft = interp_theta(torus_interpolant(fg[n % G], L, -n, ss), L, th)
        for l in range(max(abs(n), L0), L):
            for m in range(-l, l + 1):
                d = wigner_d_vec(l, m, n, th)
                out[oracle.elmn2ind(l, m, n, L, N)] = 4 * pi * pi * np.sum(w * np.sin(th) * Ft[:, m + L - 1] * d)
    return out


@pytest.mark.parametrize("ss", [0, 1])
@pytest.mark.parametrize("L,N,L0", [(4, 2, 0), (6, 3, 0), (7, 3, 2)])
def test_so3_forward_is_interpolant_integral(oracle, L, N, L0, ss):
    rng = np.random.default_rng(1000 + 100 * L + 10 * N + ss)
    f = crandn(rng, (2 * N - 1) * np.prod(samples(L, ss)))
    ref = so3_forward_by_quadrature(oracle, f, L, N, L0, ss)
    got = oracle.so3_forward(f, L, N, L0, 0, ss)
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_pinned_definition_excludes_least_squares(oracle):
    """A least-squares inverse of the MW inverse transform agrees with the quadrature on
    band-limited maps (to 1e-12) but misses it by orders of magnitude on random samples."""
    L, s = 7, 0
    A = np.stack([oracle.ssht_inverse(np.eye(L * L)[k].astype(complex), L, s) for k in range(L * L)], axis=1)
    pinv = np.linalg.pinv(A)
    rng = np.random.default_rng(5)
    flm = crandn(rng, L * L)
    f_bl = oracle.ssht_inverse(flm, L, s)
    assert np.abs(pinv @ f_bl - ssht_forward_by_quadrature(f_bl.reshape(L, 2 * L - 1), L, s, 0)).max() < 1e-12
    f = crandn(rng, L * (2 * L - 1))
    ref = ssht_forward_by_quadrature(f.reshape(L, 2 * L - 1), L, s, 0)
    assert np.abs(oracle.ssht_forward(f, L, s) - ref).max() < 1e-12
    assert np.abs(pinv @ f - ref).max() > 1e-3
