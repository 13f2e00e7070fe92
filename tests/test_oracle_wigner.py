"""Pins the oracle's Wigner Delta^l = d^l(pi/2) (Risbo recursion) against
Wigner's explicit sum (exact integer factorials) and the symmetry and
orthogonality relations; and the Delta factorisation of d^l(beta)."""
from fractions import Fraction
from math import cos, factorial, pi, sin, sqrt

import numpy as np
import pytest


def wigner_d(l, mp, m, beta):
    s = 0.0
    for k in range(0, 2 * l + 1):
        a, b, c, d = l + m - k, k, l - k - mp, k - m + mp
        if min(a, b, c, d) < 0:
            continue
        s += (-1) ** (k - m + mp) * cos(beta / 2) ** (2 * l - 2 * k + m - mp) * sin(beta / 2) ** (
            2 * k - m + mp) / (factorial(a) * factorial(b) * factorial(c) * factorial(d))
    return s * sqrt(factorial(l + m) * factorial(l - m) * factorial(l + mp) * factorial(l - mp))


def delta_exact(l, mp, m):
    """Wigner's sum at beta = pi/2 in exact rational arithmetic (cos = sin = 2^-1/2)."""
    s = Fraction(0)
    for k in range(0, 2 * l + 1):
        a, b, c, d = l + m - k, k, l - k - mp, k - m + mp
        if min(a, b, c, d) < 0:
            continue
        s += Fraction((-1) ** (k - m + mp), factorial(a) * factorial(b) * factorial(c) * factorial(d))
    return float(s) * sqrt(factorial(l + m) * factorial(l - m) * factorial(l + mp) * factorial(l - mp)) / 2.0 ** l


@pytest.mark.parametrize("l", [0, 1, 2, 3, 7, 12, 18, 40])
def test_delta_vs_wigner(oracle, l):
    D = oracle.delta_plane(l)
    for mp in range(-l, l + 1):
        for m in range(-l, l + 1):
            assert abs(D[mp + l, m + l] - delta_exact(l, mp, m)) < 1e-13


@pytest.mark.parametrize("l", [5, 31, 127, 300])
def test_delta_symmetry_orthogonality(oracle, l):
    D = oracle.delta_plane(l)
    ms = np.arange(-l, l + 1)
    # Delta^l_{m',-m} = (-1)^{l+m'} Delta^l_{m',m};  Delta^l_{m,m'} = (-1)^{m-m'} Delta^l_{m',m}
    assert np.abs(D[:, ::-1] - ((-1.0) ** (l + ms))[:, None] * D).max() < 1e-13
    assert np.abs(D.T - ((-1.0) ** (ms[:, None] - ms[None, :])) * D).max() < 1e-13
    assert np.abs(D.T @ D - np.eye(2 * l + 1)).max() < 1e-12


def test_dl_beta_and_factorisation(oracle):
    L, beta = 9, 0.73
    planes = oracle.dl_beta_full(L, beta)
    for l in range(L):
        D = oracle.delta_plane(l)
        for m in range(-l, l + 1):
            for n in range(-l, l + 1):
                ref = wigner_d(l, m, n, beta)
                assert abs(planes[l][m + l, n + l] - ref) < 1e-13
                # d^l_mn(beta) = i^{n-m} sum_{m'} Delta_{m'm} Delta_{m'n} e^{i m' beta}
                v = (1j) ** (n - m) * np.sum(D[:, m + l] * D[:, n + l] * np.exp(1j * np.arange(-l, l + 1) * beta))
                assert abs(v - ref) < 1e-13
