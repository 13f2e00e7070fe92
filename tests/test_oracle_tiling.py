"""Pins the oracle's tiling generation and scale schedule.

Checks: the reference's own tiling fixtures (src/matlab/kappas_*.mat, B=3, L=128,
J_min=2, via tests/golden/), the admissibility identities of
s2let_tiling.c:444-551 as tested in test/c/test_tiling.c:12-74 (1e-12), the
exact-vs-approximate binomial check of test/c/test_tiling.c:76-81, and the
closed-form scale schedule of the BASELINE configurations.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kappas_B3_L128_Jmin2.npz")


@pytest.mark.parametrize("wavtype,tag,tol", [(1, "s2dw", 0.0), (2, "need", 1e-12), (3, "spline", 0.0)])
def test_kappa_fixture(oracle, wavtype, tag, tol):
    d = np.load(GOLD)
    oracle.switch_wavtype(wavtype)
    try:
        p = oracle.params(128, 1, 3.0, 2)
        k, k0 = oracle.tiling_axisym(p)
    finally:
        oracle.switch_wavtype(1)
    assert k.shape == d["kappa_" + tag].shape
    assert np.abs(k[2:] - d["kappa_" + tag][2:]).max() <= tol
    assert np.abs(k0 - d["kappa0_" + tag]).max() <= tol


def _admissibility_axisym(k, k0, J_min):
    ident = k0 ** 2 + (k[J_min:] ** 2).sum(axis=0)
    return np.abs(ident - 1).max()


@pytest.mark.parametrize("B,L,J_min", [(2, 16, 2), (2, 64, 0), (3, 128, 2), (2, 512, 2), (1.5, 40, 1)])
def test_axisym_admissibility(oracle, B, L, J_min):
    p = oracle.params(L, 4, B, J_min)
    k, k0 = oracle.tiling_axisym(p)
    assert _admissibility_axisym(k, k0, J_min) < 1e-12


@pytest.mark.parametrize("L,N", [(16, 4), (16, 1), (32, 3), (64, 8)])
def test_direction_admissibility(oracle, L, N):
    p = oracle.params(L, N, 2.0, 2)
    s = oracle.tiling_direction(p)
    for el in range(1, L):
        seg = s[el * el:(el + 1) ** 2]
        assert abs(np.sum(np.abs(seg) ** 2) - 1) < 1e-12
        # azimuthal band-limit and parity: non-zero only for |m| < N, N+m odd
        ms = np.arange(-el, el + 1)
        bad = (np.abs(ms) >= N) | ((N + ms) % 2 == 0)
        assert np.all(seg[bad] == 0)


@pytest.mark.parametrize("L,N,J_min,spin", [(16, 4, 2, 0), (16, 4, 2, 2), (64, 4, 0, 0), (256, 4, 2, 2)])
def test_wavelet_admissibility(oracle, L, N, J_min, spin):
    p = oracle.params(L, N, 2.0, J_min, spin)
    psi, phi = oracle.tiling_wavelet(p)
    J = oracle.j_max(p)
    psi = psi.reshape(J + 1, L * L)
    for el in range(abs(spin), L):
        ident = 4 * np.pi / (2 * el + 1) * phi[el] ** 2
        ident += 8 * np.pi ** 2 / (2 * el + 1) * np.sum(np.abs(psi[:, el * el:(el + 1) ** 2]) ** 2)
        assert abs(ident - 1) < 1e-12


def test_binomial_exact_vs_logfact(oracle):
    # test/c/test_tiling.c:76-81 compares with cmocka's assert_float_equal, i.e. in
    # single precision with a relative FLT_EPSILON fallback.
    for n in range(1, 63):
        for k in range(n // 2 + 1):
            a, b = oracle.binomial(n, k, 0), oracle.binomial(n, k, 1)
            assert abs(a - b) <= max(1e-12, max(a, b) * np.finfo(np.float32).eps)
            assert b == round(__import__("math").comb(n, k))


@pytest.mark.parametrize("L,N,J_min,J,sched,L_s,n_wav,n_lmn", [
    (64, 4, 0, 6, [(2, 2, 1), (4, 4, 1), (8, 4, 2), (16, 4, 4), (32, 4, 8), (64, 4, 16), (64, 4, 32)], 1,
     132454, 66706),
    (512, 8, 2, 9, [(8, 8, 2), (16, 8, 4), (32, 8, 8), (64, 8, 16), (128, 8, 32), (256, 8, 64), (512, 8, 128),
                    (512, 8, 256)], 4, 18326520, 9172480),
    (2048, 8, 3, 11, None, 8, 293506800, 146796840),
])
def test_schedule(oracle, L, N, J_min, J, sched, L_s, n_wav, n_lmn):
    p = oracle.params(L, N, 2.0, J_min)
    assert oracle.j_max(p) == J
    assert min(oracle.bandlimit(J_min - 1, p), L) == L_s
    if sched is not None:
        for j, (Lj, Nj, L0) in zip(range(J_min, J + 1), sched):
            assert min(oracle.bandlimit(j, p), L) == Lj
            n = min(N, Lj)
            assert n + (n + N) % 2 == Nj
            assert oracle.L0(j, p) == L0
    assert oracle.n_wav(p) == n_wav
    assert oracle.n_lmn_wav(p) == n_lmn
