"""Pins of the oracle's axisymmetric wavelets (s2let_transform_axisym_lm.c,
s2let_transform_axisym_mw.c).

  * kernels: sum_j kappa_j(l)^2 + kappa_0(l)^2 = 1 through the lm kernels
    (admissibility, s2let_tiling.c:444-480);
  * the axisymmetric transforms equal the directional ones at N = 1, spin 0
    (test/c/test_axisym_vs_directional.c), full resolution = upsampled schedule,
    multiresolution = band-limited schedule, for all four operators;
  * round trips (test/c/test_back_and_forth.c:178-405) and adjoint identities
    (test/python/test_axisym_adjoints.py).
"""
import numpy as np
import pytest

STATES = [(16, 2.0, 2), (32, 3.0, 1), (24, 2.0, 0)]


def crandn(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.mark.parametrize("L,B,J_min", STATES)
def test_axisym_kernels_admissible(oracle, L, B, J_min):
    p = oracle.params(L, 1, B, J_min)
    wav, scal = oracle.axisym_lm_wav(p)
    J = oracle.j_max(p)
    ell = np.arange(L)
    f = np.sqrt(4 * np.pi / (2 * ell + 1))
    tot = (f * scal) ** 2 + sum((f * wav[j * L:(j + 1) * L]) ** 2 for j in range(J_min, J + 1))
    assert np.abs(tot - 1).max() < 1e-12


@pytest.mark.parametrize("L,B,J_min", STATES)
@pytest.mark.parametrize("multires", [0, 1])
def test_axisym_is_directional_n1(oracle, L, B, J_min, multires):
    p = oracle.params(L, 1, B, J_min)
    q = oracle.params(L, 1, B, J_min, 0, 0 if multires else 1)
    rng = np.random.default_rng(L + multires)
    f = crandn(rng, L * (2 * L - 1))
    aw, as_ = oracle.axisym_mw(0, multires, f, None, p)
    dw, ds = oracle.analysis_px2wav(f, q)
    assert aw.size == dw.size and np.abs(aw - dw).max() < 1e-12 * np.abs(dw).max()
    assert np.abs(as_ - ds).max() < 1e-12 * np.abs(ds).max()
    yw, ys = crandn(rng, aw.size), crandn(rng, as_.size)
    for kind, fn in ((1, oracle.analysis_adjoint_wav2px), (2, oracle.synthesis_wav2px)):
        got, want = oracle.axisym_mw(kind, multires, yw, ys, p), fn(yw, ys, q)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    bw, bs = oracle.axisym_mw(3, multires, f, None, p)
    cw, cs = oracle.synthesis_adjoint_px2wav(f, q)
    assert np.abs(bw - cw).max() < 1e-12 * np.abs(cw).max() and np.abs(bs - cs).max() < 1e-12 * np.abs(cs).max()


@pytest.mark.parametrize("L,B,J_min", STATES)
@pytest.mark.parametrize("multires", [0, 1])
def test_axisym_round_trips_and_adjoints(oracle, L, B, J_min, multires):
    p = oracle.params(L, 1, B, J_min)
    rng = np.random.default_rng(3 * L + multires)
    flm = crandn(rng, L * L)
    f = oracle.ssht_inverse(flm, L, 0)
    fw, fs = oracle.axisym_mw(0, multires, f, None, p)
    g = oracle.axisym_mw(2, multires, fw, fs, p)
    assert np.abs(oracle.ssht_forward(g, L, 0) - flm).max() < 1e-12
    # harmonic space with the transform's own kernels
    wav, scal = oracle.axisym_lm_wav(p)
    lw, ls = oracle.axisym_lm_wav_analysis(flm, wav, scal, p, multires)
    assert np.abs(oracle.axisym_lm_wav_synthesis(lw, ls, wav, scal, p, multires) - flm).max() < 1e-12
    # adjoints
    h = crandn(rng, f.size)
    yw, ys = crandn(rng, fw.size), crandn(rng, fs.size)
    lhs = np.vdot(yw, fw) + np.vdot(ys, fs)
    assert abs(lhs - np.vdot(oracle.axisym_mw(1, multires, yw, ys, p), f)) < 1e-10 * abs(lhs)
    bw, bs = oracle.axisym_mw(3, multires, h, None, p)
    lhs = np.vdot(h, oracle.axisym_mw(2, multires, yw, ys, p))
    assert abs(lhs - (np.vdot(bw, yw) + np.vdot(bs, ys))) < 1e-10 * abs(lhs)
