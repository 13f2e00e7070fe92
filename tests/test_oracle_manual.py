"""Pins of the oracle's manual-tiling restatement (s2let_analysis.c:188-276,
s2let_synthesis.c:202-289, s2let_analysis_adjoint.c:202-290,
s2let_synthesis_adjoint.c:189-290).

  * With the standard tiling shifted to j = 0..J (as test/c/test_back_and_forth.c:97-125
    builds it) the manual transforms equal the standard ones.
  * Round trip synthesis(analysis(flm)) = flm (test_back_and_forth.c:127-171, 1e-12).
  * The to-flm forms add into flm.
  * The manual adjoints are the exact adjoints of the manual transforms built on
    conj(wav_lm) (dot-product identities with random complex wavelets).
"""
import numpy as np
import pytest

STATES = [(16, 4, 2, 0, 1), (16, 4, 2, 0, 0), (16, 4, 2, 2, 1), (16, 4, 2, 2, 0), (24, 3, 1, 0, 0)]


def crandn(rng, n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)


def standard_as_manual(oracle, p):
    """test/c/test_back_and_forth.c:97-125."""
    L = p.L
    J = oracle.j_max(p) - p.J_min
    psi, phi = oracle.tiling_wavelet(p)
    scal_bl = L if p.upsample else min(oracle.bandlimit(p.J_min - 1, p), L)
    wav_bl = [L if p.upsample else min(oracle.bandlimit(p.J_min + j, p), L) for j in range(J + 1)]
    wav = psi.reshape(-1, L * L)[p.J_min:p.J_min + J + 1].ravel()
    return oracle.Manual(wav, phi, scal_bl, wav_bl, L, p.spin, p.N)


def cdot(a, b):
    return np.vdot(a, b)


@pytest.mark.parametrize("L,N,J_min,spin,up", STATES)
def test_manual_with_standard_tiling_is_standard(oracle, L, N, J_min, spin, up):
    p = oracle.params(L, N, 2.0, J_min, spin, up)
    m = standard_as_manual(oracle, p)
    rng = np.random.default_rng(L + N + spin + up)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    fw, fs = oracle.analysis_lm2wav_manual(flm, m)
    ow, os_ = oracle.analysis_lm2wav(flm, p)
    n = fw.size
    assert n <= ow.size and not np.any(ow[n:])
    assert np.abs(fw - ow[:n]).max() < 1e-13 * np.abs(ow).max()
    assert np.abs(fs - os_).max() < 1e-13 * np.abs(os_).max()
    yw, ys = crandn(rng, n), crandn(rng, fs.size)
    want = oracle.synthesis_wav2lm(np.concatenate([yw, np.zeros(ow.size - n)]), ys, p)
    got = oracle.synthesis_wav2lm_manual(yw, ys, m)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    # round trip
    rec = oracle.synthesis_wav2lm_manual(fw, fs, m)
    assert np.abs(rec - flm).max() < 1e-12
    # accumulation into flm
    base = crandn(rng, L * L)
    acc = oracle.synthesis_wav2lm_manual(fw, fs, m, flm=base)
    assert np.abs(acc - (base + rec)).max() < 1e-13
    acc = oracle.analysis_adjoint_wav2lm_manual(yw, ys, m, flm=base)
    assert np.abs(acc - (base + oracle.analysis_adjoint_wav2lm_manual(yw, ys, m))).max() < 1e-13


def random_manual(oracle, rng, L, N, spin, wav_bl, scal_bl):
    J = len(wav_bl) - 1
    wav = crandn(rng, (J + 1) * L * L)
    for j, b in enumerate(wav_bl):  # entries above each band-limit are never read
        blk = wav[j * L * L:(j + 1) * L * L]
        blk[min(b, L) ** 2:] = 0
    return oracle.Manual(wav, rng.standard_normal(L), scal_bl, wav_bl, L, spin, N)


@pytest.mark.parametrize("L,N,spin,wav_bl,scal_bl", [(12, 3, 0, [4, 8, 12], 4), (12, 4, 1, [6, 12, 12], 6),
                                                    (10, 2, 0, [10], 10)])
def test_manual_adjoints_are_adjoints_of_conjugate_wavelets(oracle, L, N, spin, wav_bl, scal_bl):
    rng = np.random.default_rng(L * N)
    m = random_manual(oracle, rng, L, N, spin, wav_bl, scal_bl)
    mc = oracle.Manual(np.conj(m.wav_lm), m.scal_l, scal_bl, wav_bl, L, spin, N)
    x = crandn(rng, L * L)
    x[:spin * spin] = 0
    yw, ys = crandn(rng, m.n_wav()), crandn(rng, m.n_scal())
    # analysis adjoint: <A_conj x, y> = <x, A*_manual y>
    aw, as_ = oracle.analysis_lm2wav_manual(x, mc)
    lhs = cdot(aw, yw) + cdot(as_, ys)
    rhs = cdot(x, oracle.analysis_adjoint_wav2lm_manual(yw, ys, m))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # synthesis adjoint: <S_conj y, x> = <y, S*_manual x>
    lhs = cdot(oracle.synthesis_wav2lm_manual(yw, ys, mc), x)
    bw, bs = oracle.synthesis_adjoint_lm2wav_manual(x, m)
    rhs = cdot(yw, bw) + cdot(ys, bs)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # and not of the unconjugated ones (complex wavelets): the quirk is real
    aw, as_ = oracle.analysis_lm2wav_manual(x, m)
    lhs2 = cdot(aw, yw) + cdot(as_, ys)
    assert abs(lhs2 - cdot(x, oracle.analysis_adjoint_wav2lm_manual(yw, ys, m))) > 1e-6 * abs(lhs2)
