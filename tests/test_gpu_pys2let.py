"""The reference's Python tests run through the pys2let-compatible module:
test/python/test_axisym_adjoints.py and test/python/test_directional_harmonic.py
(same parametrizations: L=10, B=2, J_min=2; adjoint dot products), the hybrid
manual tiling round trip of test/python/directional_manual.py (L=256, N=3) and the
axisymmetric harmonic round trip in healpy ordering."""
from functools import partial

import numpy as np
import pytest
from pytest import approx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ps():
    import s2let_amd
    s2let_amd.lib()
    from s2let_amd import pys2let
    return pys2let


def random_mw_map(ps, rng, L, spin):
    return ps.alm2map_mw(rng.uniform(size=(L * L, 2)) @ [1, 1j], L, spin)


def random_wavlet_maps(ps, rng, L, spin, nwvlts):
    maps = [random_mw_map(ps, rng, L, spin) for _ in range(nwvlts + 1)]
    return maps[0], np.concatenate(maps[1:])


PX_CASES = [("analysis_axisym_wav_mw", "analysis_adjoint_axisym_wav_mw", {}, 0),
            ("synthesis_adjoint_axisym_wav_mw", "synthesis_axisym_wav_mw", {}, 0),
            ("analysis_px2wav", "analysis_adjoint_wav2px", dict(spin=0, upsample=1, N=1), 0),
            ("analysis_px2wav", "analysis_adjoint_wav2px", dict(spin=2, upsample=1, N=1), 2),
            ("synthesis_adjoint_px2wav", "synthesis_wav2px", dict(spin=0, upsample=1, N=1), 0),
            ("synthesis_adjoint_px2wav", "synthesis_wav2px", dict(spin=2, upsample=1, N=1), 2)]


@pytest.mark.parametrize("fwd,adj,kw,spin", PX_CASES)
def test_axisym_adjoint(ps, fwd, adj, kw, spin, L=10, B=2, J_min=2):
    """test/python/test_axisym_adjoints.py"""
    rng = np.random.default_rng(1)
    px2wav, wav2px = partial(getattr(ps, fwd), **kw), partial(getattr(ps, adj), **kw)
    nwvlts = ps.pys2let_j_max(B, L, J_min) - J_min + 1
    x = random_mw_map(ps, rng, L, spin)
    y_scal, y_wav = random_wavlet_maps(ps, rng, L, spin, nwvlts)
    y = wav2px(y_wav, y_scal, B, L, J_min)
    x_wav, x_scal = px2wav(x, B, L, J_min)
    yAx = y_wav.conj().T @ x_wav + y_scal.conj() @ x_scal
    assert yAx == approx(y.conj().T @ x)


LMN_CASES = [("analysis_lm2lmn", "analysis_adjoint_lmn2lm", 0), ("analysis_lm2lmn", "analysis_adjoint_lmn2lm", 2),
             ("synthesis_adjoint_lm2lmn", "synthesis_lmn2lm", 0), ("synthesis_adjoint_lm2lmn", "synthesis_lmn2lm", 2)]


@pytest.mark.parametrize("fwd,adj,spin", LMN_CASES)
def test_directional_harmonic_adjoint(ps, fwd, adj, spin, L=10, B=2, J_min=2):
    """test/python/test_directional_harmonic.py"""
    rng = np.random.default_rng(2)
    kw = dict(spin=spin, upsample=1, N=1)
    px2wav, wav2px = partial(getattr(ps, fwd), **kw), partial(getattr(ps, adj), **kw)
    nwvlts = ps.pys2let_j_max(B, L, J_min) - J_min + 1
    lms = [rng.uniform(size=(L * L, 2)) @ [1, 1j] for _ in range(nwvlts + 2)]
    x, y_scal, y_wav = lms[0], lms[1], np.concatenate(lms[2:])
    y = wav2px(y_wav, y_scal, B, L, J_min)
    x_wav, x_scal = px2wav(x, B, L, J_min)
    yAx = y_wav.conj().T @ x_wav + y_scal.conj() @ x_scal
    assert yAx == approx(y.conj().T @ x)


def _real_hp(rng, L):
    hp = rng.standard_normal(L * (L + 1) // 2) + 1j * rng.standard_normal(L * (L + 1) // 2)
    hp[:L] = hp[:L].real
    return hp


def test_hybrid_manual_round_trip(ps):
    """test/python/directional_manual.py (second part): hybrid tiling, N=3, L=256."""
    L, N, spin = 256, 3, 0
    scal_l, wav_l, scal_bl, wav_bl, J, _ = ps.construct_hybrid_tiling(L, 0, np.array([4, 2]), np.array([64]))
    wav_l = wav_l.T.ravel()
    rng = np.random.default_rng(3)
    f_lm = _real_hp(rng, L)
    f_wav, f_scal = ps.analysis_lm2wav_manualtiling(f_lm, L, N, spin, scal_l, wav_l, scal_bl, wav_bl)
    f_lm_rec = ps.synthesis_wav2lm_manualtiling(f_wav, f_scal, L, N, spin, scal_l, wav_l, scal_bl, wav_bl)
    assert np.abs(f_lm_rec - f_lm).max() < 1e-10 * np.abs(f_lm).max()


def test_hp_transforms_round_trip(ps):
    """healpy-ordered harmonic inputs (real signals) through the axisymmetric and directional APIs."""
    L, B, J_min = 32, 2.0, 2
    rng = np.random.default_rng(4)
    f_lm = _real_hp(rng, L)
    w, s = ps.analysis_axisym_lm_wav(f_lm, B, L, J_min)
    assert w.shape == (L * (L + 1) // 2, ps.pys2let_j_max(B, L, J_min) - J_min + 1)
    assert np.abs(ps.synthesis_axisym_lm_wav(w, s, B, L, J_min) - f_lm).max() < 1e-12
    fw, fs = ps.analysis_lm2wav(f_lm, B, L, J_min, 4, 0, 0)
    assert np.abs(ps.synthesis_wav2lm(fw, fs, B, L, J_min, 4, 0, 0) - f_lm).max() < 1e-12
