"""The pys2let-compatible module (s2let_amd.pys2let, mirroring
src/pys2let/pys2let.pyx): host-side pieces on the CPU -- healpy ordering, wav_ind,
tilings against the oracle, hybrid tilings (test/python/directional_manual.py)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ps():
    import s2let_amd
    s2let_amd.lib()
    from s2let_amd import pys2let
    return pys2let


def test_healpy_ordering(ps):
    L = 9
    rng = np.random.default_rng(0)
    hp = rng.standard_normal(L * (L + 1) // 2) + 1j * rng.standard_normal(L * (L + 1) // 2)
    hp[:L] = hp[:L].real  # m = 0 entries of a real signal are real
    flm = ps.lm_hp2lm(hp, L)
    assert np.array_equal(ps.lm2lm_hp(flm, L), hp)
    for el in range(L):
        for em in range(0, el + 1):
            assert flm[el * el + el + em] == hp[ps.healpy_lm(el, em, L)]
            assert flm[el * el + el - em] == (-1) ** em * np.conj(hp[ps.healpy_lm(el, em, L)])
    assert ps.mw_lm(3, 2) == 15 and ps.mw_size(L) == L * (2 * L - 1)


def test_wav_ind(ps, oracle):
    B, L, N, J_min = 2.0, 32, 4, 2
    for up in (0, 1):
        q = oracle.params(L, N, B, J_min, 0, up)
        off = 0
        for j in range(J_min, ps.pys2let_j_max(B, L, J_min) + 1):
            bl = L if up else min(oracle.bandlimit(j, q), L)
            for n in range(2 * N - 1):
                o, b, nmw, nj = ps.wav_ind(j, n, B, L, N, J_min, up)
                assert (o, b, nmw, nj) == (off + n * bl * (2 * bl - 1), bl, bl * (2 * bl - 1),
                                           bl * (2 * bl - 1) * (2 * N - 1))
            off += bl * (2 * bl - 1) * (2 * N - 1)


def test_tilings_match_oracle(ps, oracle):
    B, L, N, J_min, spin = 2.0, 32, 4, 2, 2
    q = oracle.params(L, N, B, J_min, spin)
    psi, phi = oracle.tiling_wavelet(q)
    scal_l, wav_l = ps.wavelet_tiling(B, L, N, J_min, spin)
    J = oracle.j_max(q)
    assert np.array_equal(scal_l.real, phi)
    assert np.array_equal(wav_l[spin * spin:, :], psi.reshape(J + 1, L * L)[J_min:, spin * spin:].T)
    k, k0 = oracle.tiling_axisym(oracle.params(L, 1, B, J_min))
    s, w = ps.axisym_wav_l(B, L, J_min)
    assert np.array_equal(s, k0) and np.array_equal(w, k[J_min:].T)


# the two tilings of test/python/directional_manual.py
@pytest.mark.parametrize("L,J_min,Bs,Lt", [(256, 0, [4, 2], [64]), (3024, 3, [2, 1.4, 1.2, 1.1], [200, 1100, 1900])])
def test_hybrid_tiling_admissible(ps, L, J_min, Bs, Lt):
    scal_l, wav_l, scal_bl, wav_bl, J, bounds = ps.construct_hybrid_tiling(L, J_min, np.array(Bs), np.array(Lt))
    assert wav_l.shape == (L, J + 1) and wav_bl.size == J + 1 and bounds[-1] == L
    assert ps.verify_tiling(L, scal_l, wav_l.T.ravel(), scal_bl, wav_bl)
