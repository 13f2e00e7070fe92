"""Pins of the real-signal transforms on the CPU oracle.

The reference's real transforms (s2let_analysis.c:102-160, 371-442;
s2let_synthesis.c:101-170; s2let_analysis_adjoint.c:101-180;
s2let_synthesis_adjoint.c:103-165) store only the n >= 0 Wigner coefficients
of each scale and take the rest from the symmetry of a real signal on SO(3),
f_{l,m,-n} = (-1)^{m+n} conj(f_{l,-m,n}).  The library computes them as the
complex transforms on Hermitian harmonics / real samples.  These tests show on
the oracle that the two agree:
  * every wavelet obeys Psi_{l,-n} = (-1)^n conj(Psi_{ln});
  * the SO(3) forward transform of a real signal has that symmetry;
  * the reference's n >= 0 loops (restated below) give the complex oracle's
    result on symmetric coefficients;
  * the wavelet coefficients of a Hermitian flm are real.
"""
import numpy as np
import pytest

STATES = [(16, 4, 2, 0, 1), (16, 4, 2, 2, 1), (16, 4, 2, 2, 0),  # test/c/test_adjoint.c:461-480
          (24, 3, 1, 0, 0), (20, 5, 0, -2, 0), (32, 8, 1, 0, 0)]


def crandn(rng, n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)


def hermitian_flm(rng, L):
    """Harmonic coefficients of a real signal: f_{l,-m} = (-1)^m conj(f_lm)."""
    f = crandn(rng, L * L)
    for l in range(L):
        f[l * l + l] = f[l * l + l].real
        for m in range(1, l + 1):
            f[l * l + l - m] = (-1) ** m * np.conj(f[l * l + l + m])
    return f


def scales(oracle, p):
    """(offset, L_j, N_j) of every wavelet scale in the Wigner-coefficient array
    (s2let_synthesis.c:130-137)."""
    out, off = [], 0
    for j in range(p.J_min, oracle.j_max(p) + 1):
        Lj, Nj = p.L, p.N
        if not p.upsample:
            Lj = min(oracle.bandlimit(j, p), p.L)
            Nj = min(p.N, Lj)
            Nj += (Nj + p.N) % 2
        out.append((off, Lj, Nj))
        off += oracle.flmn_size(Lj, Nj)
    return out


def symmetrise(oracle, g, p):
    """Projects Wigner coefficients onto those of a real SO(3) signal."""
    h = g.copy()
    for off, Lj, Nj in scales(oracle, p):
        for n in range(-Nj + 1, Nj):
            for l in range(abs(n), Lj):
                for m in range(-l, l + 1):
                    a = off + oracle.elmn2ind(l, m, n, Lj, Nj)
                    b = off + oracle.elmn2ind(l, -m, -n, Lj, Nj)
                    h[a] = 0.5 * (g[a] + (-1) ** (m + n) * np.conj(g[b]))
    return h


def is_real_so3(oracle, g, p, tol):
    for off, Lj, Nj in scales(oracle, p):
        for n in range(-Nj + 1, Nj):
            for l in range(abs(n), Lj):
                for m in range(-l, l + 1):
                    a = g[off + oracle.elmn2ind(l, m, n, Lj, Nj)]
                    b = g[off + oracle.elmn2ind(l, -m, -n, Lj, Nj)]
                    if abs(a - (-1) ** (m + n) * np.conj(b)) > tol:
                        return False
    return True


def real_lmn2lm(oracle, g, s_lm, p, adjoint, flip=1):
    """The reference's real synthesis / real analysis adjoint from n >= 0 Wigner
    coefficients (s2let_synthesis.c:139-170, s2let_analysis_adjoint.c:139-176)."""
    L = p.L
    psi, phi = oracle.tiling_wavelet(p)
    psi = psi.reshape(-1, L * L)
    flm = np.zeros(L * L, complex)
    for jj, (off, Lj, Nj) in enumerate(scales(oracle, p)):
        j = p.J_min + jj
        for n in range(1 - Nj % 2, Nj, 2):
            for el in range(n, Lj):
                w = 8 * np.pi ** 2 / (2 * el + 1) if adjoint else 1.0
                ps = w * psi[j, el * el + el + n]
                nps = w * psi[j, el * el + el - n]
                for m in range(-el, el + 1):
                    flm[el * el + el + m] += g[off + oracle.elmn2ind(el, m, n, Lj, Nj)] * ps
                    if n:
                        sign = (-1 if (m + n) % 2 else 1) * flip
                        flm[el * el + el + m] += sign * np.conj(g[off + oracle.elmn2ind(el, -m, n, Lj, Nj)]) * nps
    bl = p.L if p.upsample else min(oracle.bandlimit(p.J_min - 1, p), p.L)
    for el in range(bl):
        ph = np.sqrt(4 * np.pi / (2 * el + 1)) * phi[el]
        for m in range(-el, el + 1):
            flm[el * el + el + m] += s_lm[el * el + el + m] * ph
    return flm


@pytest.mark.parametrize("L,N,J_min,spin,up", STATES)
def test_wavelets_conjugate_symmetric(oracle, L, N, J_min, spin, up):
    p = oracle.params(L, N, 2.0, J_min, spin, up)
    psi, _ = oracle.tiling_wavelet(p)
    psi = psi.reshape(-1, L * L)
    assert np.abs(psi).max() > 0.1
    for l in range(L):
        for n in range(1, min(l, N) + 1):
            np.testing.assert_allclose(psi[:, l * l + l - n], (-1) ** n * np.conj(psi[:, l * l + l + n]),
                                       rtol=0, atol=1e-15)


@pytest.mark.parametrize("L,N", [(8, 4), (9, 3), (6, 1)])
def test_so3_forward_of_real_signal_is_symmetric(oracle, L, N):
    rng = np.random.default_rng(L + N)
    f = rng.standard_normal(L * (2 * L - 1) * (2 * N - 1)).astype(complex)
    g = oracle.so3_forward(f, L, N)
    # one scale of size (L, N) at offset 0
    p = oracle.params(L, N, 2.0, 0, 0, 1)
    for n in range(-N + 1, N):
        for l in range(abs(n), L):
            for m in range(-l, l + 1):
                a = g[oracle.elmn2ind(l, m, n, L, N)]
                b = g[oracle.elmn2ind(l, -m, -n, L, N)]
                assert abs(a - (-1) ** (m + n) * np.conj(b)) < 1e-13
    assert p.L == L


@pytest.mark.parametrize("L,N,J_min,spin,up", STATES[:3] + [STATES[3]])
def test_real_lmn_loops_equal_complex(oracle, L, N, J_min, spin, up):
    p = oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(11 * L + N + spin)
    g = symmetrise(oracle, crandn(rng, oracle.n_lmn_wav(p)), p)
    assert is_real_so3(oracle, g, p, 1e-15)
    s = hermitian_flm(rng, L)[:oracle.n_lm_scal(p)]
    s = np.concatenate([s, np.zeros(L * L - s.size, complex)])
    want = oracle.synthesis_lmn2lm(g, s[:oracle.n_lm_scal(p)], p)
    got = real_lmn2lm(oracle, g, s, p, adjoint=False)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    want = oracle.analysis_adjoint_lmn2lm(g, s[:oracle.n_lm_scal(p)], p)
    got = real_lmn2lm(oracle, g, s, p, adjoint=True)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    # a plausible slip (the wrong sign on the conjugate term) is caught
    bad = real_lmn2lm(oracle, g, s, p, adjoint=True, flip=-1)
    assert np.abs(bad - want).max() > 1e-3 * np.abs(want).max()


@pytest.mark.parametrize("L,N,J_min,spin,up", STATES)
def test_hermitian_analysis_is_real(oracle, L, N, J_min, spin, up):
    p = oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(L * 7 + N)
    flm = hermitian_flm(rng, L)
    flm[:spin * spin] = 0
    lmn, _ = oracle.analysis_lm2lmn(flm, p)
    assert is_real_so3(oracle, lmn, p, 1e-14 * np.abs(lmn).max())
    fw, fs = oracle.analysis_lm2wav(flm, p)
    assert np.abs(fw.imag).max() < 1e-13 * np.abs(fw).max()
    assert np.abs(fs.imag).max() <= 1e-13 * np.abs(fs).max()  # f_scal may be all zero
    # and the complex adjoint of the synthesis too
    bw, bs = oracle.synthesis_adjoint_lm2wav(flm, p)
    assert np.abs(bw.imag).max() < 1e-13 * np.abs(bw).max()
    # a non-Hermitian flm is not: the property tests something
    fw2, _ = oracle.analysis_lm2wav(crandn(rng, L * L), p)
    assert np.abs(fw2.imag).max() > 1e-3 * np.abs(fw2).max()


def to_half(oracle, g, p):
    """n >= 0 rows of complex-layout Wigner coefficients, in the real-signal layout."""
    out, off_r = [], 0
    h = np.zeros(oracle.n_lmn_wav_real(p), complex)
    for off, Lj, Nj in scales(oracle, p):
        for n in range(0, Nj):
            for l in range(n, Lj):
                for m in range(-l, l + 1):
                    h[off_r + oracle.elmn2ind_real(l, m, n, Lj)] = g[off + oracle.elmn2ind(l, m, n, Lj, Nj)]
        off_r += oracle.flmn_size_real(Lj, Nj)
    assert off_r == h.size
    return h


@pytest.mark.parametrize("L,N", [(8, 1), (8, 4), (9, 5), (5, 5), (5, 6)])
def test_real_storage_index(oracle, L, N):
    """The n >= 0 compact order enumerates 0 .. size-1 once; size = sum_{n<N} (L^2 - n^2)."""
    seen = sorted(oracle.elmn2ind_real(l, m, n, L) for n in range(min(N, L)) for l in range(n, L)
                  for m in range(-l, l + 1))
    assert seen == list(range(oracle.flmn_size_real(L, N)))
    assert oracle.flmn_size_real(L, N) == sum(L * L - n * n for n in range(min(N, L)))


@pytest.mark.parametrize("L,N,J_min,spin,up", STATES)
def test_real_lmn_multiplies(oracle, L, N, J_min, spin, up):
    """The real multiplies (s2let_analysis.c:102-160 and siblings, in the oracle's C)
    against the complex ones: the to-lmn forms are the n >= 0 rows bit for bit; the
    to-lm forms equal the complex sums on coefficients of a real signal."""
    p = oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(5 * L + N)
    flm = hermitian_flm(rng, L)
    flm[:spin * spin] = 0
    a, b = oracle.analysis_lm2lmn_real(flm, p)
    c, d = oracle.analysis_lm2lmn(flm, p)
    assert np.array_equal(a, to_half(oracle, c, p)) and np.array_equal(b, d)
    a, b = oracle.synthesis_adjoint_lm2lmn_real(flm, p)
    c, d = oracle.synthesis_adjoint_lm2lmn(flm, p)
    assert np.array_equal(a, to_half(oracle, c, p)) and np.array_equal(b, d)
    g = symmetrise(oracle, crandn(rng, oracle.n_lmn_wav(p)), p)
    s = crandn(rng, oracle.n_lm_scal(p))
    for real_fn, cx_fn in ((oracle.synthesis_lmn2lm_real, oracle.synthesis_lmn2lm),
                           (oracle.analysis_adjoint_lmn2lm_real, oracle.analysis_adjoint_lmn2lm)):
        want = cx_fn(g, s, p)
        got = real_fn(to_half(oracle, g, p), s, p)
        assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
