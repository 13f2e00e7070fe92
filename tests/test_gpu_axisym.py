"""GPU parity of the axisymmetric wavelet API (s2let_transform_axisym_lm.c,
s2let_transform_axisym_mw.c) against the oracle's restatement
(tests/test_oracle_axisym.py pins it): harmonic multiplies bit for bit, pixel
transforms within 1e-12 (FP64), real variants, the reference's axisymmetric round
trips (test/c/test_back_and_forth.c:178-405, 1e-12) and the hard threshold."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def crandn(rng, n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)


def hermitian_flm(rng, L):
    f = crandn(rng, L * L)
    for l in range(L):
        f[l * l + l] = f[l * l + l].real
        for m in range(1, l + 1):
            f[l * l + l - m] = (-1) ** m * np.conj(f[l * l + l + m])
    return f


@pytest.fixture(scope="module")
def s2():
    import s2let_amd
    s2let_amd.lib()
    return s2let_amd


STATES = [(16, 2.0, 2), (33, 3.0, 1), (64, 2.0, 0)]


@pytest.mark.parametrize("L,B,J_min", STATES)
@pytest.mark.parametrize("multires", [False, True])
def test_axisym_lm_multiplies_bit_exact(s2, oracle, L, B, J_min, multires):
    p, q = s2.parameters(L, 1, B, J_min), oracle.params(L, 1, B, J_min)
    rng = np.random.default_rng(L)
    flm = crandn(rng, L * L)
    wav, scal = s2.axisym_lm_wav(p)
    a, b = s2.axisym_lm_wav_analysis(flm, wav, scal, p, multires)
    c, d = oracle.axisym_lm_wav_analysis(flm, wav, scal, q, int(multires))
    assert np.array_equal(a, c) and np.array_equal(b, d)
    base = crandn(rng, L * L)
    got = s2.axisym_lm_wav_synthesis(a, b, wav, scal, p, multires, flm=base)
    want = oracle.axisym_lm_wav_synthesis(c, d, wav, scal, q, int(multires), flm=base)
    assert np.array_equal(got, want)
    assert np.abs(s2.axisym_lm_wav_synthesis(a, b, wav, scal, p, multires) - flm).max() < 1e-12


@pytest.mark.parametrize("L,B,J_min", STATES)
@pytest.mark.parametrize("multires", [False, True])
def test_axisym_mw_match_oracle(s2, oracle, L, B, J_min, multires):
    print("case", L, B, J_min, multires, flush=True)
    p, q = s2.parameters(L, 1, B, J_min), oracle.params(L, 1, B, J_min)
    m = int(multires)
    rng = np.random.default_rng(L + 7 * m)
    f = crandn(rng, L * (2 * L - 1))
    fw, fs = s2.axisym_wav_analysis_mw(f, p, multires)
    ow, os_ = oracle.axisym_mw(0, m, f, None, q)
    assert rel(fw, ow) < TOL and rel(fs, os_) < TOL
    yw, ys = crandn(rng, fw.size), crandn(rng, fs.size)
    assert rel(s2.axisym_wav_analysis_adjoint_mw(yw, ys, p, multires), oracle.axisym_mw(1, m, yw, ys, q)) < TOL
    assert rel(s2.axisym_wav_synthesis_mw(yw, ys, p, multires), oracle.axisym_mw(2, m, yw, ys, q)) < TOL
    bw, bs = s2.axisym_wav_synthesis_adjoint_mw(f, p, multires)
    cw, cs = oracle.axisym_mw(3, m, f, None, q)
    assert rel(bw, cw) < TOL and rel(bs, cs) < TOL
    # real variants: real maps, real wavelet coefficients
    fr = rng.standard_normal(L * (2 * L - 1))
    rw, rs = s2.axisym_wav_analysis_mw(fr, p, multires, real=True)
    ow, os_ = oracle.axisym_mw(0, m, fr, None, q)
    assert rw.dtype == np.float64 and rel(rw, ow.real) < TOL and rel(rs, os_.real) < TOL
    yr, zr = rng.standard_normal(fw.size), rng.standard_normal(fs.size)
    assert rel(s2.axisym_wav_synthesis_mw(yr, zr, p, multires, real=True), oracle.axisym_mw(2, m, yr, zr, q).real) < TOL
    assert rel(s2.axisym_wav_analysis_adjoint_mw(yr, zr, p, multires, real=True),
               oracle.axisym_mw(1, m, yr, zr, q).real) < TOL
    bw, bs = s2.axisym_wav_synthesis_adjoint_mw(fr, p, multires, real=True)
    cw, cs = oracle.axisym_mw(3, m, fr, None, q)
    assert rel(bw, cw.real) < TOL and rel(bs, cs.real) < TOL


@pytest.mark.parametrize("multires", [False, True])
@pytest.mark.parametrize("real", [False, True])
def test_reference_axisym_round_trips(s2, multires, real):
    """test/c/test_back_and_forth.c:178-405 (L=16, B=2, J_min=2 there)."""
    L = 16
    p = s2.parameters(L, 1, 2.0, 2)
    rng = np.random.default_rng(11)
    flm = hermitian_flm(rng, L) if real else crandn(rng, L * L)
    f = s2.mw_alm2map_real(flm, L) if real else s2.mw_alm2map(flm, L, 0)
    fw, fs = s2.axisym_wav_analysis_mw(f, p, multires, real)
    g = s2.axisym_wav_synthesis_mw(fw, fs, p, multires, real)
    back = s2.mw_map2alm_real(g, L) if real else s2.mw_map2alm(g, L, 0)
    assert np.abs(back - flm).max() < 1e-12


@pytest.mark.parametrize("multires", [False, True])
def test_axisym_hardthreshold(s2, multires):
    """s2let_transform_axisym_mw.c:1183-1230: per-scale thresholds on |trunc(x)|."""
    L = 16
    p = s2.parameters(L, 1, 2.0, 2)
    bl = [min(s2.bandlimit(j, p), L) if multires else L for j in range(2, s2.j_max(p) + 1)]
    n = [b * (2 * b - 1) for b in bl]
    rng = np.random.default_rng(5)
    g = rng.standard_normal(sum(n)) * 3
    thr = np.linspace(0.5, 2.5, len(n))
    got = s2.axisym_wav_hardthreshold_real(g, thr, p, multires)
    want, off = g.copy(), 0
    for k, c in enumerate(n):
        seg = want[off:off + c]
        seg[np.abs(np.trunc(seg)) < thr[k]] = 0
        off += c
    assert np.array_equal(got, want)
    assert np.any((np.abs(g) > thr[0]) & (got == 0))  # e.g. 0.7 under a 0.5 threshold is zeroed, as there
