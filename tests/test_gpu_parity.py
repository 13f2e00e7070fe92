"""GPU parity of the C ABI entry points against the CPU FP64 oracle.

Bar: FP64 mode within 1e-12 relative (L2 over the whole output, and max-abs
relative to the largest entry) of the oracle on the same inputs; exact
reconstruction synth(anal(f)) = f to FP64 round-off; adjoint dot-product
identities; mixed mode within 1e-6 relative of FP64.  Cases cover the
reference's own test states (test/c/test_back_and_forth.c:577-611: L=16, N=4,
B=2, J_min=2, spin 0/2, upsample 0/1), odd N, N=1, and the workload sizes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def maxrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max() if b.size else 1.0
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


def crandn(rng, n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)


def flm_powerlaw(rng, L, spin=0):
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    f = crandn(rng, L * L) / (ell + 1.0)
    f[:spin * spin] = 0
    return f


@pytest.fixture(scope="module")
def s2():
    import s2let_amd
    s2let_amd.lib()
    return s2let_amd


def test_mfma_operand_layout(s2):
    # the matrix-core lane layouts the contraction kernels assume (FP64 and FP32 16x16x4)
    assert s2.lib().s2let_hip_selftest() == 0


CASES = [(16, 4, 2, s, u) for s in (0, 2) for u in (0, 1)] + [
    (64, 4, 0, 0, 0),    # L=64, N=4, J_min=0 (zero tail in f_wav)
    (32, 3, 1, 1, 0),    # odd N, spin 1
    (32, 1, 2, 0, 0),    # axisymmetric directional
    (24, 5, 1, -2, 0),   # negative spin, odd N, non power-of-two L
    (64, 8, 1, 0, 0),    # N_j > L_j at the small scales (parity-adjusted N_j)
]


@pytest.mark.parametrize("L,N,J_min,spin,up", CASES)
def test_transforms_match_oracle(s2, oracle, L, N, J_min, spin, up):
    print("case", L, N, J_min, spin, up, flush=True)
    p, q = s2.parameters(L, N, 2.0, J_min, spin, up), oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(L * 100 + N * 10 + J_min + 7 * spin + up)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    fw, fs = s2.analysis_lm2wav(flm, p)
    ow, os_ = oracle.analysis_lm2wav(flm, q)
    assert rel(fw, ow) < TOL and maxrel(fw, ow) < TOL and rel(fs, os_) < TOL
    yw, ys = crandn(rng, s2.n_wav(p)), crandn(rng, s2.n_scal(p))
    assert rel(s2.synthesis_wav2lm(yw, ys, p), oracle.synthesis_wav2lm(yw, ys, q)) < TOL
    assert rel(s2.analysis_adjoint_wav2lm(yw, ys, p), oracle.analysis_adjoint_wav2lm(yw, ys, q)) < TOL
    bw, bs = s2.synthesis_adjoint_lm2wav(flm, p)
    cw, cs = oracle.synthesis_adjoint_lm2wav(flm, q)
    assert rel(bw, cw) < TOL and rel(bs, cs) < TOL
    f = crandn(rng, L * (2 * L - 1))
    aw, as_ = s2.analysis_px2wav(f, p)
    ow, os_ = oracle.analysis_px2wav(f, q)
    assert rel(aw, ow) < TOL and rel(as_, os_) < TOL
    assert rel(s2.synthesis_wav2px(yw, ys, p), oracle.synthesis_wav2px(yw, ys, q)) < TOL
    assert rel(s2.analysis_adjoint_wav2px(yw, ys, p), oracle.analysis_adjoint_wav2px(yw, ys, q)) < TOL
    bw, bs = s2.synthesis_adjoint_px2wav(f, p)
    cw, cs = oracle.synthesis_adjoint_px2wav(f, q)
    assert rel(bw, cw) < TOL and rel(bs, cs) < TOL
    # reconstruction
    fw, fs = s2.analysis_lm2wav(flm, p)
    assert np.abs(s2.synthesis_wav2lm(fw, fs, p) - flm).max() < 1e-12


@pytest.mark.parametrize("L,N,J_min,spin,up", CASES[:5])
def test_lmn_api_matches_oracle(s2, oracle, L, N, J_min, spin, up):
    p, q = s2.parameters(L, N, 2.0, J_min, spin, up), oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(5)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    a, b = s2.analysis_lm2lmn(flm, p)
    c, d = oracle.analysis_lm2lmn(flm, q)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    a, b = s2.synthesis_adjoint_lm2lmn(flm, p)
    c, d = oracle.synthesis_adjoint_lm2lmn(flm, q)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    wl, sl = crandn(rng, a.size), crandn(rng, b.size)
    assert maxrel(s2.synthesis_lmn2lm(wl, sl, p), oracle.synthesis_lmn2lm(wl, sl, q)) < 1e-15
    assert maxrel(s2.analysis_adjoint_lmn2lm(wl, sl, p), oracle.analysis_adjoint_lmn2lm(wl, sl, q)) < 1e-15


@pytest.mark.parametrize("L,N,J_min,spin,up", CASES)
def test_real_lmn_api_matches_oracle(s2, oracle, L, N, J_min, spin, up):
    """Real-signal multiplies (n >= 0 storage), bit for bit."""
    p, q = s2.parameters(L, N, 2.0, J_min, spin, up), oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(6)
    flm = crandn(rng, L * L)
    a, b = s2.analysis_lm2lmn_real(flm, p)
    c, d = oracle.analysis_lm2lmn_real(flm, q)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    a, b = s2.synthesis_adjoint_lm2lmn_real(flm, p)
    c, d = oracle.synthesis_adjoint_lm2lmn_real(flm, q)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    wl, sl = crandn(rng, a.size), crandn(rng, b.size)
    assert np.array_equal(s2.synthesis_lmn2lm_real(wl, sl, p), oracle.synthesis_lmn2lm_real(wl, sl, q))
    assert np.array_equal(s2.analysis_adjoint_lmn2lm_real(wl, sl, p), oracle.analysis_adjoint_lmn2lm_real(wl, sl, q))


@pytest.mark.parametrize("L,spin", [(1, 0), (2, 0), (9, 1), (33, -2), (128, 2)])
def test_mw_sht_matches_oracle(s2, oracle, L, spin):
    rng = np.random.default_rng(L + spin)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    f = s2.mw_alm2map(flm, L, spin)
    assert rel(f, oracle.ssht_inverse(flm, L, spin)) < TOL
    g = crandn(rng, L * (2 * L - 1))
    assert rel(s2.mw_map2alm(g, L, spin), oracle.ssht_forward(g, L, spin)) < TOL
    assert np.abs(s2.mw_map2alm(f, L, spin) - flm).max() < 1e-12


def _dot_identities(s2, p, rng):
    L = p.L
    x = crandn(rng, L * L)
    x[:p.spin * p.spin] = 0
    yw, ys = crandn(rng, s2.n_wav(p)), crandn(rng, s2.n_scal(p))
    aw, as_ = s2.analysis_lm2wav(x, p)
    lhs = np.vdot(yw, aw) + np.vdot(ys, as_)
    rhs = np.vdot(s2.analysis_adjoint_wav2lm(yw, ys, p), x)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    y = crandn(rng, L * L)
    lhs = np.vdot(y, s2.synthesis_wav2lm(yw, ys, p))
    bw, bs = s2.synthesis_adjoint_lm2wav(y, p)
    assert abs(lhs - (np.vdot(bw, yw) + np.vdot(bs, ys))) < 1e-11 * abs(lhs)


def test_adjoint_identities_c4_shape(s2):
    # L=256, N=4, B=2, J_min=2, spin 2 (differentiable-transform workload shape)
    _dot_identities(s2, s2.parameters(256, 4, 2.0, 2, 2, 0), np.random.default_rng(4))


@pytest.mark.parametrize("L,N,J_min,spin", [(512, 8, 2, 0), (256, 4, 2, 2), (128, 8, 1, 2)])
def test_reconstruction_full_size(s2, L, N, J_min, spin):
    print("reconstruction", L, N, J_min, spin, flush=True)
    p = s2.parameters(L, N, 2.0, J_min, spin, 0)
    flm = flm_powerlaw(np.random.default_rng(2), L, spin)
    fw, fs = s2.analysis_lm2wav(flm, p)
    rec = s2.synthesis_wav2lm(fw, fs, p)
    assert rel(rec, flm) < 1e-12
    assert np.abs(rec - flm).max() < 1e-12


def test_c2_analysis_matches_oracle(s2, oracle):
    # L=512, N=8, B=2, J_min=2, spin 0, multiresolution; C_l = (l+1)^-2, seed 2
    print("c2 analysis parity", flush=True)
    L, N, J_min = 512, 8, 2
    p, q = s2.parameters(L, N, 2.0, J_min), oracle.params(L, N, 2.0, J_min)
    flm = flm_powerlaw(np.random.default_rng(2), L)
    fw, fs = s2.analysis_lm2wav(flm, p)
    ow, os_ = oracle.analysis_lm2wav(flm, q)
    assert rel(fw, ow) < TOL and maxrel(fw, ow) < TOL and rel(fs, os_) < TOL


def test_f_wav_zero_tail(s2):
    # (2N-1) gamma slices are allocated per scale, (2N_j-1) are written
    p = s2.parameters(64, 4, 2.0, 0)
    fw, fs = s2.analysis_lm2wav(crandn(np.random.default_rng(9), 64 * 64), p)
    assert fw.size == 132454 and np.all(fw[132430:] == 0)
