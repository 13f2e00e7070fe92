"""GPU parity of the transforms on the MWSS grid (S2LET_SAMPLING_MW_SS,
s2let_types.h:26-32: theta_t = pi t / L, t = 0..L; phi_p = pi p / L, p < 2L)
against the oracle (tests/test_oracle_mwss.py pins it), FP64 within 1e-12
relative; the reference's MWSS round trips (test/c/test_back_and_forth.c:445-526,
1e-12: the complex one, skipped there, and the real one); adjoint identities;
mixed precision within 1e-6 of FP64 through the device API."""
import ctypes

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


CASES = [(16, 4, 2, s, u) for s in (0, 2) for u in (0, 1)] + [
    (64, 4, 0, 0, 0), (33, 3, 1, 1, 0), (64, 8, 1, 0, 0), (2, 1, 0, 0, 1)]


@pytest.mark.parametrize("L,N,J_min,spin,up", CASES)
def test_mwss_transforms_match_oracle(s2, oracle, L, N, J_min, spin, up):
    print("case", L, N, J_min, spin, up, flush=True)
    p = s2.parameters(L, N, 2.0, J_min, spin, up, sampling_scheme=1)
    q = oracle.params(L, N, 2.0, J_min, spin, up, sampling=1)
    rng = np.random.default_rng(L * 10 + N + spin + up)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    fw, fs = s2.analysis_lm2wav(flm, p)
    ow, os_ = oracle.analysis_lm2wav(flm, q)
    assert fw.size == ow.size and fs.size == os_.size == s2.n_scal(p)
    assert rel(fw, ow) < TOL and rel(fs, os_) < TOL
    yw, ys = crandn(rng, s2.n_wav(p)), crandn(rng, s2.n_scal(p))
    assert rel(s2.synthesis_wav2lm(yw, ys, p), oracle.synthesis_wav2lm(yw, ys, q)) < TOL
    assert rel(s2.analysis_adjoint_wav2lm(yw, ys, p), oracle.analysis_adjoint_wav2lm(yw, ys, q)) < TOL
    bw, bs = s2.synthesis_adjoint_lm2wav(flm, p)
    cw, cs = oracle.synthesis_adjoint_lm2wav(flm, q)
    assert rel(bw, cw) < TOL and rel(bs, cs) < TOL
    f = crandn(rng, (L + 1) * 2 * L)
    aw, as_ = s2.analysis_px2wav(f, p)
    ow, os_ = oracle.analysis_px2wav(f, q)
    assert rel(aw, ow) < TOL and rel(as_, os_) < TOL
    assert rel(s2.synthesis_wav2px(yw, ys, p), oracle.synthesis_wav2px(yw, ys, q)) < TOL
    assert rel(s2.analysis_adjoint_wav2px(yw, ys, p), oracle.analysis_adjoint_wav2px(yw, ys, q)) < TOL
    bw, bs = s2.synthesis_adjoint_px2wav(f, p)
    cw, cs = oracle.synthesis_adjoint_px2wav(f, q)
    assert rel(bw, cw) < TOL and rel(bs, cs) < TOL
    # exact reconstruction in harmonic space
    fw, fs = s2.analysis_lm2wav(flm, p)
    assert np.abs(s2.synthesis_wav2lm(fw, fs, p) - flm).max() < 1e-12


@pytest.mark.parametrize("L,N,J_min,spin,up", CASES[:4])
def test_reference_mwss_round_trips(s2, oracle, L, N, J_min, spin, up):
    """test/c/test_back_and_forth.c:445-526: MWSS map -> wavelets -> map -> flm."""
    p = s2.parameters(L, N, 2.0, J_min, spin, up, sampling_scheme=1)
    rng = np.random.default_rng(7)
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    f = oracle.ssht_inverse(flm, L, spin, ss=1)
    fw, fs = s2.analysis_px2wav(f, p)
    g = s2.synthesis_wav2px(fw, fs, p)
    assert np.abs(oracle.ssht_forward(g, L, spin, ss=1) - flm).max() < 1e-12
    if spin == 0:  # the real variant runs at spin 0 there (:614-619)
        hlm = hermitian_flm(rng, L)
        fr = oracle.ssht_inverse(hlm, L, 0, ss=1).real
        w, s = s2.analysis_px2wav_real(fr, p)
        gr = s2.synthesis_wav2px_real(w, s, p)
        assert gr.dtype == np.float64 and gr.size == (L + 1) * 2 * L
        assert np.abs(oracle.ssht_forward(gr, L, 0, ss=1) - hlm).max() < 1e-12


@pytest.mark.parametrize("L,N,J_min,spin,up", [(16, 4, 2, 0, 1), (16, 4, 2, 2, 0), (48, 8, 2, 0, 0)])
def test_mwss_adjoint_identities(s2, L, N, J_min, spin, up):
    p = s2.parameters(L, N, 2.0, J_min, spin, up, sampling_scheme=1)
    rng = np.random.default_rng(L + N)
    f = crandn(rng, s2.n_px(p))
    yw, ys = crandn(rng, s2.n_wav(p)), crandn(rng, s2.n_scal(p))
    aw, as_ = s2.analysis_px2wav(f, p)
    lhs = np.vdot(yw, aw) + np.vdot(ys, as_)
    rhs = np.vdot(s2.analysis_adjoint_wav2px(yw, ys, p), f)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    g = s2.synthesis_wav2px(yw, ys, p)
    bw, bs = s2.synthesis_adjoint_px2wav(f, p)
    lhs = np.vdot(f, g)
    rhs = np.vdot(bw, yw) + np.vdot(bs, ys)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_mwss_device_mixed(s2):
    """Device API on the MWSS grid: FP64 equals the host API, mixed within 1e-6."""
    import torch
    L, N = 128, 4
    p = s2.parameters(L, N, 2.0, 2, 0, 0, sampling_scheme=1)
    rng = np.random.default_rng(9)
    flm = crandn(rng, L * L)
    hw, hs = s2.analysis_lm2wav(flm, p)
    dflm = torch.from_numpy(flm).cuda()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for mode, tol in ((s2.FP64, 1e-14), (s2.MIXED, 1e-6)):
        fw = torch.zeros(s2.n_wav(p), dtype=torch.complex128, device="cuda")
        fs = torch.zeros(s2.n_scal(p), dtype=torch.complex128, device="cuda")
        s2.check(s2.lib().s2let_hip_analysis_lm2wav(ctypes.c_void_p(fw.data_ptr()), ctypes.c_void_p(fs.data_ptr()),
                                                    ctypes.c_void_p(dflm.data_ptr()), s2._ref(p), mode, stream))
        torch.cuda.synchronize()
        assert rel(fw.cpu().numpy(), hw) < tol and rel(fs.cpu().numpy(), hs) < tol
        out = torch.zeros(L * L, dtype=torch.complex128, device="cuda")
        s2.check(s2.lib().s2let_hip_synthesis_wav2lm(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(fw.data_ptr()),
                                                     ctypes.c_void_p(fs.data_ptr()), s2._ref(p), mode, stream))
        torch.cuda.synchronize()
        assert np.abs(out.cpu().numpy() - flm).max() < (1e-12 if mode == s2.FP64 else 1e-5)
