"""GPU parity of the real-signal entry points (s2let_analysis.c:371-523,
s2let_synthesis.c:386-540, s2let_analysis_adjoint.c:402-556,
s2let_synthesis_adjoint.c:380-532, s2let_mw.c:51-70).

Against the oracle: the real transforms of a real signal equal the complex
oracle transforms on the same (promoted) input, with real parts kept
(tests/test_oracle_real.py pins that identity on the CPU).  Bar: FP64 within
1e-12 relative.  Also the reference's own real tests: the pixel-space adjoint
dot-product identities (test/c/test_adjoint.c:171-213, 323-366; tolerance
1e-11 as there) and the real round trip (test/c/test_back_and_forth.c:408-441,
1e-12), at its states L=16, N=4, B=2, J_min=2, spin 0 / 2, upsample 1 / 0.
"""
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


REF_STATES = [(16, 4, 2, 0, 1), (16, 4, 2, 2, 1), (16, 4, 2, 2, 0)]  # test/c/test_adjoint.c:476-480
# spin 0 runs on the half-spectrum plan (n >= 0 columns): N = 1 (the n = 0 column
# only), N = 2 (n = 1 only), odd N (n = 0, 2, ...), upsampled and multiresolution
CASES = REF_STATES + [(64, 4, 0, 0, 0), (33, 3, 1, 0, 0), (48, 8, 1, -2, 0), (32, 1, 0, 0, 0), (32, 2, 1, 0, 1),
                      (40, 5, 2, 0, 0)]


@pytest.mark.parametrize("L,N,J_min,spin,up", CASES)
def test_real_transforms_match_oracle(s2, oracle, L, N, J_min, spin, up):
    print("case", L, N, J_min, spin, up, flush=True)
    p, q = s2.parameters(L, N, 2.0, J_min, spin, up), oracle.params(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(3 * L + N + spin + up)
    flm = hermitian_flm(rng, L)
    fw, fs = s2.analysis_lm2wav_real(flm, p)
    ow, os_ = oracle.analysis_lm2wav(flm, q)
    assert rel(fw, ow.real) < TOL and rel(fs, os_.real) < TOL
    yw, ys = rng.standard_normal(s2.n_wav(p)), rng.standard_normal(s2.n_scal(p))
    assert rel(s2.synthesis_wav2lm_real(yw, ys, p), oracle.synthesis_wav2lm(yw, ys, q)) < TOL
    assert rel(s2.analysis_adjoint_wav2lm_real(yw, ys, p), oracle.analysis_adjoint_wav2lm(yw, ys, q)) < TOL
    bw, bs = s2.synthesis_adjoint_lm2wav_real(flm, p)
    cw, cs = oracle.synthesis_adjoint_lm2wav(flm, q)
    assert rel(bw, cw.real) < TOL and rel(bs, cs.real) < TOL
    # pixel space: spin-0 MW transforms of real maps
    f = rng.standard_normal(L * (2 * L - 1))
    aw, as_ = s2.analysis_px2wav_real(f, p)
    ow, os_ = oracle.analysis_lm2wav(oracle.ssht_forward(f, L, 0), q)
    assert rel(aw, ow.real) < TOL and rel(as_, os_.real) < TOL
    g = s2.synthesis_wav2px_real(yw, ys, p)
    assert rel(g, oracle.ssht_inverse(oracle.synthesis_wav2lm(yw, ys, q), L, 0).real) < TOL
    g = s2.analysis_adjoint_wav2px_real(yw, ys, p)
    assert rel(g, oracle.ssht_adjoint_forward(oracle.analysis_adjoint_wav2lm(yw, ys, q), L, 0).real) < TOL
    bw, bs = s2.synthesis_adjoint_px2wav_real(f, p)
    cw, cs = oracle.synthesis_adjoint_lm2wav(oracle.ssht_adjoint_inverse(f, L, 0), q)
    assert rel(bw, cw.real) < TOL and rel(bs, cs.real) < TOL


@pytest.mark.parametrize("L", [1, 2, 17, 64])
def test_mw_real_sht(s2, oracle, L):
    rng = np.random.default_rng(L)
    flm = hermitian_flm(rng, L)
    f = s2.mw_alm2map_real(flm, L)
    assert f.dtype == np.float64
    assert rel(f, oracle.ssht_inverse(flm, L, 0).real) < TOL
    g = rng.standard_normal(L * (2 * L - 1))
    assert rel(s2.mw_map2alm_real(g, L), oracle.ssht_forward(g, L, 0)) < TOL
    assert np.abs(s2.mw_map2alm_real(f, L) - flm).max() < 1e-12


@pytest.mark.parametrize("L,N,J_min,spin,up", REF_STATES)
def test_reference_real_adjoint_identities(s2, L, N, J_min, spin, up):
    """test/c/test_adjoint.c:171-213 and 323-366."""
    p = s2.parameters(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(1)
    f = s2.mw_alm2map_real(hermitian_flm(rng, L), L)
    f_rec = s2.mw_alm2map_real(hermitian_flm(rng, L), L)
    w_rec, s_rec = s2.analysis_px2wav_real(f_rec, p)
    w, s = s2.analysis_px2wav_real(f, p)
    f_rec = s2.analysis_adjoint_wav2px_real(w_rec, s_rec, p)
    err = np.dot(w_rec, w) + np.dot(s_rec, s) - np.dot(f_rec, f)
    assert abs(err) < 1e-11
    # synthesis: <S w, f> = <w, S^H f>
    w, s = s2.analysis_px2wav_real(s2.mw_alm2map_real(hermitian_flm(rng, L), L), p)
    g = s2.synthesis_wav2px_real(w, s, p)
    f2 = s2.mw_alm2map_real(hermitian_flm(rng, L), L)
    bw, bs = s2.synthesis_adjoint_px2wav_real(f2, p)
    err = np.dot(g, f2) - np.dot(w, bw) - np.dot(s, bs)
    assert abs(err) < 1e-11


@pytest.mark.parametrize("L,N,J_min,spin,up", [(16, 4, 2, 0, 1), (16, 4, 2, 0, 0), (16, 4, 2, 2, 1),
                                                 (64, 8, 2, 0, 0)])
def test_reference_real_round_trip(s2, L, N, J_min, spin, up):
    """test/c/test_back_and_forth.c:408-441 (run there at spin 0, :614-619):
    MW real map -> wavelets -> map -> flm.  Spin-s wavelets vanish below
    l = |s|, so at spin 2 only the l >= 2 part comes back."""
    p = s2.parameters(L, N, 2.0, J_min, spin, up)
    rng = np.random.default_rng(2)
    flm = hermitian_flm(rng, L)
    flm[:spin * spin] = 0
    f = s2.mw_alm2map_real(flm, L)
    w, s = s2.analysis_px2wav_real(f, p)
    f_rec = s2.synthesis_wav2px_real(w, s, p)
    flm_rec = s2.mw_map2alm_real(f_rec, L)
    assert np.abs(flm_rec - flm).max() < 1e-12
    # harmonic-space variant
    w, s = s2.analysis_lm2wav_real(flm, p)
    assert np.abs(s2.synthesis_wav2lm_real(w, s, p) - flm).max() < 1e-12


def _dev(p, s2, torch, fn, outs, ins, mode):
    args = [ctypes.c_void_p(t.data_ptr()) for t in outs + ins]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    s2.check(getattr(s2.lib(), fn)(*args, s2._ref(p), mode, stream))
    torch.cuda.synchronize()


def test_device_real_entry_points(s2):
    """Device-pointer real variants, FP64 (same as the host API) and mixed (1e-6)."""
    import torch
    L, N = 128, 4
    p = s2.parameters(L, N, 2.0, 2, 0, 0)
    rng = np.random.default_rng(4)
    flm = hermitian_flm(rng, L)
    hw, hs = s2.analysis_lm2wav_real(flm, p)
    dflm = torch.from_numpy(flm).cuda()
    nw, ns = s2.n_wav(p), s2.n_scal(p)
    for mode, tol in ((s2.FP64, 1e-14), (s2.MIXED, 1e-6)):
        fw = torch.zeros(nw, dtype=torch.float64, device="cuda")
        fs = torch.zeros(ns, dtype=torch.float64, device="cuda")
        _dev(p, s2, torch, "s2let_hip_analysis_lm2wav_real", [fw, fs], [dflm], mode)
        assert rel(fw.cpu().numpy(), hw) < tol and rel(fs.cpu().numpy(), hs) < tol
        out = torch.zeros(L * L, dtype=torch.complex128, device="cuda")
        _dev(p, s2, torch, "s2let_hip_synthesis_wav2lm_real", [out], [fw, fs], mode)
        assert np.abs(out.cpu().numpy() - flm).max() < (1e-12 if mode == s2.FP64 else 1e-5)
        want = s2.analysis_adjoint_wav2lm_real(hw, hs, p)
        _dev(p, s2, torch, "s2let_hip_analysis_adjoint_wav2lm_real", [out], [torch.from_numpy(hw).cuda(),
                                                                             torch.from_numpy(hs).cuda()], mode)
        assert rel(out.cpu().numpy(), want) < tol
        bw, bs = s2.synthesis_adjoint_lm2wav_real(flm, p)
        _dev(p, s2, torch, "s2let_hip_synthesis_adjoint_lm2wav_real", [fw, fs], [dflm], mode)
        assert rel(fw.cpu().numpy(), bw) < tol and rel(fs.cpu().numpy(), bs) < tol


@pytest.mark.parametrize("L,N,J_min,up", [(24, 4, 1, 0), (17, 3, 0, 1)])
def test_real_mwss_match_oracle(s2, oracle, L, N, J_min, up):
    """MWSS sampling on the half-spectrum plan: all four real transforms against the
    complex oracle on the promoted input (real parts kept)."""
    print("mwss real", L, N, J_min, up, flush=True)
    p = s2.parameters(L, N, 2.0, J_min, 0, up, sampling_scheme=1)
    q = oracle.params(L, N, 2.0, J_min, 0, up, sampling=1)
    rng = np.random.default_rng(L + N)
    flm = hermitian_flm(rng, L)
    fw, fs = s2.analysis_lm2wav_real(flm, p)
    ow, os_ = oracle.analysis_lm2wav(flm, q)
    assert rel(fw, ow.real) < TOL and rel(fs, os_.real) < TOL
    yw, ys = rng.standard_normal(s2.n_wav(p)), rng.standard_normal(s2.n_scal(p))
    assert rel(s2.synthesis_wav2lm_real(yw, ys, p), oracle.synthesis_wav2lm(yw, ys, q)) < TOL
    assert rel(s2.analysis_adjoint_wav2lm_real(yw, ys, p), oracle.analysis_adjoint_wav2lm(yw, ys, q)) < TOL
    bw, bs = s2.synthesis_adjoint_lm2wav_real(flm, p)
    cw, cs = oracle.synthesis_adjoint_lm2wav(flm, q)
    assert rel(bw, cw.real) < TOL and rel(bs, cs.real) < TOL


def test_half_spectrum_matches_complex_path_c2_shape(s2):
    """L = 512, N = 8, B = 2, J_min = 2, spin 0, a real signal: the half-spectrum
    device transforms against the complex device transforms on the promoted input,
    FP64 within 1e-12 and mixed within 1e-6 (of the FP64 complex result); synthesis
    of the analysis reproduces flm (FP64 1e-12)."""
    import torch
    from s2let_amd import torch_ops as ops
    L, N = 512, 8
    p = s2.parameters(L, N, 2.0, 2, 0)
    rng = np.random.default_rng(9)
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    flm = hermitian_flm(rng, L) / (ell + 1.0)
    dflm = torch.from_numpy(flm).cuda()
    nw, ns = s2.n_wav(p), s2.n_scal(p)
    cw, cs = ops.analysis(dflm, p, ops.FP64)
    c_rec = ops.synthesis(cw, cs, p, ops.FP64)
    ref_a = torch.cat([cw, cs]).real.cpu().numpy()
    for mode, tol in ((s2.FP64, 1e-12), (s2.MIXED, 1e-6)):
        print("half spectrum c2 shape, mode", mode, flush=True)
        fw = torch.zeros(nw, dtype=torch.float64, device="cuda")
        fs = torch.zeros(ns, dtype=torch.float64, device="cuda")
        _dev(p, s2, torch, "s2let_hip_analysis_lm2wav_real", [fw, fs], [dflm], mode)
        assert rel(torch.cat([fw, fs]).cpu().numpy(), ref_a) < tol
        out = torch.zeros(L * L, dtype=torch.complex128, device="cuda")
        _dev(p, s2, torch, "s2let_hip_synthesis_wav2lm_real", [out], [fw, fs], mode)
        assert rel(out.cpu().numpy(), c_rec.cpu().numpy()) < tol
        if mode == s2.FP64:
            assert rel(out.cpu().numpy(), flm) < 1e-12


def test_host_api_staged_transfers_match_device(s2):
    """Host-pointer entry points move arrays of 8 MB and more in 32 MB chunks through
    pinned staging buffers (L = 256, N = 8: f_wav of 73 MB, several chunks and a partial
    one): their results equal the device-pointer path's bit for bit, for the complex and
    the real transforms, twice in a row (the staging buffers reused)."""
    import torch
    from s2let_amd import torch_ops as ops
    L, N = 256, 8
    p = s2.parameters(L, N, 2.0, 2, 0)
    rng = np.random.default_rng(12)
    flm = hermitian_flm(rng, L)
    assert s2.n_wav(p) * 16 > 2 * (32 << 20)
    dflm = torch.from_numpy(flm).cuda()
    for rep in range(2):
        print("host staged transfers", rep, flush=True)
        hw, hs = s2.analysis_lm2wav(flm, p)
        dw, ds = ops.analysis(dflm, p, s2.FP64)
        assert np.array_equal(hw, dw.cpu().numpy()) and np.array_equal(hs, ds.cpu().numpy())
        hr = s2.synthesis_wav2lm(hw, hs, p)
        dr = ops.synthesis(dw, ds, p, s2.FP64)
        assert np.array_equal(hr, dr.cpu().numpy())
        rw, rs = s2.analysis_lm2wav_real(flm, p)
        fw, fs = ops.analysis_real(dflm, p, s2.FP64)
        assert np.array_equal(rw, fw.cpu().numpy()) and np.array_equal(rs, fs.cpu().numpy())
        rr = s2.synthesis_wav2lm_real(rw, rs, p)
        fr = ops.synthesis_real(fw, fs, p, s2.FP64)
        assert np.array_equal(rr, fr.cpu().numpy())
        flm = flm * (-0.5)  # new values through the same buffers
        dflm = torch.from_numpy(flm).cuda()
