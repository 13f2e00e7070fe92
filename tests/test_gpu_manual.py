"""GPU parity of the manual-tiling entry points (s2let_analysis.c:188-276,
s2let_synthesis.c:202-289, s2let_analysis_adjoint.c:202-290,
s2let_synthesis_adjoint.c:189-290) against the oracle's restatement
(tests/test_oracle_manual.py pins it), FP64 within 1e-12 relative.  Cases: the
reference's own (the standard tiling passed in as a manual one,
test/c/test_back_and_forth.c:97-175, round trip 1e-12) and random complex
wavelets with uneven band-limits (each direction then runs as two weight
columns)."""
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


@pytest.fixture(scope="module")
def s2():
    import s2let_amd
    s2let_amd.lib()
    return s2let_amd


def both(s2, oracle, wav, phi, scal_bl, wav_bl, L, spin, N):
    return (s2.ManualTiling(wav, phi, scal_bl, wav_bl, L, spin, N),
            oracle.Manual(wav, phi, scal_bl, wav_bl, L, spin, N))


def standard(s2, oracle, L, N, J_min, spin, up):
    p = s2.parameters(L, N, 2.0, J_min, spin, up)
    J = s2.j_max(p) - J_min
    psi, phi = s2.tiling_wavelet(p)
    scal_bl = L if up else min(s2.bandlimit(J_min - 1, p), L)
    wav_bl = [L if up else min(s2.bandlimit(J_min + j, p), L) for j in range(J + 1)]
    wav = np.asarray(psi).reshape(-1, L * L)[J_min:J_min + J + 1].ravel()
    return both(s2, oracle, wav, phi, scal_bl, wav_bl, L, spin, N)


def check_all(s2, oracle, t, m, rng, spin):
    L = t.L
    flm = crandn(rng, L * L)
    flm[:spin * spin] = 0
    fw, fs = s2.analysis_lm2wav_manual(flm, t)
    ow, os_ = oracle.analysis_lm2wav_manual(flm, m)
    assert rel(fw, ow) < TOL and rel(fs, os_) < TOL
    yw, ys = crandn(rng, fw.size), crandn(rng, fs.size)
    base = crandn(rng, L * L)
    assert rel(s2.synthesis_wav2lm_manual(yw, ys, t, flm=base), oracle.synthesis_wav2lm_manual(yw, ys, m, flm=base)) < TOL
    assert rel(s2.analysis_adjoint_wav2lm_manual(yw, ys, t, flm=base),
               oracle.analysis_adjoint_wav2lm_manual(yw, ys, m, flm=base)) < TOL
    assert rel(s2.analysis_adjoint_wav2lm_manual(yw, ys, t), oracle.analysis_adjoint_wav2lm_manual(yw, ys, m)) < TOL
    bw, bs = s2.synthesis_adjoint_lm2wav_manual(flm, t)
    cw, cs = oracle.synthesis_adjoint_lm2wav_manual(flm, m)
    assert rel(bw, cw) < TOL and rel(bs, cs) < TOL
    return flm, fw, fs


@pytest.mark.parametrize("L,N,J_min,spin,up", [(16, 4, 2, 0, 1), (16, 4, 2, 0, 0), (16, 4, 2, 2, 1),
                                                 (16, 4, 2, 2, 0), (64, 8, 2, 0, 0), (33, 3, 1, 1, 0)])
def test_standard_tiling_as_manual(s2, oracle, L, N, J_min, spin, up):
    print("case", L, N, J_min, spin, up, flush=True)
    t, m = standard(s2, oracle, L, N, J_min, spin, up)
    rng = np.random.default_rng(L + N + spin)
    flm, fw, fs = check_all(s2, oracle, t, m, rng, spin)
    # test/c/test_back_and_forth.c:97-175: exact reconstruction
    assert np.abs(s2.synthesis_wav2lm_manual(fw, fs, t) - flm).max() < 1e-12


@pytest.mark.parametrize("L,N,spin,wav_bl,scal_bl", [(24, 3, 0, [8, 16, 24], 8), (32, 4, 2, [12, 20, 32, 40], 12),
                                                    (20, 5, -1, [20], 30), (16, 16, 0, [16, 9], 5)])
def test_complex_manual_wavelets(s2, oracle, L, N, spin, wav_bl, scal_bl):
    print("case", L, N, spin, wav_bl, scal_bl, flush=True)
    rng = np.random.default_rng(L * N + 1)
    J = len(wav_bl) - 1
    wav = crandn(rng, (J + 1) * L * L)
    t, m = both(s2, oracle, wav, rng.standard_normal(L), scal_bl, wav_bl, L, spin, N)
    check_all(s2, oracle, t, m, rng, spin)
