"""s2let_hip_analysis_lm2wav_noalloc (the allocation-free analysis): called
through the C ABI, it must give bit for bit what s2let_hip_analysis_lm2wav
gives (the same kernels on the same plan, scratch in the library's static
device array), on two calls in a row, the second with a smaller input (the
array keeps the first call's contents); it must refuse, before launching
anything, parameter sets without a prepared plan and those whose scratch
exceeds the array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def powerlaw(rng, L, spin):
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    f = (rng.standard_normal(L * L) + 1j * rng.standard_normal(L * L)) / np.sqrt(2) / (ell + 1.0)
    f[:spin * spin] = 0
    return f


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(s2, torch, fn, flm, p, mode):
    fw = torch.zeros(s2.n_wav(p), dtype=torch.complex128, device="cuda")
    fs = torch.zeros(s2.n_scal(p), dtype=torch.complex128, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(_ptr(fw), _ptr(fs), _ptr(flm), ctypes.byref(p), int(mode), st)
    return rc, fw, fs


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_matches_analysis(mode):
    import torch

    import s2let_amd as s2
    lib = s2.lib()
    # first call: the c4 shape (L=256, N=4, spin 2); second: a smaller map (L=64, N=8)
    cases = [(256, 4, 2, 2), (64, 8, 1, 0)]
    rng = np.random.default_rng(7 + mode)
    for L, N, J_min, spin in cases:
        print("noalloc", mode, L, N, J_min, spin, flush=True)
        p = s2.parameters(L, N, 2.0, J_min, spin)
        assert lib.s2let_hip_prepare(ctypes.byref(p), mode) == 0
        flm = torch.from_numpy(powerlaw(rng, L, spin)).cuda()
        rc, fw, fs = _run(s2, torch, lib.s2let_hip_analysis_lm2wav_noalloc, flm, p, mode)
        assert rc == 0, s2.lib().s2let_hip_last_error_string()
        rc, rw, rs = _run(s2, torch, lib.s2let_hip_analysis_lm2wav, flm, p, mode)
        assert rc == 0
        torch.cuda.synchronize()
        assert torch.equal(fw, rw) and torch.equal(fs, rs), (L, N, mode)
        assert float(fw.abs().max()) > 0


def test_noalloc_refuses_without_plan_or_room():
    import torch

    import s2let_amd as s2
    lib = s2.lib()
    # no plan was prepared for this parameter set: refused, nothing is built
    p = s2.parameters(48, 3, 2.0, 1, 0)
    flm = torch.zeros(48 * 48, dtype=torch.complex128, device="cuda")
    rc, _, _ = _run(s2, torch, lib.s2let_hip_analysis_lm2wav_noalloc, flm, p, 0)
    assert rc == -4 and b"s2let_hip_prepare" in lib.s2let_hip_last_error_string()
    # L=512, N=8 in FP64 needs about 315 MB of scratch: refused before any launch
    p = s2.parameters(512, 8, 2.0, 2, 0)
    assert lib.s2let_hip_prepare(ctypes.byref(p), 0) == 0
    flm = torch.zeros(512 * 512, dtype=torch.complex128, device="cuda")
    rc, _, _ = _run(s2, torch, lib.s2let_hip_analysis_lm2wav_noalloc, flm, p, 0)
    assert rc == -5 and b"maximum" in lib.s2let_hip_last_error_string()
