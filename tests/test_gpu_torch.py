"""Device-pointer entry points through torch tensors: FP64 and mixed
precision against the oracle, and autograd through the adjoint kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crandn(rng, n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def ops():
    import torch  # noqa: F401
    from s2let_amd import torch_ops
    return torch_ops


@pytest.mark.parametrize("L,N,J_min,spin", [(64, 4, 0, 0), (128, 8, 2, 2), (512, 8, 2, 0)])
def test_device_api_fp64_and_mixed(ops, oracle, L, N, J_min, spin):
    import torch
    print("device api", L, N, J_min, spin, flush=True)
    p = ops.parameters(L, N, 2.0, J_min, spin)
    rng = np.random.default_rng(L)
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    flm = crandn(rng, L * L) / (ell + 1.0)
    flm[:spin * spin] = 0
    x = torch.from_numpy(flm).cuda()
    w64, s64 = ops.analysis(x, p, ops.FP64)
    w32, s32 = ops.analysis(x, p, ops.MIXED)
    torch.cuda.synchronize()
    a64 = np.concatenate([w64.cpu().numpy(), s64.cpu().numpy()])
    a32 = np.concatenate([w32.cpu().numpy(), s32.cpu().numpy()])
    if L <= 128:
        ow, os_ = oracle.analysis_lm2wav(flm, oracle.params(L, N, 2.0, J_min, spin))
        assert rel(a64, np.concatenate([ow, os_])) < 1e-12
    assert rel(a32, a64) < 1e-6
    r64 = ops.synthesis(w64, s64, p, ops.FP64)
    r32 = ops.synthesis(w32, s32, p, ops.MIXED)
    torch.cuda.synchronize()
    assert rel(r64.cpu().numpy(), flm) < 1e-12
    assert rel(r32.cpu().numpy(), r64.cpu().numpy()) < 1e-6


def test_autograd_backward_is_adjoint(ops, oracle):
    import torch
    L, N = 32, 4
    p = ops.parameters(L, N, 2.0, 2, 2)
    rng = np.random.default_rng(3)
    flm = crandn(rng, L * L)
    flm[:4] = 0
    x = torch.from_numpy(flm).cuda().requires_grad_(True)
    tw = torch.from_numpy(crandn(rng, ops.n_wav(p))).cuda()
    ts = torch.from_numpy(crandn(rng, ops.n_scal(p))).cuda()
    w, s = ops.wavelet_analysis(x, p)
    loss = (w - tw).abs().pow(2).sum() + (s - ts).abs().pow(2).sum()
    loss.backward()
    g = x.grad.detach().cpu().numpy()
    # the reference adjoint routine (oracle restatement of s2let_analysis_adjoint_wav2lm)
    gw = 2 * (w.detach() - tw).cpu().numpy()
    gs = 2 * (s.detach() - ts).cpu().numpy()
    ref = oracle.analysis_adjoint_wav2lm(gw, gs, oracle.params(L, N, 2.0, 2, 2))
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-12
    l0 = float(loss.item())
    # central finite differences (FP64, step 1e-6) on a few coefficients, real and imaginary parts
    def lossf(v):
        w2, s2 = ops.analysis(torch.from_numpy(v).cuda(), p)
        return float(((w2 - tw).abs().pow(2).sum() + (s2 - ts).abs().pow(2).sum()).item())
    h = 1e-6
    for idx in [5, 17, 100, 511, 1000]:
        for d in (1.0, 1j):
            e = np.zeros_like(flm)
            e[idx] = d * h
            fd = (lossf(flm + e) - lossf(flm - e)) / (2 * h)
            # torch: grad = dL/d(re) + i dL/d(im)
            an = g[idx].real if d == 1.0 else g[idx].imag
            # central difference error: O(h^2) truncation (zero for a quadratic loss) plus
            # cancellation ~ eps * loss / h
            tol = 1e-6 * abs(an) + 20 * np.finfo(float).eps * l0 / h
            assert abs(fd - an) < tol, (idx, d, fd, an, tol)


def real_field_alm(rng, L, lmin=2):
    """Harmonics of a real Gaussian field with C_l = (l+1)^-2 for l >= lmin:
    a_l,-m = (-1)^m conj(a_lm), a_l0 real."""
    a = np.zeros(L * L, np.complex128)
    for l in range(lmin, L):
        sd = 1.0 / (l + 1.0)
        a[l * l + l] = rng.standard_normal() * sd
        m = np.arange(1, l + 1)
        v = (rng.standard_normal(l) + 1j * rng.standard_normal(l)) * sd / np.sqrt(2)
        a[l * l + l + m] = v
        a[l * l + l - m] = ((-1.0) ** m) * np.conj(v)
    return a


def test_c4_spin2_gradient(ops, oracle):
    """c4: L=256, N=4, B=2, J_min=2, spin 2 (E/B fields, seed 4; target seed 40).
    Loss sum_j ||W_j(f) - T_j||^2 (+ the scaling part); its autograd gradient
    (hand-written adjoint) against the reference adjoint routine
    (s2let_analysis_adjoint.c:304-369, the oracle's restatement) at this shape,
    and against a central finite difference (FP64, step 1e-6) along two random
    directions."""
    import torch
    L, N = 256, 4
    p = ops.parameters(L, N, 2.0, 2, 2)
    rng = np.random.default_rng(4)
    E, B = real_field_alm(rng, L), real_field_alm(rng, L)
    flm = -(E + 1j * B)
    rt = np.random.default_rng(40)
    Et, Bt = real_field_alm(rt, L), real_field_alm(rt, L)
    tw, ts = ops.analysis(torch.from_numpy(-(Et + 1j * Bt)).cuda(), p)
    x = torch.from_numpy(flm).cuda().requires_grad_(True)
    w, s = ops.wavelet_analysis(x, p)
    loss = (w - tw).abs().pow(2).sum() + (s - ts).abs().pow(2).sum()
    loss.backward()
    g = x.grad.detach()
    gw = (2 * (w.detach() - tw)).cpu().numpy()
    gs = (2 * (s.detach() - ts)).cpu().numpy()
    print("c4 oracle adjoint", flush=True)
    ref = oracle.analysis_adjoint_wav2lm(gw, gs, oracle.params(L, N, 2.0, 2, 2))
    e_adj = np.linalg.norm(g.cpu().numpy() - ref) / np.linalg.norm(ref)
    print("c4 gradient vs oracle adjoint %.3e" % e_adj, flush=True)
    assert e_adj < 1e-12

    def lossf(v):
        w2, s2 = ops.analysis(v, p)
        return float(((w2 - tw).abs().pow(2).sum() + (s2 - ts).abs().pow(2).sum()).item())

    h = 1e-6
    l0 = float(loss.item())
    for k in range(2):
        d = torch.from_numpy(rng.standard_normal(L * L) + 1j * rng.standard_normal(L * L)).cuda()
        d[:4] = 0
        d /= torch.linalg.vector_norm(d)
        fd = (lossf(x.detach() + h * d) - lossf(x.detach() - h * d)) / (2 * h)
        an = float(torch.real(torch.vdot(g, d)).item())  # d/dt L(f + t d) = Re <grad, d>
        tol = 1e-6 * abs(an) + 20 * np.finfo(float).eps * l0 / h
        print("c4 directional derivative", k, fd, an, tol, flush=True)
        assert abs(fd - an) < tol, (k, fd, an, tol)
