"""Regression tests for the batch-graph cache and the plan scratch.

A captured batch graph holds the device addresses of the Delta tables, which
are reallocated when a larger band-limit (or a mode that needs another tiled
copy) is first planned; it also depends on the global wavelet type.  The cache
must recapture in both cases: replays after those events must equal the eager
per-map results bit for bit.  Calls with the same parameters on two streams
share one plan's scratch and must be ordered by the library."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def powerlaw(rng, L, spin):
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    f = (rng.standard_normal(L * L) + 1j * rng.standard_normal(L * L)) / np.sqrt(2) / (ell + 1.0)
    f[:spin * spin] = 0
    return f


@pytest.fixture(scope="module")
def ops():
    import torch  # noqa: F401
    from s2let_amd import torch_ops
    return torch_ops


def small_batch(ops, mode, seed):
    import torch
    rng = np.random.default_rng(seed)
    shapes = [(32, 4, 0), (64, 1, 2), (48, 8, 0)]
    ps = [ops.parameters(L, N, 2.0, 1, s) for L, N, s in shapes]
    x = [torch.from_numpy(powerlaw(rng, L, s)).cuda() for L, _, s in shapes]
    fw = [torch.zeros(ops.n_wav(p), dtype=torch.complex128, device="cuda") for p in ps]
    fs = [torch.zeros(ops.n_scal(p), dtype=torch.complex128, device="cuda") for p in ps]
    return ps, x, fw, fs


def check_equal_eager(ops, ps, x, fw, fs, mode, tag):
    import torch
    torch.cuda.synchronize()
    for k, p in enumerate(ps):
        w, s = ops.analysis(x[k], p, mode)
        torch.cuda.synchronize()
        assert torch.equal(fw[k], w) and torch.equal(fs[k], s), (tag, k)


@pytest.mark.parametrize("mode", [1, 0])
def test_graph_replay_after_table_growth(ops, mode):
    """Capture at L <= 64, then plan L = 512 (tables grow and move), replay."""
    import s2let_amd as s2
    print("graph replay after table growth, mode", mode, flush=True)
    s2.release()  # tables start at the small band-limits of this batch
    ps, x, fw, fs = small_batch(ops, mode, 21)
    ops.batch(0, fw, fs, x, None, ps, mode, graph=True)
    check_equal_eager(ops, ps, x, fw, fs, mode, "first")
    # a plan of the other mode at a larger L moves the FP64 table (and leaves this mode's
    # tiled copy to be rebuilt by the next plan of this mode), then one of this mode
    big = ops.parameters(528, 1, 2.0, 7, 0)
    s2.prepare(big, 1 - mode)
    s2.prepare(big, mode)
    for t in fw + fs:
        t.zero_()
    for k in range(len(x)):
        x[k].mul_(-1.5)
    ops.batch(0, fw, fs, x, None, ps, mode, graph=True)
    check_equal_eager(ops, ps, x, fw, fs, mode, "after growth")


def test_graph_replay_after_wavelet_switch(ops):
    import torch
    import s2let_amd as s2
    print("graph replay after wavelet switch", flush=True)
    mode = 0
    ps, x, fw, fs = small_batch(ops, mode, 22)
    try:
        ops.batch(0, fw, fs, x, None, ps, mode, graph=True)
        check_equal_eager(ops, ps, x, fw, fs, mode, "s2dw")
        ref_s2dw = [t.clone() for t in fw]
        sizes = [(s2.n_wav(p), s2.n_scal(p)) for p in ps]
        s2.switch_wavtype(2)  # needlets: the S2DW scale schedule, so the same buffer sizes
        assert sizes == [(s2.n_wav(p), s2.n_scal(p)) for p in ps]
        ops.batch(0, fw, fs, x, None, ps, mode, graph=True)
        check_equal_eager(ops, ps, x, fw, fs, mode, "needlet")
        assert any(not torch.equal(a, b) for a, b in zip(ref_s2dw, fw))
    finally:
        s2.switch_wavtype(1)  # S2DW, the default


def test_two_streams_same_plan(ops):
    """Two analyses with the same parameters on two streams, back to back: each
    result equals the one computed alone."""
    import torch
    print("two streams, one plan", flush=True)
    L, N = 128, 8
    p = ops.parameters(L, N, 2.0, 2, 0)
    rng = np.random.default_rng(23)
    xs = [torch.from_numpy(powerlaw(rng, L, 0)).cuda() for _ in range(2)]
    refs = []
    for x in xs:
        refs.append(ops.analysis(x, p, 1))
        torch.cuda.synchronize()
    s1, s2_ = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for rep in range(3):
        with torch.cuda.stream(s1):
            a = ops.analysis(xs[0], p, 1)
        with torch.cuda.stream(s2_):
            b = ops.analysis(xs[1], p, 1)
        torch.cuda.synchronize()
        outs.append((a, b))
    for a, b in outs:
        assert torch.equal(a[0], refs[0][0]) and torch.equal(a[1], refs[0][1])
        assert torch.equal(b[0], refs[1][0]) and torch.equal(b[1], refs[1][1])
