"""Throughput of the directional wavelet transform on one MI355X.

Workload (the metric's configuration): one map = s2let_analysis_lm2wav +
s2let_synthesis_wav2lm at L=512, N=8, B=2, J_min=2, spin 0, MW sampling,
multiresolution, on complex128 f_lm with C_l = (l+1)^-2 (numpy seed 2).  A
step transforms one map; inputs and outputs stay resident in HBM.  The
headline value is the mixed-precision mode's maps/s; the FP64 mode is timed
in the same run, and the mixed result's error against FP64 is reported.

Prints one JSON line (see the fields below).  --gpus N > 1 is not supported
(one GPU per box).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CFG = dict(L=512, N=8, B=2.0, J_min=2, spin=0, upsample=0)
HBM_PEAK = 8.0e12        # B/s
FP64_PEAK = 78.6e12      # FLOP/s (dense)
FP32_PEAK = 157.3e12
THETA_BF_MAX_L = 512      # the FP32-mode theta stage runs on the bf16 matrix cores up to here
# roofs by bound: the ell-contractions on the MFMA roof of their precision, the
# streaming stages (DFTs, theta products, mixing) on HBM
PEAKS = {"mfma_fp32": (FP32_PEAK / 1e12, "TFLOP/s"), "mfma_fp64": (FP64_PEAK / 1e12, "TFLOP/s"),
         "hbm": (HBM_PEAK / 1e9, "GB/s")}
PEAK_BASIS = {"mfma_fp32": "FP32 dense MFMA 157.3 TFLOP/s", "mfma_fp64": "FP64 dense MFMA 78.6 TFLOP/s",
              "hbm": "HBM3E 8 TB/s"}


def powerlaw_flm(L, seed):
    rng = np.random.default_rng(seed)
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    return (rng.standard_normal(L * L) + 1j * rng.standard_normal(L * L)) / np.sqrt(2) / (ell + 1.0)


def algorithmic_work(s2, p):
    """F (flop/map, analysis+synthesis, 4 real flop per real-coefficient x complex
    MAC with the m' fold) and Q (unavoidable HBM bytes/map)."""
    J = s2.j_max(p)
    F = 0
    fwav = 0
    for j in range(p.J_min, J + 1):
        Lj = min(s2.bandlimit(j, p), p.L)
        Nj = min(p.N, Lj)
        Nj += (Nj + p.N) % 2
        L0 = s2.L0(j, p)
        for n in range(-Nj + 1, Nj, 2):
            for l in range(max(abs(n), abs(p.spin), L0), Lj):
                F += (l + 1) * (2 * l + 1)
        fwav += Lj * (2 * Lj - 1) * (2 * Nj - 1)
    Ls = min(s2.bandlimit(p.J_min - 1, p), p.L)
    for l in range(abs(p.spin), Ls):
        F += (l + 1) * (2 * l + 1)
    F *= 8
    Q = 16 * (2 * p.L * p.L + 2 * fwav + 2 * Ls * (2 * Ls - 1))
    return F, Q


# one kernel function per timed stage of the library (s2let_hip_timing)
STAGE_KERNEL = {
    "contract_inv": "k_inv3", "contract_fwd": "k_fwd3", "contract_fwd_sum": "k_sum_parts",
    "beta_inv": "k_theta", "beta_fwd": "k_theta", "beta_inv_adj": "k_theta", "beta_fwd_adj": "k_theta",
    "rows": "k_rows4", "gamma": "k_mix", "xtile": "k_xtile",
}


def stage_kernel(name, mixed, L):
    k = STAGE_KERNEL.get(name, name)
    if k == "k_theta" and mixed and L <= THETA_BF_MAX_L:
        return "k_theta_bf"
    return k


def kernel_work(info, F, mixed):
    """Algorithmic work per step (one analysis + one synthesis) of each kernel function
    and the roof it is scored on, from the plan's layout (s2let_hip_plan_info):
    the ell-contractions (k_inv3, k_fwd3) are dense contractions, F/2 each on the MFMA
    roof of the mode's precision; every other stage is a streaming pass scored in HBM
    bytes (its FFT / matrix-product flops are overhead):
      k_theta(_bf): the column spectra read and written once per direction;
      k_rows4:      the same for the alpha / phi lines (the scaling function's map is
                    complex128 on the caller's side);
      k_mix:        the column spectra and the f_wav slices, once per direction;
      k_sum_parts:  the m' split partials read once and flm written once;
      k_xtile:      flm read (complex128) and its two operand tiles written once."""
    e = info["complex_bytes"]
    groups = info["groups"]
    spec = info["spec_elems"]
    L = info["L"]
    scal_plane = groups[0][0] * (2 * groups[0][0] - 1)
    theta = 2 * 2 * spec * e
    rows = 2 * (2 * (spec - scal_plane) * e + scal_plane * (e + 16))
    mix = sum(2 * Lg * (2 * Lg - 1) * (nc * e + ng * 16) for Lg, nc, ng in groups if ng)
    S = info["splits"]
    nx = (L + 15) // 16
    peak_mfma = "mfma_fp32" if mixed else "mfma_fp64"
    return {"k_inv3": (peak_mfma, F / 2), "k_fwd3": (peak_mfma, F / 2), "k_theta": ("hbm", theta),
            "k_theta_bf": ("hbm", theta), "k_rows4": ("hbm", rows), "k_mix": ("hbm", mix),
            "k_sum_parts": ("hbm", (S + 1) * L * L * 16 if S > 1 else 0),
            "k_xtile": ("hbm", L * L * 16 + 2 * nx * nx * 256 * e)}


def pmc_traffic(path):
    """HBM bytes per launch of each kernel function (mixed mode) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, or {} if absent."""
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch_mixed", {})
    except (OSError, ValueError):
        return {}


def timed(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p_cfg, flm, reps):
    """The oracle (FP64 C restatement of the reference algorithm) on the host cores:
    all the cores this process may use (OpenMP) on one map of the workload, the
    median of `reps` runs, with and without the per-call tiling generation (the
    reference regenerates the tiling in every lm2wav / wav2lm call); and one thread,
    as the reference runs, on a bounded sample: one map at L=256 with the same N, B,
    J_min (an eighth of the L=512 map's ell-sum work), median of three, to keep the run
    short (about 30 s of CPU work in all)."""
    from oracle import oracle as O
    affinity = len(os.sched_getaffinity(0))
    cores = int(os.environ.get("S2LET_CPU_THREADS", "0")) or int(os.environ.get("OMP_NUM_THREADS", "0")) or affinity
    q = O.params(p_cfg["L"], p_cfg["N"], p_cfg["B"], p_cfg["J_min"], p_cfg["spin"], p_cfg["upsample"])

    def one_map(qq, x):
        t0 = time.perf_counter()
        fw, fs = O.analysis_lm2wav(x, qq)
        O.synthesis_wav2lm(fw, fs, qq)
        return time.perf_counter() - t0

    O.set_threads(cores)
    t_tiling = sorted(_timeit(lambda: O.tiling_wavelet(q)) for _ in range(5))[2]
    ts = sorted(one_map(q, flm) for _ in range(max(1, reps)))
    t_all = ts[len(ts) // 2]
    O.set_threads(1)
    L1 = min(256, p_cfg["L"])
    q1 = O.params(L1, p_cfg["N"], p_cfg["B"], p_cfg["J_min"], p_cfg["spin"], p_cfg["upsample"])
    x1 = powerlaw_flm(L1, 2)
    t_one = sorted(one_map(q1, x1) for _ in range(3))[1]
    O.set_threads(cores)
    return {"value": 1.0 / t_all, "unit": "maps/s", "cores": cores, "kind": "port",
            "sample": "median of %d map(s) (analysis+synthesis) at L=%d N=%d J_min=%d, CPU FP64 oracle, %d OpenMP "
                      "threads, %.1f s per map" % (len(ts), p_cfg["L"], p_cfg["N"], p_cfg["J_min"], cores, t_all),
            "cpu_model": cpu_model(), "cores_available": affinity,
            "excluding_tiling": {"value": 1.0 / max(1e-9, t_all - 2 * t_tiling), "unit": "maps/s",
                                 "tiling_s_per_call": t_tiling},
            "single_thread": {"value": 1.0 / t_one, "unit": "maps/s", "cores": 1,
                              "sample": "median of 3 maps at L=%d N=%d J_min=%d, one thread, %.1f s per map" % (
                                  L1, p_cfg["N"], p_cfg["J_min"], t_one)}}


def _timeit(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def c3_batch(s2, ops, torch, steps, warmup, modes=("fp64", "mixed"), before_timed=None):
    """The uneven batch: 256 maps, L in {64,128,256,512}, N in {1,4,8}, spin in {0,2},
    B=2, J_min=1, f_lm with C_l = (l+1)^-2, drawn with numpy default_rng(3); one step =
    the analysis batch then the synthesis batch, each one cached HIP graph.
    before_timed() runs right before each mode's warm-up and timed batches (after the
    autotuner's trial launches), e.g. to mark where a counter profile starts."""
    rng = np.random.default_rng(3)
    shapes = [(int(rng.choice([64, 128, 256, 512])), int(rng.choice([1, 4, 8])), int(rng.choice([0, 2])))
              for _ in range(256)]
    flms = []
    for L, _, spin in shapes:
        ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
        f = (rng.standard_normal(L * L) + 1j * rng.standard_normal(L * L)) / np.sqrt(2) / (ell + 1.0)
        f[:spin * spin] = 0
        flms.append(f)
    ps = [s2.parameters(L, N, 2.0, 1, sp) for L, N, sp in shapes]
    x = [torch.from_numpy(f).cuda() for f in flms]
    out = {"maps": len(ps)}
    # launch shapes per band-limit bucket (outside the timed region): one autotune per
    # distinct parameter set and mode
    distinct = {(L, N, sp): p for (L, N, sp), p in zip(shapes, ps)}
    out["buckets"] = len(distinct)
    for mode in (s2.FP64, s2.MIXED):
        if ("fp64" if mode == s2.FP64 else "mixed") in modes:
            for p in distinct.values():
                ops.autotune(p, mode)
    F_total = 0
    for p in ps:
        F_total += algorithmic_work(s2, p)[0]
    for mode, name in ((s2.FP64, "fp64"), (s2.MIXED, "mixed")):
        if name not in modes:
            continue
        fw = [torch.zeros(s2.n_wav(p), dtype=torch.complex128, device="cuda") for p in ps]
        fs = [torch.zeros(s2.n_scal(p), dtype=torch.complex128, device="cuda") for p in ps]
        rec = [torch.zeros_like(t) for t in x]

        def step():
            ops.batch(0, fw, fs, x, None, ps, mode, graph=True)
            ops.batch(1, rec, None, fw, fs, ps, mode, graph=True)

        if before_timed is not None:
            before_timed()
        dt = timed(step, steps, warmup, torch)
        worst = max(float((rec[k] - x[k]).abs().max() / x[k].abs().max()) for k in range(len(ps)))
        out[name] = {"value": steps * len(ps) / dt, "unit": "maps/s", "ms_per_batch": 1e3 * dt / steps,
                     "max_reconstruction_rel_err": worst}
        del fw, fs, rec
    out["flop_per_batch"] = F_total
    pmc = pmc_c3(os.path.join(ROOT, "profiles", "pmc_traffic_c3.json"))
    for name in modes:
        if name in pmc:
            out[name]["hbm_bytes_per_map"] = pmc[name]
    return out


def c2_batched(s2, ops, torch, p, nmaps, steps, warmup):
    """The headline shape as a batch: nmaps maps of L=512, N=8 (f_lm drawn like the
    headline's, numpy seeds 2, 3, ...) through the batched entry point (one bucket: every
    launch covers all maps; one cached HIP graph per direction).  Maps/s per mode and the
    worst FP64 reconstruction error; informational (the headline is one map per step)."""
    x = [torch.from_numpy(powerlaw_flm(p.L, 2 + k)).cuda() for k in range(nmaps)]
    ps = [p] * nmaps
    out = {"maps": nmaps}
    for mode, name in ((s2.FP64, "fp64"), (s2.MIXED, "mixed")):
        fw = [torch.zeros(s2.n_wav(p), dtype=torch.complex128, device="cuda") for _ in ps]
        fs = [torch.zeros(s2.n_scal(p), dtype=torch.complex128, device="cuda") for _ in ps]
        rec = [torch.zeros_like(t) for t in x]

        def step():
            ops.batch(0, fw, fs, x, None, ps, mode, graph=True)
            ops.batch(1, rec, None, fw, fs, ps, mode, graph=True)

        dt = timed(step, steps, warmup, torch)
        worst = max(float((rec[k] - x[k]).abs().max() / x[k].abs().max()) for k in range(nmaps))
        out[name] = {"value": steps * nmaps / dt, "unit": "maps/s", "ms_per_batch": 1e3 * dt / steps,
                     "max_reconstruction_rel_err": worst}
        del fw, fs, rec
    return out


def pmc_c3(path):
    """HBM bytes per map of the 256-map batch per mode, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of `bench.py --c3-only`, or {} if absent."""
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_map", {})
    except (OSError, ValueError):
        return {}


def real_half_spectrum(s2, ops, torch, p, flm_h, steps, warmup):
    """A real signal at the headline shape: analysis + synthesis through the real
    entry points (half-spectrum plan, n >= 0 directions) against the complex entry
    points on the same (Hermitian) input, per mode: maps/s, speed-up, and the
    reconstruction error of the real path."""
    L = p.L
    h = flm_h.copy()
    for el in range(L):  # Hermitian part: f_l,-m = (-1)^m conj(f_lm)
        m = np.arange(1, el + 1)
        pos, neg = el * el + el + m, el * el + el - m
        a, b = h[pos].copy(), h[neg].copy()
        h[pos] = 0.5 * (a + (-1.0) ** m * np.conj(b))
        h[neg] = (-1.0) ** m * np.conj(h[pos])
        h[el * el + el] = h[el * el + el].real
    x = torch.from_numpy(h).cuda()
    out = {}
    for mode, name in ((s2.FP64, "fp64"), (s2.MIXED, "mixed")):
        fw, fs = ops.analysis_real(x, p, mode)
        rec = ops.synthesis_real(fw, fs, p, mode)
        cw, cs = ops.analysis(x, p, mode)
        crec = ops.synthesis(cw, cs, p, mode)

        def real_step():
            ops.analysis_real(x, p, mode, out=(fw, fs))
            ops.synthesis_real(fw, fs, p, mode, out=rec)

        def cx_step():
            ops.analysis(x, p, mode, out=(cw, cs))
            ops.synthesis(cw, cs, p, mode, out=crec)

        tr = timed(real_step, steps, warmup, torch)
        tc = timed(cx_step, steps, warmup, torch)
        err = float(torch.linalg.norm(rec - x) / torch.linalg.norm(x))
        out[name] = {"value": steps / tr, "unit": "maps/s", "complex_maps_per_s": steps / tc, "speedup": tc / tr,
                     "reconstruction_rel_err": err}
        del fw, fs, rec, cw, cs, crec
    return out


def c5_single(s2, ops, torch, steps, warmup):
    """L = 2048, N = 8, B = 2, J_min = 3, spin 0, multiresolution, f_lm with C_l = (l+1)^-2
    (numpy default_rng(5)): analysis + synthesis maps/s in both modes, the FP64 round trip
    and the mixed error against FP64, and the stage times."""
    L, N, J_min = 2048, 8, 3
    p = s2.parameters(L, N, 2.0, J_min, 0)
    rng = np.random.default_rng(5)
    ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
    flm = (rng.standard_normal(L * L) + 1j * rng.standard_normal(L * L)) / np.sqrt(2) / (ell + 1.0)
    x = torch.from_numpy(flm).cuda()
    out = {"L": L, "N": N, "J_min": J_min}
    res = {}
    for mode, name in ((s2.FP64, "fp64"), (s2.MIXED, "mixed")):
        t0 = time.perf_counter()
        s2.prepare(p, mode)
        ops.autotune(p, mode)
        prep = time.perf_counter() - t0
        fw, fs = ops.analysis(x, p, mode)
        rec = ops.synthesis(fw, fs, p, mode)

        def step():
            ops.analysis(x, p, mode, out=(fw, fs))
            ops.synthesis(fw, fs, p, mode, out=rec)

        dt = timed(step, steps, warmup, torch)
        s2.timing(True)
        step()
        torch.cuda.synchronize()
        st = s2.timing_report()
        s2.timing(False)
        res[name] = (torch.cat([fw, fs]).cpu().numpy(), rec.cpu().numpy())
        out[name] = {"value": steps / dt, "unit": "maps/s", "ms_per_step": 1e3 * dt / steps,
                     "plan_and_autotune_s": prep, "stages_ms": {k: v[1] for k, v in st.items()}}
        del fw, fs, rec
    a64, r64 = res["fp64"]
    a32, r32 = res["mixed"]
    out["fp64"]["reconstruction_rel_err"] = float(np.linalg.norm(r64 - flm) / np.linalg.norm(flm))
    out["mixed"]["analysis_rel_err_vs_fp64"] = float(np.linalg.norm(a32 - a64) / np.linalg.norm(a64))
    out["mixed"]["reconstruction_rel_err"] = float(np.linalg.norm(r32 - flm) / np.linalg.norm(flm))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 256-map uneven batch")
    ap.add_argument("--no-real", action="store_true", help="skip the real-signal (half-spectrum) comparison")
    ap.add_argument("--no-c2-batch", action="store_true", help="skip the 8-map batch of the headline shape")
    ap.add_argument("--c5", action="store_true", help="also the L = 2048 single map (a few minutes of table builds)")
    ap.add_argument("--c3-only", choices=("fp64", "mixed", "both"), default=None,
                    help="run only the 256-map batch (in the given modes, --steps batches after --warmup)")
    ap.add_argument("--cpu-reps", type=int, default=1, help="CPU baseline maps (median)")
    ap.add_argument("--L", type=int, default=CFG["L"])
    ap.add_argument("--N", type=int, default=CFG["N"])
    ap.add_argument("--J_min", type=int, default=CFG["J_min"])
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="copy MIB MiB device-to-device 3 times (reference kernel for PMC byte counters): first "
                         "thing, or with --c3-only right before each mode's warm-up and timed batches")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import s2let_amd as s2
    from s2let_amd import torch_ops as ops

    def copy_calib():
        src = torch.ones(args.copy_calib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.c3_only:
        modes = ("fp64", "mixed") if args.c3_only == "both" else (args.c3_only,)
        mark = copy_calib if args.copy_calib > 0 else None
        print(json.dumps({"c3_batch": c3_batch(s2, ops, torch, args.steps, args.warmup, modes, mark)}))
        return

    cfg = dict(CFG, L=args.L, N=args.N, J_min=args.J_min)
    p = s2.parameters(cfg["L"], cfg["N"], cfg["B"], cfg["J_min"], cfg["spin"], cfg["upsample"])
    flm_h = powerlaw_flm(cfg["L"], 2)
    flm = torch.from_numpy(flm_h).cuda()
    nw, ns = s2.n_wav(p), s2.n_scal(p)
    bufs = {}
    for mode in (s2.FP64, s2.MIXED):
        s2.prepare(p, mode)
        ops.autotune(p, mode)  # launch shapes per band-limit (outside the timed region)
        bufs[mode] = (torch.zeros(nw, dtype=torch.complex128, device="cuda"),
                      torch.zeros(ns, dtype=torch.complex128, device="cuda"),
                      torch.zeros(cfg["L"] ** 2, dtype=torch.complex128, device="cuda"))
    # the copy that calibrates the rocprofv3 byte counters marks the end of the set-up
    # (plans, tables, autotuner trials): the dispatches after it are the measured ones
    if args.copy_calib > 0:
        copy_calib()

    def step(mode):
        fw, fs, rec = bufs[mode]
        ops.analysis(flm, p, mode, out=(fw, fs))
        ops.synthesis(fw, fs, p, mode, out=rec)

    res = {}
    for mode in (s2.FP64, s2.MIXED):
        dt = timed(lambda: step(mode), args.steps, args.warmup, torch)
        s2.timing(True)
        for _ in range(args.steps):
            step(mode)
        torch.cuda.synchronize()
        stages = s2.timing_report()
        s2.timing(False)
        res[mode] = (dt, stages)

    # the reference's host-pointer entry points (FP64, synchronous): rate including
    # the PCIe copies of f_lm in and f_wav / f_scal out and back (informational)
    fw_h, fs_h = s2.analysis_lm2wav(flm_h, p)
    s2.synthesis_wav2lm(fw_h, fs_h, p)
    t0 = time.perf_counter()
    for _ in range(3):
        fw_h, fs_h = s2.analysis_lm2wav(flm_h, p)
        s2.synthesis_wav2lm(fw_h, fs_h, p)
    host_rate = 3 / (time.perf_counter() - t0)
    del fw_h, fs_h

    # errors (outside the timed region)
    fw64, fs64, rec64 = bufs[s2.FP64]
    fw32, fs32, rec32 = bufs[s2.MIXED]
    a64 = torch.cat([fw64, fs64]).cpu().numpy()
    a32 = torch.cat([fw32, fs32]).cpu().numpy()
    err_ana = float(np.linalg.norm(a32 - a64) / np.linalg.norm(a64))
    r64, r32 = rec64.cpu().numpy(), rec32.cpu().numpy()
    err_syn = float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))
    err_rec64 = float(np.linalg.norm(r64 - flm_h) / np.linalg.norm(flm_h))

    F, Q = algorithmic_work(s2, p)
    dt32, st32 = res[s2.MIXED]
    dt64, st64 = res[s2.FP64]
    K = args.steps
    maps_s = K / dt32

    def kernel_table(stages, mixed):
        """per kernel function: launches and ms per step (the library's per-stage HIP
        events on the stream the kernels run on), its algorithmic work and roof"""
        info = s2.plan_info(p, s2.MIXED if mixed else s2.FP64)
        work = kernel_work(info, F, mixed)
        per_k = {}
        for name, (cnt, ms) in stages.items():
            kf = stage_kernel(name, mixed, p.L)
            c0, m0 = per_k.get(kf, (0, 0.0))
            per_k[kf] = (c0 + cnt, m0 + ms)
        table = {}
        for kf, (cnt, ms) in per_k.items():
            b, w = work.get(kf, ("hbm", None))
            t = ms / K / 1e3
            rate = None if w is None else w / t / (1e9 if b == "hbm" else 1e12)
            table[kf] = {"ms_per_step": ms / K, "launches_per_step": cnt / K, "bound": b,
                         "work_per_launch": None if w is None else w * K / cnt,
                         "unit": "B" if b == "hbm" else "flop",
                         "achieved": rate, "frac": None if rate is None else rate / PEAKS[b][0]}
        return table, info

    kernels32, info32 = kernel_table(st32, True)
    kernels64, _ = kernel_table(st64, False)
    # dominant kernel of the headline (mixed) mode: the largest time per step
    dk = max(kernels32, key=lambda k: kernels32[k]["ms_per_step"])
    d = kernels32[dk]
    peak, unit = PEAKS[d["bound"]]
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", "pmc_traffic_L512_N8.json")).get(dk)
    # path roofline: the time the map would take at the binding roof (F on FP32 MFMA, Q on HBM)
    t32 = dt32 / K
    t64 = dt64 / K
    bound32 = max(F / FP32_PEAK, Q / HBM_PEAK)
    bound64 = max(F / FP64_PEAK, Q / HBM_PEAK)
    roof = {"bound": "hbm" if d["bound"] == "hbm" else "mfma", "achieved": d["achieved"], "peak": peak,
            "unit": unit, "peak_basis": PEAK_BASIS[d["bound"]], "frac": d["frac"], "traffic": traffic,
            "kernel": dk, "launches_per_step": d["launches_per_step"],
            "avg_launch_ms": d["ms_per_step"] / d["launches_per_step"], "work_per_launch": d["work_per_launch"],
            "path": {"mixed": {"bound_us": 1e6 * bound32, "binding": "hbm" if Q / HBM_PEAK > F / FP32_PEAK
                               else "mfma_fp32", "measured_us": 1e6 * t32, "frac": bound32 / t32},
                     "fp64": {"bound_us": 1e6 * bound64, "binding": "hbm" if Q / HBM_PEAK > F / FP64_PEAK
                              else "mfma_fp64", "measured_us": 1e6 * t64, "frac": bound64 / t64}}}
    out = {
        "metric": "analysis+synthesis maps/s at L=512,N=8: mixed-precision vs FP64, err<=1e-6",
        "value": maps_s,
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt32 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (bf16x3 and f32 MFMA) + f64",
        "data": "synthetic: complex128 f_lm, C_l=(l+1)^-2, numpy default_rng(2)",
        "config": {"workload": "directional wavelet analysis+synthesis, one map per step", "L": cfg["L"],
                   "N": cfg["N"], "B": cfg["B"], "J_min": cfg["J_min"], "spin": cfg["spin"],
                   "sampling": "MW", "multiresolution": True},
        "fp64": {"value": K / dt64, "ms_per_step": 1e3 * dt64 / K, "reconstruction_rel_err": err_rec64},
        "host_api_fp64_maps_per_s": host_rate,
        "mixed_err_vs_fp64": {"analysis": err_ana, "synthesis": err_syn},
        "stages_ms_per_step": {"mixed": {k: v[1] / K for k, v in st32.items()},
                               "fp64": {k: v[1] / K for k, v in st64.items()}},
        "algorithmic": {"flop_per_map": F, "bytes_per_map": Q},
        "roofline": roof,
        "kernels_mixed": kernels32,
        "kernels_fp64": kernels64,
        "plan_mixed": {"splits": info32["splits"], "spec_elems": info32["spec_elems"]},
    }
    if not args.no_real:
        out["real_signal"] = real_half_spectrum(s2, ops, torch, p, flm_h, max(2, args.steps // 2), 1)
    if not args.no_c2_batch:
        out["c2_batch8"] = c2_batched(s2, ops, torch, p, 8, max(2, args.steps // 4), 1)
    if not args.no_c3:
        out["c3_batch"] = c3_batch(s2, ops, torch, max(2, args.steps // 4), 1)
    if args.c5:
        out["c5_L2048"] = c5_single(s2, ops, torch, 3, 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, flm_h, args.cpu_reps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
