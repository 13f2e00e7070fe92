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
BF16_PEAK = 2.5e15        # dense bf16 MFMA
THETA_BF_MAX_L = 512      # the FP32-mode theta stage runs on the bf16 matrix cores up to here
# roofs by bound: FP32 MFMA; the bf16 three-part products (six bf16 products per
# FP32-equivalent product); HBM
PEAKS = {"mfma": (FP32_PEAK / 1e12, "TFLOP/s"), "mfma_bf16x3": (BF16_PEAK / 6 / 1e12, "TFLOP/s"),
         "hbm": (HBM_PEAK / 1e9, "GB/s")}


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


def dft_factors(M):
    """M = M1 M2 with M1 the largest divisor <= sqrt(M) (the row stage's two-level DFT)."""
    m1 = int(M ** 0.5)
    while M % m1:
        m1 -= 1
    return m1, M // m1


def column_groups(s2, p):
    """(L, columns, gamma samples) of the scaling function and of every scale."""
    out = [(min(s2.bandlimit(p.J_min - 1, p), p.L), 1, 0)]
    for j in range(p.J_min, s2.j_max(p) + 1):
        Lj = min(s2.bandlimit(j, p), p.L)
        Nj = min(p.N, Lj)
        Nj += (Nj + p.N) % 2
        out.append((Lj, Nj, 2 * Nj - 1))
    return out


def kernel_work(s2, p, F, e):
    """Algorithmic work per step (one analysis + one synthesis) of each kernel
    function, with its roof: ("mfma", flop) or ("hbm", bytes).  e = bytes of one
    complex intermediate (8 mixed, 16 FP64); f_wav, f_scal and flm are complex128.
      k_inv3 / k_fwd3: the analysis / synthesis ell-sums, F/2 each (the path's
        algorithmic flop count);
      k_theta: the theta stage as real matrix products, 2 (directions) x 4 L^2 M ncol
        (k_theta_bf: the same products on the bf16 matrix cores, roof = bf16 peak / 6);
      k_rows4: the two-level DFT of each alpha line of the wavelet columns and phi line
        of the scaling function, M = M1 M2: 8 M1^2 M2 + 8 M2^2 M1 real flop (one product
        of 8 M^2 when M is prime);
      k_mix:   the column spectra and the f_wav slices read or written once;
      k_sum_parts: the m' split partials read once and flm written once;
      k_xtile: flm read (complex128) and its two operand tiles written once."""
    g = column_groups(s2, p)
    L = p.L
    theta = sum(2 * 4 * Lg * Lg * (2 * Lg - 1) * nc for Lg, nc, ng in g)
    rows = 0
    mix = 0
    for Lg, nc, ng in g:
        M = 2 * Lg - 1
        m1, m2 = dft_factors(M)
        per_line = 8 * M * M if m1 == 1 else 8 * m1 * m1 * m2 + 8 * m2 * m2 * m1
        rows += 2 * (nc if ng else 1) * Lg * per_line
        if ng:
            mix += 2 * Lg * M * (nc * e + ng * 16)
    S = max(1, min(8, L // 64))
    work = {"k_inv3": ("mfma", F / 2), "k_fwd3": ("mfma", F / 2), "k_theta": ("mfma", theta),
            "k_theta_bf": ("mfma_bf16x3", theta), "k_rows4": ("mfma", rows), "k_mix": ("hbm", mix),
            "k_sum_parts": ("hbm", (S + 1) * L * L * 16)}
    nx = (L + 15) // 16
    work["k_xtile"] = ("hbm", L * L * 16 + 2 * nx * nx * 256 * e)
    return work


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


def cpu_baseline(p_cfg, flm):
    """The oracle (FP64 C restatement of the reference algorithm, OpenMP) on one
    map of the same workload: analysis + synthesis."""
    from oracle import oracle as O
    threads = int(os.environ.get("S2LET_CPU_THREADS", "0")) or min(16, len(os.sched_getaffinity(0)))
    O.set_threads(threads)
    q = O.params(p_cfg["L"], p_cfg["N"], p_cfg["B"], p_cfg["J_min"], p_cfg["spin"], p_cfg["upsample"])
    t0 = time.perf_counter()
    fw, fs = O.analysis_lm2wav(flm, q)
    O.synthesis_wav2lm(fw, fs, q)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "maps/s", "cores": threads, "kind": "port",
            "sample": "1 map (analysis+synthesis) at L=%d N=%d J_min=%d, CPU FP64 oracle, %d OpenMP threads, %.1f s"
                      % (p_cfg["L"], p_cfg["N"], p_cfg["J_min"], threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--L", type=int, default=CFG["L"])
    ap.add_argument("--N", type=int, default=CFG["N"])
    ap.add_argument("--J_min", type=int, default=CFG["J_min"])
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device 3 times (reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import s2let_amd as s2
    from s2let_amd import torch_ops as ops

    if args.copy_calib > 0:
        src = torch.ones(args.copy_calib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

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
    # dominant kernel of the headline (mixed) mode: the kernel function with the
    # largest time per step, from the per-stage HIP events recorded on the
    # library's stream (each stage is one kernel function)
    per_k = {}
    for name, (cnt, ms) in st32.items():
        kf = stage_kernel(name, True, p.L)
        c0, m0 = per_k.get(kf, (0, 0.0))
        per_k[kf] = (c0 + cnt, m0 + ms)
    dk, (dcount, dms) = max(per_k.items(), key=lambda kv: kv[1][1])
    work = kernel_work(s2, p, F, 8)
    bound, w_step = work.get(dk, ("hbm", None))
    t_step = dms / K / 1e3  # s
    peak, unit = PEAKS[bound]
    achieved = None if w_step is None else w_step / t_step / (1e9 if bound == "hbm" else 1e12)
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", "pmc_traffic_L512_N8.json")).get(dk)
    roof = {"bound": "hbm" if bound == "hbm" else "mfma", "achieved": achieved, "peak": peak, "unit": unit,
            "peak_basis": {"mfma": "FP32 dense MFMA 157.3 TFLOP/s",
                           "mfma_bf16x3": "bf16 dense MFMA 2.5 PFLOP/s / 6 (three-part bf16 products per "
                                          "FP32-equivalent product)", "hbm": "HBM3E 8 TB/s"}[bound],
            "frac": None if achieved is None else achieved / peak, "traffic": traffic,
            "kernel": dk, "launches_per_step": dcount / K, "avg_launch_ms": dms / dcount,
            "work_per_launch": None if w_step is None else w_step * K / dcount}
    kernels = {}
    for kf, (cnt, ms) in per_k.items():
        b, w = work.get(kf, ("hbm", None))
        t = ms / K / 1e3
        rate = None if w is None else w / t / (1e9 if b == "hbm" else 1e12)
        kernels[kf] = {"ms_per_step": ms / K, "launches_per_step": cnt / K, "bound": b,
                       "achieved": rate, "frac": None if rate is None else rate / PEAKS[b][0]}
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
        "kernels_mixed": kernels,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, flm_h)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
