"""Writes tests/golden/kappas_B3_L128_Jmin2.npz from the reference's own tiling
fixtures src/matlab/kappas_{s2dw,need,spline}.mat (MATLAB v5 numeric arrays,
loaded with scipy.io.loadmat, which executes nothing from the file).

These arrays were produced by the reference's axisymmetric tiling at B=3,
L=128, J_min=2 (src/matlab/s2let_plot_wavs.m:7-9).  Run in a container that has
the reference checkout:  python tests/golden/make_kappa_fixtures.py
"""
import os

import numpy as np
import scipy.io as sio

SRC = "/root/reference/src/matlab"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kappas_B3_L128_Jmin2.npz")

arrays = {}
for tag in ("s2dw", "need", "spline"):
    d = sio.loadmat(os.path.join(SRC, "kappas_%s.mat" % tag))
    arrays["kappa_" + tag] = np.asarray(d["kappa_" + tag], dtype=np.float64)
    arrays["kappa0_" + tag] = np.asarray(d["kappa0_" + tag], dtype=np.float64).ravel()
np.savez(OUT, B=3.0, L=128, J_min=2, **arrays)
print("wrote", OUT, {k: v.shape for k, v in arrays.items()})
