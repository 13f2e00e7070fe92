"""The FP32 theta kernel (k_theta<float>) that the mixed mode runs instead of the
bf16 three-part products when S2LET_THETA_BF=0: mixed analysis and synthesis
within 1e-6 of FP64, on the MW grid (L = 64, 96, both row-tile widths of the
launch, K not a multiple of the chunk at L = 96 + 1 rows on MWSS) and on the MWSS
grid.  The switch is read once per process, so each case runs in a child process."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, ROOT)
import torch
from s2let_amd import torch_ops as ops
L, N, ss = int(sys.argv[1]), int(sys.argv[2]), int(sys.argv[3])
p = ops.parameters(L, N, 2.0, 2, 0, sampling_scheme=ss)
rng = np.random.default_rng(L + 7 * N + ss)
ell = np.repeat(np.arange(L), 2 * np.arange(L) + 1)
flm = (rng.standard_normal(L * L) + 1j * rng.standard_normal(L * L)) / np.sqrt(2) / (ell + 1.0)
x = torch.from_numpy(flm).cuda()
w64, s64 = ops.analysis(x, p, ops.FP64)
w32, s32 = ops.analysis(x, p, ops.MIXED)
r64 = ops.synthesis(w64, s64, p, ops.FP64)
r32 = ops.synthesis(w64, s64, p, ops.MIXED)
torch.cuda.synchronize()
def rel(a, b):
    a = a.cpu().numpy(); b = b.cpu().numpy()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
e_ana = rel(torch.cat([w32, s32]), torch.cat([w64, s64]))
e_syn = rel(r32, r64)
e_rec = float(np.linalg.norm(r64.cpu().numpy() - flm) / np.linalg.norm(flm))
print("RESULT", e_ana, e_syn, e_rec, flush=True)
"""


@pytest.mark.parametrize("L,N,ss", [(64, 4, 0), (96, 3, 0), (96, 4, 1)])
def test_theta_fp32_kernel(L, N, ss):
    env = dict(os.environ, S2LET_THETA_BF="0")
    print("theta fp32 kernel", L, N, ss, flush=True)
    r = subprocess.run([sys.executable, "-c", "ROOT=%r\n" % ROOT + CHILD, str(L), str(N), str(ss)], env=env,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    line = [s for s in r.stdout.splitlines() if s.startswith("RESULT")][-1]
    e_ana, e_syn, e_rec = (float(v) for v in line.split()[1:])
    print("errors", e_ana, e_syn, e_rec, flush=True)
    assert e_ana < 1e-6 and e_syn < 1e-6
    assert e_rec < 1e-12
