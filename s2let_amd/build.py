"""Builds s2let_amd/libs2let_amd.so (HIP kernels for gfx950 + C ABI) in-tree.

    python -m s2let_amd.build          # or s2let_amd.build.build()

Every HIP source is compiled with packed-FP32 instructions disabled; the
tiling host code and the harmonic multiplies are compiled with
-ffp-contract=off so that they round like the reference's C.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libs2let_amd.so")
ARCH = os.environ.get("S2LET_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# matrix-core accumulators in ordinary vector registers (no copies out of the
# accumulation registers before the vector units read a product)
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-mfma-vgpr-form", "-Wno-unused-result"]

SOURCES = [
    ("tiling.cpp", ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]),
    ("fits.cpp", ["g++", "-O2", "-std=c++17", "-fPIC"]),
    ("lm.cpp", ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]),
    ("multiply.hip", [HIPCC] + HIP_FLAGS + ["-ffp-contract=off"]),
    ("kernels.hip", [HIPCC] + HIP_FLAGS),
    ("contract_mfma.hip", [HIPCC] + HIP_FLAGS),
    ("theta_mfma.hip", [HIPCC] + HIP_FLAGS),
    ("theta_bf16.hip", [HIPCC] + HIP_FLAGS),
    ("rows_mfma.hip", [HIPCC] + HIP_FLAGS),
    ("plan.hip", [HIPCC] + HIP_FLAGS),
    ("pipeline.hip", [HIPCC] + HIP_FLAGS),
    ("api.hip", [HIPCC] + HIP_FLAGS),
]


def _compile(src, cmd):
    obj = os.path.join(BUILD, src + ".o")
    full = cmd + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(full, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(full), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, c) for s, c in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda sc: _compile(*sc), srcs))
    tmp = LIB + ".tmp"
    link = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
