"""Build the gfx950 shared library ``libzandhip.so`` in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc", "zh_api.hip")
DEPS = [os.path.join(PKG, "csrc", f)
        for f in ("zh_api.hip", "zh_gemm.hpp", "zh_kernels.hpp", "zh_scf.hpp", "zh_omega.hpp", "zh_omgxi.hpp", "zh_psi.hpp")] + [
    os.path.join(ROOT, "include", "zandhip.h")]
LIB = os.path.join(PKG, "libzandhip.so")
ARCH = os.environ.get("ZH_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 VALU ops stay off (the compiler would otherwise form
    # v_pk_{add,mul,fma}_f32 from float2 code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libzandhip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
