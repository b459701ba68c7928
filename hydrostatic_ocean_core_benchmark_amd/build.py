"""Build libhoc.so (the HIP kernels and the C-ABI) in-tree for gfx950.

    python -m hydrostatic_ocean_core_benchmark_amd.build

Flags that matter for results, not only speed:
  -ffp-contract=off   no fused multiply-add: every product is rounded on its
                      own, as in NumPy, so results are bitwise the oracle's
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions (fp32 paths only)
No -ffast-math: divisions and square roots stay IEEE correctly rounded.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libhoc.so"
SOURCES = ("hoc_abi.hip", "hoc_step_kernels.hip", "hoc_tiled.hip", "hoc_helmholtz.hip", "hoc_mg.hip",
           "hoc_adjoint.hip", "hoc_sw2d.hip", "hoc_helm2d.hip",
           "hoc_model3d.hip")
ARCH = os.environ.get("HOC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def command(out: Path = LIB) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I", str(CSRC), "-I", str(PKG.parent / "include"),
            *[str(CSRC / s) for s in SOURCES], "-o", str(out)]


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
