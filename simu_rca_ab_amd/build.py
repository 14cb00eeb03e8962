"""In-tree build of the gfx950 HIP library (``lib/librca_das.so``).

hipcc cross-compiles for gfx950 without a GPU.  Packed-FP32 instructions are
disabled for every HIP translation unit.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = [PKG / "csrc" / n for n in ("rca_das.hip", "rca_plan.hip", "rca_svd.hip", "rca_sim.hip",
                                      "rca_probe.hip", "rca_noalloc.hip", "rca_quant.hip")]
OUT = PKG / "lib" / "librca_das.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", str(ROOT / "include")]
# the clutter filter calls rocBLAS (ZGEMM) and rocSOLVER (ZHEEVD); the
# simulator's inverse transform is hipFFT
LIBS = ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]


def build_library(verbose: bool = False) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-o", str(tmp), *map(str, SOURCES), *LIBS]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
