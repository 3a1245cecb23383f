"""Build liblbm_hip.so for gfx950 with hipcc, in-tree.

The library is compiled from lbmengine.jl_amd/csrc: the step kernels once
per population storage (FP64, FP32, FP16, custom 16-bit) and part (one
lattice of D2Q9/D3Q19, batch, one lattice of D3Q15/D3Q27, the D3Q19
kernels that keep the moments for the adjoint), the
particle kernels, the adjoint kernels, and the host runtime that implements
include/lbm_hip.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent          # lbmengine.jl_amd/
CSRC = ROOT / "csrc"
LIB_PATH = ROOT / "liblbm_hip.so"
OBJ_DIR = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link liblbm_hip.so; returns its path."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ_DIR.mkdir(exist_ok=True)
    jobs = []
    for store, tag in ((0, "f64"), (1, "f32"), (2, "f16"), (3, "f16c")):
        for part, ptag in ((0, ""), (1, "_batch"), (2, "_q15q27"), (3, "_keep")):
            jobs.append([HIPCC, *COMMON, "-c", f"-DLBM_STORE={store}", f"-DLBM_PART={part}",
                         str(CSRC / "lbm_kernels.hip"), "-o", str(OBJ_DIR / f"lbm_kernels_{tag}{ptag}.o")])
    jobs.append([HIPCC, *COMMON, "-c", str(CSRC / "lbm_particles.hip"), "-o", str(OBJ_DIR / "lbm_particles.o")])
    jobs.append([HIPCC, *COMMON, "-c", str(CSRC / "lbm_adjoint.hip"), "-o", str(OBJ_DIR / "lbm_adjoint.o")])
    jobs.append([HIPCC, *COMMON, "-x", "hip", "-c", str(CSRC / "lbm_capi.cpp"), "-o", str(OBJ_DIR / "lbm_capi.o")])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for fut in [ex.submit(_run, j) for j in jobs]:
            fut.result()
    objs = sorted(str(p) for p in OBJ_DIR.glob("*.o"))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
