"""Building blocks of the discrete adjoint, checked on the CPU against
central finite differences of the forward routines they reverse
(tests/native/adjoint_check.cpp, compiled with hipcc; host code only runs):
the collision VJP (BGK/TRT, with and without a force field), the squirmer
node-velocity VJP, and the dual-number Jacobian of the particle move, whose
primal must equal move_gpu bit for bit."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(shutil.which(HIPCC) is None and not Path(HIPCC).exists(), reason="hipcc not installed")
def test_adjoint_building_blocks(tmp_path):
    exe = tmp_path / "adjoint_check"
    cmd = [HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", "-Wno-unused-command-line-argument",
           f"-I{ROOT / 'lbmengine.jl_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "adjoint_check.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"
