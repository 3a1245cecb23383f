"""The node band shared by box blocks and tiles, checked on the CPU: the
helpers in lbm_step.hpp are host-callable, so tests/native/band_check.cpp
compiles them (hipcc, host code only runs) and checks the enumeration, the
coverage of every node two ticks need, and the node bound;
tests/native/link_box_check.cpp checks that the link box and the band rows
stay inside the lattice at any particle centre (huge, infinite, NaN) and
equal the integer formulas at ordinary ones."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(shutil.which(HIPCC) is None and not Path(HIPCC).exists(), reason="hipcc not installed")
def test_band_partition(tmp_path):
    exe = tmp_path / "band_check"
    cmd = [HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", "-Wno-unused-command-line-argument",
           f"-I{ROOT / 'lbmengine.jl_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "band_check.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe), "800"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok "), out.stdout
    assert int(out.stdout.split()[1]) > 500


@pytest.mark.skipif(shutil.which(HIPCC) is None and not Path(HIPCC).exists(), reason="hipcc not installed")
def test_link_box_and_band_rows_at_any_centre(tmp_path):
    exe = tmp_path / "link_box_check"
    cmd = [HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", "-Wno-unused-command-line-argument",
           f"-I{ROOT / 'lbmengine.jl_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "link_box_check.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok "), out.stdout
    assert int(out.stdout.split()[1]) == 20000
