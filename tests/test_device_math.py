"""Host-side checks of the device math helpers shared by the HIP kernels
(botorch_amd/csrc/bta_common.h), compiled with hipcc and run on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_exp_nonpos_matches_libm(tmp_path):
    # Cody-Waite + degree-12 Taylor e^x for x in [-745, 0]: within ~2 ulp of
    # libm over 2M points, exact 1 at 0, 0 below the underflow threshold,
    # NaN propagated (the program's own exit status encodes the bounds)
    exe = tmp_path / "exp_nonpos_check"
    src = os.path.join(HERE, "native", "exp_nonpos_check.hip")
    subprocess.run([HIPCC, "-O2", "--offload-arch=gfx950", "-std=c++17", src, "-o", str(exe)],
                   check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
