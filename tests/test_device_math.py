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


def test_goldberg_log1p_float32():
    # The mixed modes' FP32 sample arithmetic (csrc/qlognei_step.h,
    # log_fatplus_grad) takes softplus's log1p(e), e = exp(z) for z <= 20, as
    # log(u) e / (u - 1) with u = fl(1 + e) (Goldberg 1991, thm. 4), and e itself
    # when u rounds to 1.  With every step rounded to FP32 it stays within 4 ulp
    # of the FP64 log1p over the range; the plain log(1 + e) it replaces does not.
    import numpy as np
    z = np.linspace(-30.0, 20.0, 200001)
    e = np.exp(z).astype(np.float32)
    u = (np.float32(1) + e).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = (e / (u - np.float32(1))).astype(np.float32)
        sp = np.where(u == np.float32(1), e, (np.log(u).astype(np.float32) * q).astype(np.float32))
    ref = np.log1p(e.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    err = np.abs(sp.astype(np.float64) - ref) / ulp
    assert err.max() <= 4, err.max()
    naive = np.log(u).astype(np.float32)
    assert (np.abs(naive.astype(np.float64) - ref) / ulp).max() > 1000
