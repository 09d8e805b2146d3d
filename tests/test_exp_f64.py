"""The step kernel's double-precision exponential (mf_common.h exp_f64:
Cody-Waite reduction by ln 2, degree-13 Taylor polynomial) against the C
library's exp, on the host: within 2 ulp over the normal range, within one
subnormal step below it, inf above 709.78, NaN through, 0 below -745.13 and
exactly 1 at 0.  The device evaluates the same operations (fma, rint,
ldexp) on the same coefficients."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_exp_f64_matches_libm(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "exp_check")
    r = subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
                        "-I", os.path.join(ROOT, "multifil_jax_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                        os.path.join(HERE, "exp_f64_check.cpp"), "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    worst, sub = int(out[0]), float(out[2])
    assert worst <= 2, f"{worst} ulp at x = {out[1]}"
    assert sub <= 5e-324
    assert out[3:] == ["1", "1", "1", "1"]
