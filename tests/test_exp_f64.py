"""The step kernel's double-precision exponential and crossbridge angle
(mf_common.h exp_f64: Cody-Waite reduction by ln 2, degree-13 Taylor
polynomial; atan2_pos_f64: octant reduction, 22-term odd Taylor series on
|u| <= tan(pi/8)) against the C library, on the host.  exp: within 2 ulp
over the normal range, within one subnormal step below it, inf above
709.78, NaN through, 0 below -745.13 and exactly 1 at 0.  atan2: within 4
ulp over spacings 0.5-40 nm and offsets -60..60 nm, exact on the axis and
the diagonals.  The device evaluates the same operations (fma, rint, ldexp,
division) on the same coefficients."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_exp_and_atan2_match_libm(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "exp_check")
    r = subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
                        "-I", os.path.join(ROOT, "multifil_jax_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                        os.path.join(HERE, "exp_f64_check.cpp"), "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines()
    out = lines[0].split()
    worst, sub = int(out[0]), float(out[2])
    assert worst <= 2, f"{worst} ulp at x = {out[1]}"
    assert sub <= 5e-324
    assert out[3:] == ["1", "1", "1", "1"]
    at = lines[1].split()
    assert int(at[0]) <= 4, f"atan2: {at[0]} ulp at y = {at[1]}, x = {at[2]}"
    assert at[3] == "3"
