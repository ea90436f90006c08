"""closure_dc, the closure difference of the mixed-precision defect
correction (bgy3d_amd/csrc/closure_math.hpp, shared with the kernels),
compiled for the host with g++ and checked against a long-double reference
f(x_b + d) - f(x_b) - d over x_b in [-30, 10], |d| in [1e-6, 1] for HNC,
KH and PSE-2/3/6: its FP32 error stays at FP32 rounding of the step size,
where the plain FP32 difference loses all digits for small steps; and it
stays finite deep in a repulsive core (x_b down to -1.9e9, |d| up to 1e3)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_closure_dc_rounding(tmp_path):
    exe = tmp_path / "closure_math_check"
    try:
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                        os.path.join(HERE, "closure_math_check.cpp")], check=True)
    except FileNotFoundError:
        pytest.skip("g++ not available")
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = {ln.split()[0]: [float(x) for x in ln.split()[1:]] for ln in out.strip().splitlines()}
    assert set(rows) == {"HNC", "KH", "PSE2", "PSE3", "PSE6"}
    for name, (e32, e64, naive, nonfinite) in rows.items():
        assert nonfinite == 0, name
        assert e32 < 1e-6, (name, e32)
        assert e64 < 1e-11, (name, e64)
        assert naive > 1e-2, (name, naive)
