"""The FP64 10^x / log10 of the speciation kernels (clb_fastmath.h), built for the host
from the same source the device compiles (explicit fma, so the roundings agree), against
long-double references: a few ulp everywhere, relative accuracy near y = 1 where the
convergence judge takes the log of S_P / S_R."""
import os
import subprocess

from conftest import PKG, ROOT


def test_fastmath_accuracy(tmp_path):
    exe = tmp_path / "fastmath_check"
    src = os.path.join(ROOT, "tests", "cpp", "fastmath_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(PKG, "csrc"), src, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    e_exp, e_log, e_near1, e_pts, log1 = (float(v) for v in out)
    # a few ulp (1 ulp = 2.2e-16): six orders below anything the 1e-10 judge can see
    assert e_exp < 8e-16
    assert e_log < 6e-16
    assert e_near1 < 1e-15
    assert e_pts < 8e-16
    assert log1 == 0.0
