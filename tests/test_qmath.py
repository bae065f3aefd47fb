"""qmath.h's exp and erf (the functions the 2D Gaussian MLE evaluates on the GPU and in the
oracle) against the exact values and against glibc's expf/erff, on the CPU.

The exact value stands in as glibc's double exp/erf; every 61st float of each range is
checked.  A sweep of every float (|x| <= 104 for exp, <= 5 for erf), run once, gave at most
0.951 ulp (exp) and 1.501 ulp (erf) from the exact values and 1 and 2 ulp from glibc's."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_expf_error(oracle):
    n, exact, libc, worst = oracle.qm_check(0, 104.0, 61)
    print(f"qm_expf: {n} floats, max {exact:.3f} ulp from exact (at {worst}), {libc:.3f} ulp from glibc expf")
    assert n > 3e7
    assert exact <= 1.0 and libc <= 1.0


def test_erff_error(oracle):
    n, exact, libc, worst = oracle.qm_check(1, 8.0, 61)
    print(f"qm_erff: {n} floats, max {exact:.3f} ulp from exact (at {worst}), {libc:.3f} ulp from glibc erff")
    assert n > 3e7
    assert exact <= 2.0 and libc <= 2.0


def test_special_values(oracle):
    e, f = oracle.L.oq_qm_expf, oracle.L.oq_qm_erff
    assert e(0.0) == 1.0 and e(float("-inf")) == 0.0 and e(float("inf")) == float("inf")
    assert math.isnan(e(float("nan"))) and math.isnan(f(float("nan")))
    assert e(89.0) == float("inf") and e(-104.0) == 0.0
    # the last finite result and the subnormal range, each within 1 ulp
    for x in (88.72283, -87.5, -100.0, -103.9):
        t = math.exp(np.float32(x))
        u = float(np.spacing(np.float32(t)))
        assert abs(e(x) - t) <= u, x
    assert f(0.0) == 0.0 and f(float("inf")) == 1.0 and f(float("-inf")) == -1.0
    assert f(4.0) == 1.0 and f(-4.0) == -1.0 and f(3.9) == np.float32(math.erf(3.9))
    # odd symmetry and the interval boundaries (1, 2, 4) within 2 ulp from either side
    for x in (0.3, 0.999999, 1.0, 1.0000001, 1.999999, 2.0, 3.99999):
        assert f(-x) == -f(x)
        t = math.erf(np.float32(x))
        assert abs(f(x) - t) <= 2 * float(np.spacing(np.float32(t))), x


def test_coefficients_match_their_derivation():
    """qmath.h holds exactly the block qmath_fit.py derives."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "qtrk_amd", "csrc", "qmath_fit.py")],
                       capture_output=True, text=True, check=True)
    src = open(os.path.join(ROOT, "qtrk_amd", "csrc", "qmath.h")).read()
    block = src.split("/* ---- generated by qmath_fit.py ---- */\n")[1].split("/* ---- end of generated block")[0]
    assert block.strip() == r.stdout.strip()


def test_gaussian_mle_uses_them(oracle):
    """The oracle's Gaussian model takes its erf from qmath.h: a spot's pixel equals the product of
    qm_erff differences formed in float (cputrack/utils.cpp:393-402)."""
    img = oracle.gaussian_spot(8, 8, 3.3, 4.1, 1.5, 1000.0, 10.0)
    ed = np.float32(1.0) / np.sqrt(np.float32(2 * 1.5 * 1.5))
    f = oracle.L.oq_qm_erff
    x, y = 5, 2
    h = np.float32(0.5)
    dx = h * np.float32(f(float((np.float32(x) - np.float32(3.3) + h) * ed))) - \
        h * np.float32(f(float((np.float32(x) - np.float32(3.3) - h) * ed)))
    dy = h * np.float32(f(float((np.float32(y) - np.float32(4.1) + h) * ed))) - \
        h * np.float32(f(float((np.float32(y) - np.float32(4.1) - h) * ed)))
    assert img[y, x] == np.float32(10.0) + np.float32(1000.0) * np.float32(dx) * np.float32(dy)
