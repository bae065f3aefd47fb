"""CPU pins of two exactness facts the kernels rely on to stay bit-identical to the reference.

* LUT scores (cputrack/cpu_tracker.cpp:829-836): d = (double)(float)(p - l); without radial weights
  the kernels compute the term and the sum as one fused multiply-add, fma(-d, d, s).  That rounds
  as the reference's s + (-d*d) because d*d is exact in double (a 24-bit significand squared, far
  inside double's exponent range): the fused and unfused forms then add the same exact value.
* NormalizeRadialProfile (cputrack/utils.cpp:230-261) sums floats in double in index order; the
  kernel sums them as a tree when their exponents span at most 28 - ceil(log2 n) binades, where
  every partial sum of any subset is exact, so every order gives the sequential result.
"""
import math
from fractions import Fraction

import numpy as np
import pytest


def _rand_floats(rng, n, emin, emax):
    """float32 values with random signs, mantissas and exponents in [emin, emax] (unbiased)."""
    m = rng.uniform(1.0, 2.0, n)
    e = rng.integers(emin, emax + 1, n)
    s = rng.choice([-1.0, 1.0], n)
    return (s * np.ldexp(m, e)).astype(np.float32)


def test_square_of_float_difference_is_exact_in_double():
    rng = np.random.default_rng(7)
    p = _rand_floats(rng, 4000, -30, 30)
    l = _rand_floats(rng, 4000, -30, 30)
    d = (p - l).astype(np.float32)                     # the float difference
    for v in d.tolist():                               # as double
        assert Fraction(v) ** 2 == Fraction(v * v), v  # d*d rounds to itself: exact
    # the extremes of the float range stay exact too (no overflow or underflow in double)
    for v in (float(np.finfo(np.float32).max), float(np.finfo(np.float32).smallest_subnormal), 1.0 + 2.0 ** -23):
        assert Fraction(v) ** 2 == Fraction(v * v)


def test_unfused_term_equals_exact_sum_rounded():
    """s + (-d*d) rounds the exact s - d^2 once, as a fused multiply-add would."""
    rng = np.random.default_rng(8)
    d = _rand_floats(rng, 3000, -12, 12).astype(np.float64)
    s = -np.abs(_rand_floats(rng, 3000, -8, 20).astype(np.float64))
    for dv, sv in zip(d.tolist(), s.tolist()):
        exact = Fraction(sv) - Fraction(dv) ** 2
        assert sv + (-dv * dv) == float(exact)  # Python's float(Fraction) rounds to nearest even


def _seq(v):
    s = 0.0
    for x in v.tolist():
        s += x
    return s


def _tree(v):
    v = [float(x) for x in v.tolist()]
    while len(v) > 1:
        v = [v[i] + v[i + 1] if i + 1 < len(v) else v[i] for i in range(0, len(v), 2)]
    return v[0] if v else 0.0


def _lg(n):
    return max(0, math.ceil(math.log2(n)))


@pytest.mark.parametrize("n", [16, 64, 100, 256])
def test_tree_sum_equals_sequential_within_exponent_span(n):
    rng = np.random.default_rng(n)
    span = 28 - _lg(n)
    for _ in range(300):
        e0 = int(rng.integers(-40, 40))
        v = _rand_floats(rng, n, e0, e0 + span)
        order = rng.permutation(n)
        assert _tree(v) == _seq(v) == _seq(v[order]) == float(sum(Fraction(float(x)) for x in v.tolist()))


def test_wider_spans_can_round():
    """The bound is not vacuous: with spans well past it, orders disagree for some inputs."""
    rng = np.random.default_rng(3)
    differ = 0
    for _ in range(300):
        v = _rand_floats(rng, 64, -40, 40)
        differ += _tree(v) != _seq(v)
    assert differ > 0
