"""Derives the polynomial coefficients of qmath.h (single-precision exp and erf).

Run `python qtrk_amd/csrc/qmath_fit.py` to print the coefficient block that
qmath.h holds.  Method: each polynomial is the minimax fit (Lawson's
iteratively reweighted least squares, C. L. Lawson 1961; see also Rice &
Usow 1968) of the target function on a dense grid of its interval, computed in
double precision from Python's math.exp / math.erf / math.erfc, then rounded to
float.  The intervals and the argument reductions are described in qmath.h.

    exp(r) = 1 + r + r^2 * E(r),               |r| <= ln(2)/2 (+ margin)
    erf(a) = a + a * S(a^2),                   0 <= a < 1
    erfc(a) = exp(-a^2) * R1(a - 1.5),         1 <= a < 2
    erfc(a) = exp(-a^2) * R2(a - 3),           2 <= a < 4
"""
import math

import numpy as np

F32 = np.float32

# (name, degree) of each polynomial; the degrees are the smallest whose fit error is
# well below float resolution on the interval.
DEGREES = {"E": 5, "S": 6, "R1": 8, "R2": 9}


def lawson(t, y, deg, relative):
    A = np.vander(t, deg + 1, increasing=True)
    scale = 1.0 / np.abs(y) if relative else np.ones_like(y)
    wt = np.full_like(t, 1.0 / len(t))
    c = None
    for _ in range(80):
        sw = np.sqrt(wt)
        c, *_ = np.linalg.lstsq(A * (sw * scale)[:, None], y * sw * scale, rcond=None)
        e = np.abs((A @ c - y) * scale)
        wt = wt * e
        wt /= wt.sum()
    return c, float(np.max(np.abs((A @ c - y) * scale)))


def fits():
    out = {}
    # E(r) = (exp(r) - 1 - r) / r^2 on |r| <= 0.3466 (ln 2 / 2 = 0.34657...), with margin
    r = np.linspace(-0.35, 0.35, 40001)
    r = r[np.abs(r) > 1e-6]
    y = np.array([(math.expm1(v) - v) / (v * v) for v in r])
    out["E"] = lawson(r, y, DEGREES["E"], relative=True)
    # S(s) = erf(sqrt(s)) / sqrt(s) - 1 on s in [0, 1]
    s = np.linspace(0.0, 1.0, 40001)
    a = np.sqrt(s)
    y = np.array([math.erf(v) / v - 1.0 if v > 0 else 2.0 / math.sqrt(math.pi) - 1.0 for v in a])
    out["S"] = lawson(s, y, DEGREES["S"], relative=False)
    # R(a) = erfc(a) exp(a^2), centred variables
    for name, lo, hi, mid in (("R1", 1.0, 2.0, 1.5), ("R2", 2.0, 4.0, 3.0)):
        a = np.linspace(lo, hi, 40001)
        y = np.array([math.erfc(v) * math.exp(v * v) for v in a])
        out[name] = lawson(a - mid, y, DEGREES[name], relative=True)
    return out


def hexf(v):
    """The float nearest v as a C hexadecimal float literal."""
    h = float(F32(v)).hex()  # e.g. -0x1.5555560000000p-3
    mant, exp = h.split("p")
    mant = mant.rstrip("0")
    if mant.endswith("."):
        mant = mant[:-1]
    return f"{mant}p{exp}f"


def cody_waite_ln2():
    """ln 2 split as HI + LO, HI with 16 significant bits so that k * HI is exact for |k| < 256."""
    hi = round(math.log(2.0) * 2**16) / 2**16
    return hi, float(F32(math.log(2.0) - hi))


def coefficient_block():
    lines = []
    for name, (c, err) in fits().items():
        lines.append(f"/* {name}: degree {len(c) - 1}, fit error {err:.2e} */")
        vals = ", ".join(hexf(v) for v in c)
        lines.append(f"#define QM_{name} {vals}")
    hi, lo = cody_waite_ln2()
    lines.append(f"#define QM_LN2_HI {hexf(hi)}")
    lines.append(f"#define QM_LN2_LO {hexf(lo)}")
    lines.append(f"#define QM_LOG2E {hexf(1.0 / math.log(2.0))}")
    return "\n".join(lines)


if __name__ == "__main__":
    print(coefficient_block())
