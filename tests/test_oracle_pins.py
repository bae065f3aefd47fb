"""Pins the CPU oracle (oracle/qtrk_oracle.c) before it is trusted as the checker.

Each test checks one restated reference function against something the oracle
does not share code with: the reference's own headers compiled unmodified
(oracle/_ref), numpy's FFT, closed forms, or invariants chosen so that a
plausible slip (sign, index, dropped term, swapped operand) fails.
"""
import re

import numpy as np
import pytest

REF_QT = "/root/reference/cputrack/QueuedTracker.h"


def _ref_weights(name):
    """The fit weights as written in the reference header (cputrack/QueuedTracker.h:106,116)."""
    txt = open(REF_QT).read()
    m = re.search(r"^#define\s+%s\s+\{([^}]*)\}" % name, txt, re.M)  # skips the //#define variants
    return np.array([float(v.strip().rstrip("f")) for v in m.group(1).split(",")], np.float32)


# ---------------------------------------------------------------- a1: config

def test_config_appendix_c(oracle):
    c1 = oracle.config(width=64, height=64, qi_minradius=2, qi_roi_coverage=0.875, qi_radial_coverage=1.25,
                       qi_angular_coverage=1.46, qi_angstep_factor=1, qi_iterations=5, zlut_minradius=2,
                       zlut_roi_coverage=0.875, zlut_radial_coverage=1.25)
    assert (c1.qi_maxradius, c1.qi_radialsteps, c1.qi_angstepspq) == (28.0, 32, 64)
    assert (c1.zlut_maxradius, c1.zlut_radialsteps, c1.zlut_angularsteps) == (28.0, 32, 123)
    c2 = oracle.config(width=100, height=100, qi_minradius=1, qi_roi_coverage=1, qi_radial_coverage=1.5,
                       qi_angular_coverage=1.635, qi_iterations=5, zlut_minradius=1, zlut_roi_coverage=1,
                       zlut_radial_coverage=1.31, zlut_angular_coverage=0.7)
    assert (c2.qi_maxradius, c2.qi_radialsteps, c2.qi_angstepspq) == (50.0, 64, 128)
    assert (c2.zlut_radialsteps, c2.zlut_angularsteps) == (64, 219)
    d = oracle.config()  # reference defaults, 150x150 (cputrack/qtrk_c_api.h:60-81)
    assert (d.zlut_maxradius, d.zlut_radialsteps, d.zlut_angularsteps) == (75.0, 222, 329)
    assert (d.qi_radialsteps, d.qi_angstepspq) == (256, 82)


def test_nearest_pow2(oracle):
    # ties go to the lower power (strict < at cputrack/utils.cpp:649)
    cases = {1: 1, 2: 2, 3: 2, 5: 4, 6: 4, 7: 8, 12: 8, 13: 16, 73: 64, 96: 64, 97: 128}
    for v, e in cases.items():
        assert oracle.L.oq_nearest_pow2(v) == e, v


# ------------------------------------------------------------- a5: bilinear

def test_interpolate_affine_exact(oracle):
    h, w = 20, 30
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    img = (3 + 2 * xx - 5 * yy).astype(np.float32)
    rng = np.random.default_rng(1)
    for _ in range(200):
        x = rng.integers(-127, (w - 1) * 128) / 128.0
        y = rng.integers(-127, (h - 1) * 128) / 128.0
        v, out = oracle.interpolate(img, x, y)
        assert not out
        assert v == np.float32(3 + 2 * x - 5 * y)


def test_interpolate_outside_rule(oracle):
    img = np.arange(12, dtype=np.float32).reshape(3, 4)
    # x in (-1, 0) is inside and extrapolates (trunc toward zero)
    v, out = oracle.interpolate(img, -0.5, 0.0)
    assert not out and v == np.float32(-0.5)
    assert oracle.interpolate(img, -1.0, 0.0)[1]
    assert oracle.interpolate(img, 3.0, 0.0)[1]          # rx+1 >= w
    assert not oracle.interpolate(img, 2.999, 1.0)[1]
    assert oracle.interpolate(img, 0.0, 2.0)[1]          # ry+1 >= h
    assert oracle.interpolate(img, float("nan"), 1.0)[1]
    assert oracle.interpolate(img, 1e12, 1.0)[1]


def test_radial_bin_window_closed_form(oracle):
    n = 64
    x = np.arange(n) / n
    ref = np.sqrt((1 - np.exp(-(1 - x) ** 2 / 0.05)) * (1 - np.exp(-x ** 2 / 0.01)) * x * 2)
    np.testing.assert_allclose(oracle.bin_window(n), ref, atol=2e-6)


# ------------------------------------------------------------------ a7: FFT

@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 6, 10, 12])
def test_fft_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    tol = 1e-5 * n
    np.testing.assert_allclose(oracle.fft(x), np.fft.fft(x.astype(np.complex128)), atol=tol)
    # inverse is unnormalised (cputrack/kissfft.h)
    np.testing.assert_allclose(oracle.fft(x, inverse=True), n * np.fft.ifft(x.astype(np.complex128)), atol=tol)


def test_fft_leaf_order_is_digit_reversal(oracle):
    for n in (8, 32, 64, 128, 256):
        p = oracle.fft_leaf_order(n)
        assert sorted(p) == list(range(n))
        # kissfft factors 4s first, then one 2; leaves read the mixed-radix digit reversal
        radices, m = [], n
        while m % 4 == 0:
            radices.append(4)
            m //= 4
        if m == 2:
            radices.append(2)
        exp = []
        for j in range(n):
            rem, src, stride = j, 0, 1
            sizes = [int(np.prod(radices[i + 1:])) for i in range(len(radices))]
            for r, sz in zip(radices, sizes):
                q, rem = divmod(rem, sz)
                src += q * stride
                stride *= r
            exp.append(src)
        assert list(p) == exp


# ---------------------------------------------------------- a8: peak fits

def test_lsqfit_bit_exact_vs_reference_header(oracle, ref_headers):
    rng = np.random.default_rng(5)
    for n in (3, 4, 5, 6, 7):
        for _ in range(50):
            x = np.arange(n, dtype=np.float32) - rng.integers(0, n)
            y = rng.standard_normal(n).astype(np.float32) * 100
            wts = rng.uniform(0.1, 1, n).astype(np.float32)
            assert np.array_equal(oracle.lsqfit(x, y, wts), ref_headers.lsqfit(x, y, wts))
            assert np.array_equal(oracle.lsqfit(x, y, wts, double=True),
                                  ref_headers.lsqfit(x, y, wts, double=True))


def test_max_interp_bit_exact_vs_reference_header(oracle, ref_headers):
    qiw = _ref_weights("QI_LSQFIT_WEIGHTS")
    zw = _ref_weights("ZLUT_LSQFIT_WEIGHTS").astype(np.float64)
    rng = np.random.default_rng(6)
    for length in (3, 5, 8, 64, 128, 200):
        for _ in range(40):
            d = rng.standard_normal(length).astype(np.float32)
            k = rng.integers(0, length)
            d[k] += 5  # clear peak anywhere, including the edges
            got, _ = oracle.max_interp_qi(d)
            assert np.float32(got) == np.float32(ref_headers.max_interp_f7(d, qiw))
            dd = -np.cumsum(np.abs(rng.standard_normal(length)))[::-1] * 1e3
            dd = np.roll(dd, rng.integers(0, length))
            got, _ = oracle.max_interp_zlut(dd)
            assert got == ref_headers.max_interp_d7(dd, zw)
    flat = np.zeros(16, np.float64)  # a == 0 -> the integer argmax (|a| < 1e-9 switch)
    assert oracle.max_interp_zlut(flat)[0] == 0.0


def test_peak_fit_closed_form(oracle):
    # an exact parabola has its vertex recovered whatever the weights or window
    for v in (10.25, 0.5, 62.75, 40.0):
        x = np.arange(64, dtype=np.float64)
        y = -3.0 * (x - v) ** 2 + 7.0
        z, imax = oracle.max_interp_zlut(y)
        assert abs(z - v) < 1e-9, (v, z)
        # a constant shift changes nothing
        assert abs(oracle.max_interp_zlut(y + 1234.5)[0] - v) < 1e-9


# ------------------------------------------------------------ a14: B-spline

def test_bspline_weights(oracle, ref_headers):
    for t in np.linspace(0, 1, 33, dtype=np.float32):
        w = oracle.bspline_weights(float(t))
        assert np.array_equal(w, ref_headers.bspline_weights(float(t)))
        assert abs(w.sum() - 1) < 1e-6
        # uniform cubic B-spline: w0 = (1-t)^3/6, w3 = t^3/6
        assert abs(w[0] - (1 - t) ** 3 / 6) < 1e-7 and abs(w[3] - t ** 3 / 6) < 1e-7


def test_generate_image_constant_and_radial(oracle):
    lut = np.full((10, 16), 0.75, np.float32)
    img = oracle.generate_image_from_lut(20, 20, lut, 1.0, 9.0, 10.3, 9.6, 4.5)
    assert np.all(img == np.float32(0.75))
    # LUT linear in bin index -> pixel value = mean over sub-samples of (r - rmin) * steps/(R - rmin)
    lut = np.tile(np.arange(16, dtype=np.float32), (10, 1))
    img = oracle.generate_image_from_lut(20, 20, lut, 1.0, 9.0, 10.0, 10.0, 5.0, ovs=4)
    o = (np.arange(4) + 0.5) / 4 - 0.5
    X, Y = np.meshgrid(14 + o - 10, 11 + o - 10)
    r = np.clip((np.hypot(X, Y) - 1.0) * 2.0, 0, 14)
    assert abs(img[11, 14] - r.mean()) < 1e-5


# ----------------------------------------------------------------- a4: COM

def _com(oracle, img, bg=0.0):
    cfg = oracle.config(width=img.shape[1], height=img.shape[0], com_bgcorrection=bg)
    jobs = np.zeros(1, dtype=[("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                              ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
    r = oracle.localize(cfg, 0, None, img[None].astype(np.float32), jobs)
    return r[0]


def test_com_symmetric_single_pixel_constant(oracle):
    yy, xx = np.mgrid[0:64, 0:64]
    img = (np.abs(xx - 31.5) * 3 + np.abs(yy - 31.5) ** 2).astype(np.float32)  # symmetric about 31.5
    r = _com(oracle, img)
    assert r["comx"] == 31.5 and r["comy"] == 31.5
    one = np.zeros((40, 48), np.float32)
    one[20, 10] = 1000
    r = _com(oracle, one, bg=1.0)  # background weights are zeroed by |v-mean| - sigma
    assert (r["comx"], r["comy"]) == (10.0, 20.0)
    r = _com(oracle, np.full((30, 30), 5, np.float32))
    assert (r["comx"], r["comy"]) == (15.0, 15.0)  # 0/0 COM falls back to the ROI centre
    assert r["imageMean"] == 5.0


# ------------------------------------------------------------------ a6: QI

def _bead(oracle, x, y, z=20.0, w=64):
    from qtrk_amd import synth
    lut = synth.analytic_lut(50, 40, 0.0, 30.0, 8.0)
    return (1000 + 500 * oracle.generate_image_from_lut(w, w, lut, 0.0, 30.0, x, y, z)).astype(np.float32)


def _qi(oracle, rois, iters):
    cfg = oracle.config(width=64, height=64, qi_iterations=iters, qi_minradius=2, qi_roi_coverage=0.8,
                        qi_radial_coverage=1.5, qi_angular_coverage=1.0)
    jobs = np.zeros(len(rois), dtype=[("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                                      ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
    return oracle.localize(cfg, 2, None, rois, jobs)  # LT_QI


def test_qi_converges_to_known_position(oracle):
    truth = np.array([(32.3, 31.7), (30.9, 33.4), (32.0, 32.0), (33.6, 30.2), (32.5, 32.5), (31.25, 32.75)])
    rois = np.stack([_bead(oracle, x, y) for x, y in truth])
    # each QI iteration shrinks the error of the COM start towards the true centre
    errs = []
    for it in (1, 3, 12):
        r = _qi(oracle, rois, it)
        errs.append(np.hypot(r["x"] - truth[:, 0], r["y"] - truth[:, 1]))
    assert np.all(errs[1] < errs[0]) and np.all(errs[2] < errs[1])
    assert errs[2].max() < 0.015
    # translation equivariance at convergence (the COM start is not equivariant: ROI edges)
    r = _qi(oracle, rois, 40)
    r2 = _qi(oracle, np.stack([_bead(oracle, x + 2, y - 3) for x, y in truth]), 40)
    np.testing.assert_allclose(r2["x"] - r["x"], 2, atol=2e-3)
    np.testing.assert_allclose(r2["y"] - r["y"], -3, atol=2e-3)


# --------------------------------------------------------- a9: radial profile

def test_radial_profile_analytic_and_normalised(oracle):
    yy, xx = np.mgrid[0:64, 0:64]
    r = np.hypot(xx - 32.0, yy - 32.0)
    img = (100 + 50 * np.cos(r / 3.0)).astype(np.float32)
    prof = oracle.radial_profile(img, 32, 150, 2.0, 26.0, 32.0, 32.0, 0.0, False)
    rr = 2.0 + np.arange(32) * 24.0 / 32
    np.testing.assert_allclose(prof, 100 + 50 * np.cos(rr / 3.0), atol=1.5)
    n = oracle.radial_profile(img, 32, 150, 2.0, 26.0, 32.0, 32.0, 0.0, True)
    assert abs(n.astype(np.float64).mean()) < 1e-6
    assert abs(np.sqrt((n.astype(np.float64) ** 2).mean()) - 1) < 1e-5
    # too few in-bounds samples (<= 4) -> the supplied mean
    far = oracle.radial_profile(img, 4, 8, 40.0, 60.0, 32.0, 32.0, -7.0, False)
    assert np.all(far == -7.0)


# ----------------------------------------------------------- a10: LUT match

def test_lut_compare_exact_parabola(oracle):
    rng = np.random.default_rng(3)
    res, planes, z0 = 32, 100, 37.25
    prof = (rng.integers(-64, 64, res) / 16).astype(np.float32)
    v = np.full(res, 0.125, np.float32)
    lut = np.stack([prof - (k - z0) * v for k in range(planes)]).astype(np.float32)
    z, best, sc = oracle.lut_compare(prof, lut)
    assert best == 37 and abs(z - z0) < 1e-5
    np.testing.assert_allclose(sc, -((np.arange(planes) - z0) ** 2) * res / 64, rtol=0, atol=1e-9)
    # radial weights multiply the difference before squaring: zero weight hides a component
    lut2 = np.stack([prof, prof]).astype(np.float32)
    lut2[0, 0] += 4
    lut2[1, 1] += 1
    w = np.ones(res, np.float32)
    assert oracle.lut_compare(prof, lut2)[1] == 1
    w[0] = 0
    assert oracle.lut_compare(prof, lut2, w)[1] == 0


# --------------------------------------------------------------- a13: MLE

def test_gauss2d_mle_noise_free(oracle):
    img = oracle.gaussian_spot(16, 16, 7.3, 8.6, 1.5, 1000.0, 10.0)
    out = oracle.gauss2d_mle(img, float(img.mean()), 7.0, 8.0, 20, 1.5)
    assert abs(out[0] - 7.3) < 2e-3 and abs(out[1] - 8.6) < 2e-3


# ----------------------------------------------- whole pipeline + FP64 oracle

def test_fp64_oracle_agrees(oracle, oracle64):
    from qtrk_amd import synth
    w = synth.make_workload("plain64", n=64)
    r32 = oracle.localize(oracle.config_from_bytes(bytes(w["cfg"])), w["mode"], w["luts"], w["rois"], w["jobs"])
    r64 = oracle64.localize(oracle64.config_from_bytes(bytes(w["cfg"])), w["mode"], w["luts"], w["rois"],
                            w["jobs"])
    for k in "xy":
        assert np.abs(r32[k] - r64[k]).max() < 1e-3
    assert np.abs(r32["z"] - r64["z"]).max() < 1e-3
    # and both localise the simulated beads
    assert np.abs(r32["x"] - w["truth"][:, 0]).max() < 0.15


def test_queued_cpu_tracker_restatement(oracle):
    from qtrk_amd import synth
    w = synth.make_workload("plain64", n=120)
    cfg = oracle.config_from_bytes(bytes(w["cfg"]))
    single = oracle.localize(cfg, w["mode"], w["luts"], w["rois"], w["jobs"])
    t, q = oracle.queued_run(cfg, w["mode"], w["luts"], w["rois"], w["jobs"], 4)
    assert t > 0 and len(q) == len(single)
    q = q[np.argsort(q["frame"])]
    assert np.array_equal(q, single)


# --------------------------------------------- BuildLUT composition (f2)

def _plain64_cfg(oracle, **kw):
    s = dict(width=64, height=64, qi_minradius=2, qi_roi_coverage=0.875, qi_radial_coverage=1.25,
             qi_angular_coverage=1.46, qi_angstep_factor=1, qi_iterations=5, zlut_minradius=2,
             zlut_roi_coverage=0.875, zlut_radial_coverage=1.25, zlut_angular_coverage=0.7)
    s.update(kw)
    return oracle.config(**s)


def test_build_lut_known_position_closed_form(oracle):
    """BuildLUT with known positions (cputrack/QueuedCPUTracker.cpp:580-645) on the quadratic bowl
    f = A + B |p - c|^2: the angular mean about a known position k is A + B (r^2 + |k - c|^2) (the
    cross term averages out over the uniform directions), and bilinear sampling adds t(1 - t) per
    axis, so each bin lies in A + B (r_i^2 + d^2 + [0, 0.5]) with r_i = minr + i (maxr - minr) / n.
    The row is added into its plane (twice -> doubled) and other planes stay zero."""
    cfg = _plain64_cfg(oracle)
    nb = cfg.zlut_radialsteps
    A, B, c = 100.0, 0.5, (31.3, 32.6)
    y, x = np.mgrid[0:64, 0:64].astype(np.float64)
    img = (A + B * ((x - c[0]) ** 2 + (y - c[1]) ** 2)).astype(np.float32)
    k = np.array([[28.0, 35.0]], np.float32)
    zlut = np.zeros((1, 3, nb), np.float32)
    oracle.build_lut(cfg, img[None], 1, zlut, known_pos=k)
    r = cfg.zlut_minradius + np.arange(nb) * (cfg.zlut_maxradius - cfg.zlut_minradius) / nb
    d2 = float(((k[0] - np.array(c)) ** 2).sum())
    lo, hi = A + B * (r * r + d2), A + B * (r * r + d2 + 0.5)
    row = zlut[0, 1].astype(np.float64)
    tol = 2e-6 * hi
    assert np.all(row >= lo - tol) and np.all(row <= hi + tol)
    assert np.all(zlut[0, 0] == 0) and np.all(zlut[0, 2] == 0)
    first = zlut[0, 1].copy()
    oracle.build_lut(cfg, img[None], 1, zlut, known_pos=k)
    assert np.array_equal(zlut[0, 1], first + first)


def test_build_lut_position_is_com_then_qi_without_window(oracle):
    """Without known positions BuildLUT profiles at ComputeQI(ComputeMeanAndCOM()) with neither the
    configured COM background correction nor the QI radial bin window (cputrack/
    QueuedCPUTracker.cpp:598-603); ProcessJob applies both (:291, :306).  On an asymmetric image the
    window moves QI's result, and the row equals the profile at the unwindowed position only."""
    cfg = _plain64_cfg(oracle, com_bgcorrection=1.5)
    nb, na = cfg.zlut_radialsteps, cfg.zlut_angularsteps
    y, x = np.mgrid[0:64, 0:64].astype(np.float64)
    img = (50 + 4000 * np.exp(-((x - 32.4) ** 2 + (y - 31.7) ** 2) / 32)
           + 1500 * np.exp(-((x - 50.0) ** 2 + (y - 36.0) ** 2) / 6)).astype(np.float32)
    pw = oracle.com_qi(cfg, img, 0.0, window=True)
    pn = oracle.com_qi(cfg, img, 0.0, window=False)
    pb = oracle.com_qi(cfg, img, 1.5, window=False)
    assert abs(pw[0] - pn[0]) + abs(pw[1] - pn[1]) > 1e-2
    zlut = np.zeros((1, 1, nb), np.float32)
    oracle.build_lut(cfg, img[None], 0, zlut)
    prof = lambda p: oracle.radial_profile(img, nb, na, cfg.zlut_minradius, cfg.zlut_maxradius, p[0], p[1],
                                           float(img.mean()), False)
    assert np.array_equal(zlut[0, 0], prof(pn))
    assert not np.array_equal(zlut[0, 0], prof(pw))
    if pb[:2] != pn[:2]:
        assert not np.array_equal(zlut[0, 0], prof(pb))


# ------------------------------------- sample-count thresholds (a6, a9)

def _affine(w, h, a=10.0, b=0.7, c=-0.3):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    return (a + b * x + c * y).astype(np.float32), (a, b, c)


def test_quadrant_profile_takes_four_samples(oracle):
    """ComputeQuadrantProfile averages when at least MIN_RADPROFILE_SMP_COUNT = 4 samples are inside
    (n >= 4, cputrack/cpu_tracker.cpp:670) and falls back to the mean below.  Quadrant 0, 4 angles
    (pi/2 (j + 0.5)/4), radius 10 on a 40x40 affine image: centre x 28.5 keeps all 4 (average,
    exact for an affine image), 29.5 puts the first angle's sample at x = 39.31 >= W - 1 (3 left:
    the fallback)."""
    img, (a, b, c) = _affine(40, 40)
    ang = np.float32(0.5 * np.float64(np.float32(3.141593)) * (np.arange(4) + 0.5) / 4).astype(np.float64)
    for cx, inside in ((28.5, 4), (29.5, 3)):
        xs = cx + np.cos(ang) * 10
        assert int((xs < 39).sum()) == inside
        p = oracle.quadrant_profile(img, -999.0, 4, 4, 0, 1, 10.0, 20.0, cx, 20.0)
        if inside == 4:
            want = a + b * np.mean(xs) + c * np.mean(20.0 + np.sin(ang) * 10)
            assert abs(p[0] - want) < 1e-3
        else:
            assert p[0] == np.float32(-999.0)


def test_radial_profile_needs_more_than_four_samples(oracle):
    """::ComputeRadialProfile averages only when more than 4 samples are inside (n > 4,
    cputrack/utils.cpp:313): 5 directions 2 pi j/5 at radius 10 on a 40x40 affine image, centre x
    28.5 (all 5 inside: the average) and 29.5 (the j = 0 sample at x = 39.5 is outside, 4 left: the
    fallback)."""
    img, (a, b, c) = _affine(40, 40)
    ang = (2 * np.float32(3.141593) * np.arange(5) / np.float32(5)).astype(np.float64)
    for cx, inside in ((28.5, 5), (29.5, 4)):
        xs = cx + np.cos(ang) * 10
        assert int((xs < 39).sum()) == inside
        p = oracle.radial_profile(img, 1, 5, 10.0, 20.0, cx, 20.0, -999.0, False)
        if inside == 5:
            want = a + b * np.mean(xs) + c * np.mean(20.0 + np.sin(ang) * 10)
            assert abs(p[0] - want) < 1e-3
        else:
            assert p[0] == np.float32(-999.0)


# ------------------------------------ QI angular index at factor != 1 (a6)

def test_quadrant_profile_angular_index_truncates_the_step(oracle):
    """With fewer angular samples than table directions (qi_angstep_factor > 1), sample a uses
    direction (int)angstepf * a, the step truncated before the multiply (cputrack/cpu_tracker.cpp:
    655-657), not (int)(angstepf * a).  64 directions, 17 samples (64 / 1.3^5 = 17.2): step 3, so
    directions 0..48 -- the other rule reaches 60.  Exact on an affine image."""
    img, (a, b, c) = _affine(100, 100, b=0.9, c=0.4)
    A, n, r, cx, cy = 64, 17, 30.0, 50.0, 50.0
    ang = np.float32(0.5 * np.float64(np.float32(3.141593)) * (np.arange(A) + 0.5) / A).astype(np.float64)
    step = 64 / np.float32(17)
    val = lambda j: a + b * np.mean(cx + np.cos(ang[j]) * r) + c * np.mean(cy + np.sin(ang[j]) * r)
    want = val(int(step) * np.arange(n))
    other = val((step * np.arange(n)).astype(int))
    assert abs(want - other) > 1.0
    p = oracle.quadrant_profile(img, 0.0, A, n, 0, 1, r, r + 1, cx, cy)
    assert abs(p[0] - want) < 1e-3
