"""Pins of the oracle's optional localization modes (LT_XCor1D, LT_ZLUTAlign,
LT_LocalizeZWeighted) and of the mixed-radix kissfft butterflies they reach.

Each check is against something the restatement does not share code with:
numpy's FFT, closed-form vertices of exact parabolas, or the true position of
a symmetric synthetic bead, chosen so that a plausible slip (dropped fabs, the
wrong LUT rows, a sign or index error, an off-by-one shift) fails.
"""
import numpy as np
import pytest

from qtrk_amd import synth

LT_XCOR1D, LT_QI, LT_ZLUTALIGN, LT_LOCALIZEZ, LT_NORMPROF, LT_ZWEIGHTED = 1, 2, 8, 16, 64, 512


@pytest.mark.parametrize("n", [3, 5, 9, 15, 25, 30, 60, 90, 96, 150, 186, 7, 14, 62, 98])
def test_fft_mixed_radix_matches_numpy(oracle, n):
    """kissfft's radix-3, radix-5 and generic butterflies (cputrack/kissfft.h:154-290)."""
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    tol = 2e-6 * n
    np.testing.assert_allclose(oracle.fft(x), np.fft.fft(x.astype(np.complex128)), atol=tol)
    np.testing.assert_allclose(oracle.fft(x, inverse=True), n * np.fft.ifft(x.astype(np.complex128)), atol=tol)


# ------------------------------------------------------- weighted compare

def test_weighted_compare_zero_weight_bin_is_ignored(oracle):
    """A bin whose values at planes zi and zi+1 agree gets weight |L[zi+1]-L[zi]| * r = 0, so an
    arbitrary mismatch there leaves the exact parabola of the other bins (cputrack/cpu_tracker.cpp:871-901)."""
    res, planes, k = 12, 60, 23
    minr, maxr = 1.0, 13.0
    rng = np.random.default_rng(5)
    base = (rng.integers(-32, 32, res) / 8).astype(np.float32)
    slope = (rng.integers(1, 8, res) / 16).astype(np.float32)
    j = np.arange(planes, dtype=np.float32)[:, None]
    lut = (base[None, :] + (j - k) * slope[None, :]).astype(np.float32)
    lut[:, 0] = 4.0 * (np.arange(planes) - k - 0.5) ** 2  # equal at planes k and k+1
    prof = lut[k].copy()
    prof[0] = 1000.0
    z, best, sc = oracle.lut_compare_weighted(prof, lut, minr, maxr, k + 0.4)
    assert best == k and abs(z - k) < 1e-5
    r = minr + (maxr - minr) * np.arange(res) / res
    c = float(np.sum(np.abs(slope[1:].astype(np.float64)) ** 3 * r[1:]))
    np.testing.assert_allclose(sc, -((np.arange(planes) - k) ** 2) * c, rtol=1e-12, atol=1e-9)
    # the unweighted compare is dominated by the mismatched bin
    assert oracle.lut_compare(prof, lut)[1] != k
    # the weights come from the rows around z_estim: from plane k+5, bin 0 has weight and dominates
    assert oracle.lut_compare_weighted(prof, lut, minr, maxr, k + 5.2)[1] != k


def test_weighted_compare_radius_factor_and_abs(oracle):
    """Two bins, at radii 2 and 6, matching planes 40 and 43 with slopes +s and -s: the score is
    -(s^3 2 (j-40)^2 + s^3 6 (j-43)^2) with its vertex at (2*40 + 6*43)/8 = 42.25."""
    res, planes = 8, 90
    minr, maxr = 1.0, 9.0           # bin radii 1 + i
    s = 0.5
    j = np.arange(planes, dtype=np.float64)
    lut = np.zeros((planes, res), np.float32)
    lut[:, 1] = (j - 40) * s
    lut[:, 5] = -(j - 43) * s
    prof = np.zeros(res, np.float32)
    z, best, sc = oracle.lut_compare_weighted(prof, lut, minr, maxr, 42.0)
    assert best == 42 and abs(z - 42.25) < 1e-5
    np.testing.assert_allclose(sc, -(s ** 3) * (2 * (j - 40) ** 2 + 6 * (j - 43) ** 2), rtol=1e-12)


# ------------------------------------------------------------------ XCor1D

def _spot(w, h, x, y, sigma=3.0):
    yy, xx = np.mgrid[0:h, 0:w]
    return (100 + 1000 * np.exp(-((xx - x) ** 2 + (yy - y) ** 2) / (2 * sigma ** 2))).astype(np.float32)


def test_xcor1d_symmetric_spot(oracle):
    """A profile symmetric about index m correlates with its reverse at shift 2m - N + 1, so one
    iteration moves the centre onto the spot (cputrack/cpu_tracker.cpp:190-270)."""
    truth = [(31.3, 32.6), (30.0, 33.0), (33.7, 29.2), (32.45, 31.55)]
    for x0, y0 in truth:
        img = _spot(64, 64, x0, y0)
        x1, y1, b = oracle.xcor1d(img, x0 - 1.4, y0 + 0.9, 1, 16, 32)
        assert not b
        assert abs(x1 - x0) < 0.05 and abs(y1 - y0) < 0.05, (x0, y0, x1, y1)
        x3, y3, _ = oracle.xcor1d(img, x0 - 1.4, y0 + 0.9, 3, 16, 32)
        assert abs(x3 - x0) < 0.03 and abs(y3 - y0) < 0.03
    # translation equivariance at a fixed start offset
    a = oracle.xcor1d(_spot(64, 64, 31.3, 32.6), 30.3, 33.6, 2, 16, 32)
    b = oracle.xcor1d(_spot(64, 64, 33.3, 29.6), 32.3, 30.6, 2, 16, 32)
    assert abs((b[0] - a[0]) - 2) < 2e-3 and abs((b[1] - a[1]) + 3) < 2e-3


def test_xcor1d_window_and_boundary(oracle):
    img = _spot(64, 64, 32.0, 32.0)
    # a profile longer than the ROI: boundary hit, start returned unchanged
    assert oracle.xcor1d(img, 30.5, 33.25, 2, 32, 128) == (30.5, 33.25, True)
    # a window crossing the ROI edge raises the boundary flag but still localizes
    x, y, b = oracle.xcor1d(_spot(64, 64, 20.0, 32.0), 19.0, 32.0, 2, 16, 48)
    assert b and abs(x - 20.0) < 0.1


def test_xcor1d_localize_mode(oracle):
    """LT_XCor1D replaces QI in the job (cputrack/QueuedCPUTracker.cpp:299-303)."""
    cfg = oracle.config(width=64, height=64, xc1_profileLength=32, xc1_profileWidth=16, xc1_iterations=2)
    truth = np.array([(31.3, 32.6), (33.7, 29.2)])
    rois = np.stack([_spot(64, 64, x, y) for x, y in truth])
    jobs = np.zeros(2, dtype=[("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                              ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
    r = oracle.localize(cfg, LT_XCOR1D | LT_QI, None, rois, jobs)
    for i in range(2):
        ref = oracle.xcor1d(rois[i], float(r["comx"][i]), float(r["comy"][i]), 2, 16, 32)
        assert (r["x"][i], r["y"][i]) == (np.float32(ref[0]), np.float32(ref[1]))
    assert np.abs(r["x"] - truth[:, 0]).max() < 0.05 and np.abs(r["y"] - truth[:, 1]).max() < 0.05


# ---------------------------------------------------------- QuadrantAlign

def _lut_bead(oracle, x, y, z, w=64, res=40, minr=0.0, maxr=30.0):
    lut = synth.analytic_lut(50, res, minr, maxr, 8.0)
    img = (1000 + 500 * oracle.generate_image_from_lut(w, w, lut, minr, maxr, x, y, z)).astype(np.float32)
    return lut, img


def test_quadrant_align_moves_onto_bead(oracle):
    """QuadrantAlign correlates [rev(q1+q2), q0+q3] against the bead's unmirrored LUT profile
    (cputrack/cpu_tracker.cpp:409-512).  A centre offset d shifts the quadrant sums by (2/pi) d, the
    correlation peaks at that shift (half QI's autocorrelation shift) and the pi/4 scale shared with QI
    turns it into a step of d/2: every call halves the offset, in the right direction."""
    for x0, y0, z0 in [(32.3, 31.7, 20.0), (30.6, 33.2, 12.4), (33.1, 32.8, 31.7)]:
        lut, img = _lut_bead(oracle, x0, y0, z0)
        x, y = x0 + 0.8, y0 - 0.6
        ex, ey = [x - x0], [y - y0]
        for _ in range(8):
            x, y, b = oracle.quadrant_align(img, float(img.mean()), lut, 0.0, 30.0, 20, x, y, z0)
            assert not b
            ex.append(x - x0)
            ey.append(y - y0)
        rx, ry = np.array(ex[1:4]) / np.array(ex[:3]), np.array(ey[1:4]) / np.array(ey[:3])
        assert np.all((rx > 0.4) & (rx < 0.65)) and np.all((ry > 0.4) & (ry < 0.65)), (rx, ry)
        assert abs(ex[-1]) < 0.01 and abs(ey[-1]) < 0.01, (ex, ey)
    # LUT rows are interpolated at z: a bead between rows still aligns
    lut, img = _lut_bead(oracle, 31.4, 32.5, 20.5)
    x, y = 32.0, 32.0
    for _ in range(8):
        x, y, _ = oracle.quadrant_align(img, float(img.mean()), lut, 0.0, 30.0, 20, x, y, 20.5)
    assert abs(x - 31.4) < 0.01 and abs(y - 32.5) < 0.01


def test_zlut_align_job(oracle):
    """LT_ZLUTAlign: 5 profile/align/compare passes; the boundary flag ends cleared
    (cputrack/QueuedCPUTracker.cpp:319-348, cputrack/cpu_tracker.cpp:484)."""
    res = 40
    cfg = oracle.config(width=64, height=64, qi_iterations=2, zlut_minradius=0.0, zlut_roi_coverage=30.0 / 32,
                        zlut_radial_coverage=res / 30.0)
    assert cfg.zlut_radialsteps == res and cfg.zlut_maxradius == 30.0
    lut = synth.analytic_lut(50, res, 0.0, 30.0, 8.0)
    luts = np.stack([synth.normalize_profiles(lut)])
    truth = [(32.3, 31.7, 20.2), (30.6, 33.2, 12.6), (25.0, 40.0, 31.3)]
    rois = np.stack([(1000 + 500 * oracle.generate_image_from_lut(64, 64, lut, 0.0, 30.0, *t)).astype(np.float32)
                     for t in truth])
    jobs = np.zeros(3, dtype=[("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                              ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
    plain = oracle.localize(cfg, LT_QI | LT_LOCALIZEZ | LT_NORMPROF, luts, rois, jobs)
    aligned = oracle.localize(cfg, LT_QI | LT_LOCALIZEZ | LT_NORMPROF | LT_ZLUTALIGN, luts, rois, jobs)
    t = np.array(truth)
    e_plain = np.hypot(plain["x"] - t[:, 0], plain["y"] - t[:, 1])
    e_al = np.hypot(aligned["x"] - t[:, 0], aligned["y"] - t[:, 1])
    print("QI error", e_plain, "ZLUTAlign error", e_al)
    # 4 alignment passes after 2 QI iterations: each pass halves the offset
    assert np.all(e_al < e_plain / 4) and e_al[:2].max() < 0.015 and e_al[2] < 0.06, e_al
    assert plain["error"][2] == 1 and np.all(aligned["error"] == 0)
    # z: the rendered images are blurred by pixel integration against the analytic LUT, which biases
    # the plane by under one; the alignment must not move it further
    assert np.abs(aligned["z"] - t[:, 2]).max() < 1.0
    # the weighted compare runs after the passes on the last profile; x/y are unchanged
    wz = oracle.localize(cfg, LT_QI | LT_LOCALIZEZ | LT_NORMPROF | LT_ZLUTALIGN | LT_ZWEIGHTED, luts, rois, jobs)
    assert np.array_equal(wz["x"], aligned["x"]) and np.abs(wz["z"] - t[:, 2]).max() < 1.5


# ------------------------------------------------------------ Z-bias table

def test_zbias_apply_closed_forms(oracle):
    """ZLUTBiasCorrection (cputrack/QueuedTracker.cpp:115-140): 4 fixed-point steps pos = z - bias(pos)."""
    planes, w = 100, 50
    assert oracle.zbias_apply(np.zeros(w), planes, 37.3) == np.float32(37.3)
    # a constant bias is subtracted exactly
    assert oracle.zbias_apply(np.full(w, 0.75), planes, 37.25) == np.float32(36.5)
    # bias(pos) = c pos (the table samples planes j * planes / w): pos_4 = z (1 - c + c^2 - c^3 + c^4)
    c = 0.125
    row = (c * np.arange(w) * planes / w).astype(np.float32)
    for z in (10.0, 41.5, 80.25):
        exp = z * (1 - c + c ** 2 - c ** 3 + c ** 4)
        assert abs(oracle.zbias_apply(row, planes, z) - exp) < 1e-4 * z
    # outside the table the end entries hold
    ramp = np.arange(w, dtype=np.float32)
    assert oracle.zbias_apply(ramp, planes, -5.0) == -5.0 - ramp[0]
    assert oracle.zbias_apply(ramp, planes, 500.0) == 500.0 - ramp[-1]


def _bias_cfg(oracle, w):
    """The localization ComputeZBiasCorrection applies to its renders, as a ProcessJob configuration:
    COM without background correction, 2 QI iterations, the normalised profile."""
    cfg = oracle.config_from_bytes(bytes(w["cfg"]))
    cfg.qi_iterations = 2
    cfg.com_bgcorrection = 0.0
    return cfg


def test_zbias_table_composition_and_round_trip(oracle):
    """ComputeZBiasCorrection (cputrack/QueuedTracker.cpp:143-189): table = z measured on the bead's
    own LUT render at the centre minus the true z.  The same measurement through ProcessJob (2 QI
    iterations, radial weights = the Z-LUT bin window) reproduces every entry, and inverting the
    table with ZLUTBiasCorrection returns the true z, on and between the table's planes."""
    w = synth.make_workload("plain64", n=1)
    lut = w["luts"]
    planes, bias_planes = lut.shape[1], 200
    table = oracle.compute_zbias_table(oracle.config_from_bytes(bytes(w["cfg"])), lut, bias_planes)
    assert table.shape == (1, bias_planes) and np.all(np.isfinite(table))
    cfg = _bias_cfg(oracle, w)
    zwin = oracle.bin_window(cfg.zlut_radialsteps).astype(np.float32)
    jobs = np.zeros(1, dtype=[("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                              ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
    errs = []
    for p in (17, 64, 101, 150, 183):
        zp = np.float32(np.float32(p) / np.float32(bias_planes) * np.float32(planes))
        img = oracle.generate_image_from_lut(64, 64, lut[0], cfg.zlut_minradius, cfg.zlut_maxradius, 32.0, 32.0,
                                             float(zp)).astype(np.float32)
        r = oracle.localize(cfg, synth.MODE_QI_Z, lut, img[None], jobs, zcmp=zwin)
        assert np.float32(r["z"][0] - zp) == table[0, p]
        errs.append(abs(oracle.zbias_apply(table[0], planes, float(r["z"][0])) - zp))
        # between table planes: the correction interpolates the table
        zq = float(zp) + 0.23
        img = oracle.generate_image_from_lut(64, 64, lut[0], cfg.zlut_minradius, cfg.zlut_maxradius, 32.0, 32.0,
                                             zq).astype(np.float32)
        rq = oracle.localize(cfg, synth.MODE_QI_Z, lut, img[None], jobs, zcmp=zwin)
        errs.append(abs(oracle.zbias_apply(table[0], planes, float(rq["z"][0])) - zq))
    print("corrected |z - truth|", np.round(errs, 5))
    assert max(errs[0::2]) < 1e-3 and max(errs[1::2]) < 0.02


# ------------------------------------------------------------ LT_FourierLUT

@pytest.mark.parametrize("w,h", [(64, 64), (100, 100), (96, 96), (62, 58)])
def test_power_spectrum_matches_numpy(oracle, w, h):
    """FFT2D::Apply (cputrack/cpu_tracker.cpp:920-952) is |FFT2|^2 with the zero frequency at (w/2, h/2):
    numpy's fftshift for even and odd sizes."""
    rng = np.random.default_rng(w * h)
    img = rng.uniform(0, 100, size=(h, w)).astype(np.float32)
    ps = oracle.power_spectrum_2d(img)
    ref = np.fft.fftshift(np.abs(np.fft.fft2(img.astype(np.float64))) ** 2)
    np.testing.assert_allclose(ps, ref, rtol=1e-4, atol=1e-5 * ref.max())
    assert ps[h // 2, w // 2] == np.float32(img.astype(np.float64).sum() ** 2) or \
        abs(ps[h // 2, w // 2] / img.astype(np.float64).sum() ** 2 - 1) < 1e-6  # DC term


def test_fourier_profile_composition_and_shift_invariance(oracle):
    """FourierRadialProfile (:954-963): log of the radial profile of the power spectrum about (w/2, h/2),
    radii 1 .. 0.3 w, fallback 0, no normalisation; a circular shift of the image leaves it unchanged."""
    lut = synth.analytic_lut(60, 40, 0.0, 30.0, 8.0)
    img = (500 + 1000 * oracle.generate_image_from_lut(64, 64, lut, 0.0, 30.0, 31.6, 32.3, 25.0)).astype(np.float32)
    prof = oracle.fourier_radial_profile(img, 40, 120)
    ps = oracle.power_spectrum_2d(img)
    exp = np.log(oracle.radial_profile(ps, 40, 120, 1.0, np.float32(64 * np.float32(0.3)), 32.0, 32.0, 0.0, False))
    np.testing.assert_allclose(prof, exp, rtol=1e-6)
    rolled = oracle.fourier_radial_profile(np.roll(img, (5, -7), axis=(0, 1)), 40, 120)
    np.testing.assert_allclose(rolled, prof, atol=2e-4)


def test_fourier_lut_localizes_z_independent_of_position(oracle):
    """BUILDLUT_FOURIER (cputrack/QueuedCPUTracker.cpp:630-637) + LT_FourierLUT: a Fourier LUT built from
    centred beads gives the same plane for beads moved by whole pixels (the power spectrum ignores the
    phase; a sub-pixel move resamples the bead and shifts z by up to a quarter plane)."""
    cfg = oracle.config(width=64, height=64, zlut_minradius=0.0, zlut_roi_coverage=30.0 / 32, zlut_radial_coverage=1.0)
    res, planes = cfg.zlut_radialsteps, 60
    lut = synth.analytic_lut(planes, 40, 0.0, 30.0, 8.0)
    flut = np.zeros((1, planes, res), np.float32)
    for p in range(planes):
        img = (500 + 1000 * oracle.generate_image_from_lut(64, 64, lut, 0.0, 30.0, 32.0, 32.0, float(p)))
        oracle.build_lut(cfg, img[None].astype(np.float32), p, flut, fourier=True)
    jobs = np.zeros(4, dtype=[("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                              ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
    truth = [(32.0, 32.0, 20.3), (33.0, 31.0, 20.3), (32.0, 32.0, 41.7), (30.0, 34.0, 41.7)]
    rois = np.stack([(500 + 1000 * oracle.generate_image_from_lut(64, 64, lut, 0.0, 30.0, *t)).astype(np.float32)
                     for t in truth])
    r = oracle.localize(cfg, LT_QI | LT_LOCALIZEZ | 256, flut, rois, jobs)
    print("Fourier-LUT z", r["z"])
    assert abs(r["z"][0] - 20.3) < 0.05 and abs(r["z"][2] - 41.7) < 0.05
    assert abs(r["z"][1] - r["z"][0]) < 2e-3 and abs(r["z"][3] - r["z"][2]) < 2e-3


def test_image_lut_differences_closed_form(oracle):
    """FinalizeLUT's image LUT differences (cputrack/QueuedCPUTracker.cpp:661-696) on planes that are
    quadratic in z, v = a + b z + c z^2 per pixel: the central difference is b + 2 c z and the second
    difference 2 c (exact here: small integers), the first and last planes equal their neighbours'."""
    rng = np.random.default_rng(7)
    beads, planes, h, w = 2, 6, 3, 4
    a, b, c = (rng.integers(-20, 20, (beads, 1, h, w)).astype(np.float32) for _ in range(3))
    z = np.arange(planes, dtype=np.float32)[None, :, None, None]
    lut = (a + b * z + c * z * z).astype(np.float32)
    dz, dz2 = oracle.image_lut_diff(lut)
    zi = z[:, 1:-1]
    assert np.array_equal(dz[:, 1:-1], np.broadcast_to(b + 2 * c * zi, dz[:, 1:-1].shape))
    assert np.array_equal(dz2[:, 1:-1], np.broadcast_to(2 * c, dz2[:, 1:-1].shape))
    assert np.array_equal(dz[:, 0], dz[:, 1]) and np.array_equal(dz[:, -1], dz[:, -2])
    assert np.array_equal(dz2[:, 0], dz2[:, 1]) and np.array_equal(dz2[:, -1], dz2[:, -2])
