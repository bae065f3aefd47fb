"""Parity of the optional localization modes with the CPU oracle, through the C ABI.

LT_XCor1D (cputrack/cpu_tracker.cpp:190-270), LT_ZLUTAlign's five profile /
QuadrantAlign / compare passes (cputrack/QueuedCPUTracker.cpp:319-344,
cputrack/cpu_tracker.cpp:409-512) and LT_LocalizeZWeighted
(cputrack/cpu_tracker.cpp:871-901), alone and combined, on power-of-two and
mixed-radix transform lengths: 2^k, 4*4*3, 4*3*5, 4*2*7 (generic radix 7),
2*3*3*5 and 2*3*31 (generic radix 31).

Bars as for the common modes: x/y within 1e-3 px, z within 1e-3 planes, the
boundary flag, best plane and radial profile identical.  The kernels follow
the oracle's float rounding operation for operation (the FFTs reproduce
kissfft's butterflies), so the counts of identical values are printed too.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth

from test_gpu_parity import _assert_parity, _batch, _oracle, _report, _tracker

pytestmark = pytest.mark.gpu

Z = api.LT_LocalizeZ | api.LT_NormalizeProfile
XCOR_Z = api.LT_XCor1D | Z
ALIGN = api.LT_QI | Z | api.LT_ZLUTAlign
WEIGHTED = api.LT_QI | Z | api.LT_LocalizeZWeighted
ALIGN_WEIGHTED = ALIGN | api.LT_LocalizeZWeighted


def _parity(oracle, w, mode, tag):
    print(f"[{tag}] mode {mode:#x} n {len(w['jobs'])} xc {w['cfg'].xc1_profileLength} "
          f"zsteps {w['cfg'].zlut_radialsteps}", flush=True)
    ref, rbest, rprof = _oracle(oracle, w, mode=mode)
    trk = _tracker(w, mode=mode)
    try:
        got, best, prof = _batch(trk, w["rois"], w["jobs"])
    finally:
        trk.free()
    _report(tag, got, ref)
    _assert_parity(got, ref, best, rbest)
    if mode & api.LT_LocalizeZ:
        np.testing.assert_allclose(prof, rprof, atol=1e-5)
    return got, ref


@pytest.mark.parametrize("length", [32, 48, 60, 56])
def test_xcor1d_parity(gpu, oracle, length):
    w = synth.make_workload("plain64", n=384, xc1_profileLength=length, xc1_profileWidth=16, xc1_iterations=3)
    got, ref = _parity(oracle, w, XCOR_Z, f"xcor {length}")
    # XCor1D localises the beads (the COM start is off by up to a pixel)
    e = np.hypot(got["x"] - w["truth"][:, 0], got["y"] - w["truth"][:, 1])
    print(f"[xcor {length}] median |xy - truth| {np.median(e):.4f} px", flush=True)
    assert np.median(e) < 0.1


def test_xcor1d_profile_longer_than_roi(gpu, oracle):
    """The default 128-sample profile does not fit a 64x64 ROI: the COM is kept, boundary set."""
    w = synth.make_workload("plain64", n=64)
    got, _ = _parity(oracle, w, api.LT_XCor1D | api.LT_QI, "xcor 128 on 64px")
    assert np.array_equal(got["x"], got["comx"]) and np.all(got["error"] == 1)


@pytest.mark.parametrize("mode", [ALIGN, WEIGHTED, ALIGN_WEIGHTED], ids=["align", "weighted", "align+weighted"])
def test_zlut_align_and_weighted_realtime100(gpu, oracle, mode):
    w = synth.make_workload("realtime100", n=512)
    got, ref = _parity(oracle, w, mode, f"realtime100 {mode:#x}")
    if mode & api.LT_ZLUTAlign:
        assert np.all(got["error"] == 0)  # QuadrantAlign clears the flag


@pytest.mark.parametrize("coverage,steps", [(1.25, 32), (1.75, 45), (93.5 / 26, 93)], ids=["n64", "n90", "n186"])
def test_zlut_align_mixed_radix(gpu, oracle, coverage, steps):
    w = synth.make_workload("plain64", n=256, zlut_radial_coverage=coverage)
    assert w["cfg"].zlut_radialsteps == steps
    _parity(oracle, w, ALIGN_WEIGHTED, f"align {2 * steps}")


def test_xcor1d_with_zlut_align(gpu, oracle):
    """XCor1D start, then the alignment passes (QI's direction table sets the quadrant sampling)."""
    w = synth.make_workload("plain64", n=256, xc1_profileLength=48, xc1_profileWidth=16, xc1_iterations=2)
    _parity(oracle, w, api.LT_XCor1D | Z | api.LT_ZLUTAlign | api.LT_LocalizeZWeighted, "xcor+align")


def test_optional_mode_edges(gpu, oracle):
    """Constant, zero, ramp and edge-spot ROIs and beads beyond the LUT through the optional modes."""
    w = synth.make_workload("plain64", n=32, xc1_profileLength=48, xc1_profileWidth=16)
    rois = w["rois"].copy()
    rois[0] = 123.0
    rois[1] = 0.0
    rois[2][:, :] = np.arange(64, dtype=np.float32)[None, :]
    rois[3][30:34, 1:3] = 50000
    jobs = w["jobs"].copy()
    jobs["zlutIndex"][4:8] = 7  # one bead only: no compare, no alignment
    for mode in (XCOR_Z, ALIGN_WEIGHTED):
        ref, rbest, _ = _oracle(oracle, w, mode=mode, rois=rois, jobs=jobs)
        trk = _tracker(w, mode=mode)
        try:
            got, best, _ = _batch(trk, rois, jobs)
        finally:
            trk.free()
        _report(f"edges {mode:#x}", got, ref)
        _assert_parity(got, ref, best, rbest)
        assert np.all(best[4:8] == -1) and np.all(got["z"][4:8] == 0)


def test_optional_mode_validation(gpu):
    w = synth.make_workload("plain64", n=2, xc1_profileLength=600)
    trk = api.QTrkInstance(w["settings"])
    try:
        with pytest.raises(api.QTrkError, match="xc1_profileLength"):
            trk.set_localization_mode(api.LT_XCor1D)
        with pytest.raises(api.QTrkError, match="not implemented"):
            trk.set_localization_mode(api.LT_QI | 32)  # a bit LocalizeModeEnum does not define
        trk.set_localization_mode(api.LT_QI | api.LT_ZLUTAlign | Z)
    finally:
        trk.free()
    w = synth.make_workload("plain64", n=2, zlut_radial_coverage=10.0)
    trk = api.QTrkInstance(w["settings"])
    try:
        with pytest.raises(api.QTrkError, match="zlut_radialsteps"):
            trk.set_localization_mode(ALIGN)
        trk.set_localization_mode(WEIGHTED)
    finally:
        trk.free()


def test_pipeline_rejects_per_roi_compare_modes(gpu):
    w = synth.batching_tracker(32, 16, 4, 40, 32, np.random.default_rng(3), render=False)
    trk = api.QTrkInstance(w["settings"])
    p = api.Pipeline()
    try:
        trk.set_radial_zlut(w["lut"])
        trk.set_localization_mode(ALIGN)
        with pytest.raises(api.QTrkError, match="LT_ZLUTAlign"):
            p.add(trk, 16, api.QTrkU16)
    finally:
        p.free()
        trk.free()


# ------------------------------------------------------------ Z-bias table

def _bias_workloads():
    """plain64 (1 bead, 100 planes, 64x64) and 4 beads of realtime100 (200 planes, 100x100)."""
    a = synth.make_workload("plain64", n=64)
    b = synth.make_workload("realtime100", n=64)
    b["luts"] = b["luts"][:4].copy()
    b["jobs"]["zlutIndex"] %= 4
    return [a, b]


@pytest.mark.parametrize("spline,ovs", [(True, 4), (False, 2)], ids=["spline-ovs4", "linear-ovs2"])
def test_zbias_table_matches_oracle(gpu, oracle, spline, ovs):
    """QueuedTracker::ComputeZBiasCorrection (cputrack/QueuedTracker.cpp:143-189) on the GPU: the LUT
    renders and their localization against the oracle's restatement, entry by entry."""
    for w in _bias_workloads():
        planes = w["luts"].shape[1]
        bias_planes = 2 * planes + 7  # table planes fall between LUT planes too
        print(f"[zbias] {w['name']} beads {w['luts'].shape[0]} bias_planes {bias_planes}", flush=True)
        ref = oracle.compute_zbias_table(oracle.config_from_bytes(bytes(w["cfg"])), w["luts"], bias_planes, ovs,
                                         spline)
        trk = _tracker(w)
        try:
            got = trk.compute_zlut_bias_table(bias_planes, ovs, spline)
            assert np.array_equal(trk.get_zlut_bias_table(), got)
        finally:
            trk.free()
        same = int((got == ref).sum())
        print(f"[zbias] {same}/{got.size} identical, max|d| {np.abs(got - ref).max():.2e} planes, "
              f"bias range [{ref.min():.3f}, {ref.max():.3f}]", flush=True)
        assert np.abs(got - ref).max() <= 1e-3


def test_zbias_correction_applied_to_results(gpu, oracle):
    """With a table installed every result's z is ZLUTBiasCorrection of the uncorrected z
    (cputrack/QueuedCPUTracker.cpp:346-347), through the queue and the device-resident batch."""
    w = synth.make_workload("realtime100", n=256)
    trk = _tracker(w)
    try:
        plain, _, _ = _batch(trk, w["rois"], w["jobs"], diag=False)
        rng = np.random.default_rng(9)
        table = rng.uniform(-1.5, 1.5, size=(w["luts"].shape[0], 57)).astype(np.float32)
        trk.set_zlut_bias_table(table)
        assert np.array_equal(trk.get_zlut_bias_table(), table)
        batch, _, _ = _batch(trk, w["rois"], w["jobs"], diag=False)
        queued = trk.localize_all(w["rois"], w["jobs"])
        planes = w["luts"].shape[1]
        exp = np.array([oracle.zbias_apply(table[b], planes, float(z))
                        for b, z in zip(plain["zlutIndex"], plain["z"])], np.float32)
        assert np.array_equal(batch["z"], exp)
        assert np.array_equal(queued["z"], exp)  # localize_all returns job order
        for k in ("x", "y", "error"):
            assert np.array_equal(batch[k], plain[k])
        with pytest.raises(api.QTrkError, match="numbeads"):
            trk.set_zlut_bias_table(table[:3])
        trk.set_zlut_bias_table(None)
        assert trk.get_zlut_bias_table() is None
        again, _, _ = _batch(trk, w["rois"], w["jobs"], diag=False)
        assert np.array_equal(again["z"], plain["z"])
        # a new LUT drops the table (cputrack/QueuedCPUTracker.cpp:368-369)
        trk.set_zlut_bias_table(table)
        trk.set_radial_zlut(w["luts"])
        assert trk.get_zlut_bias_table() is None
    finally:
        trk.free()


# ------------------------------------------------------------ LT_FourierLUT

@pytest.mark.parametrize("size", [64, 100, 96], ids=["64px", "100px", "96px"])
def test_fourier_lut_build_and_localize(gpu, oracle, size):
    """BUILDLUT_FOURIER and LT_FourierLUT (cputrack/QueuedCPUTracker.cpp:324-326, 630-637;
    cputrack/cpu_tracker.cpp:920-963): the 2D power spectrum is kissfft's in rows and columns
    (lengths 64 = 4^3, 100 = 4*5*5, 96 = 4*4*2*3), its profile is taken in log.  The device logf and
    glibc's may differ in the last place, so profiles and z are compared within tolerance."""
    w = synth.make_workload("plain64", n=16, width=size, height=size)
    cfg = w["cfg"]
    o_cfg = oracle.config_from_bytes(bytes(cfg))
    planes, count = 40, 2
    lut_src = w["render_luts"][0]
    trk = api.QTrkInstance(w["settings"])
    try:
        trk.set_radial_zlut((count, planes))
        trk.begin_lut(api.BUILDLUT_FOURIER | api.BUILDLUT_NORMALIZE)  # the Fourier profile ignores NORMALIZE
        ref = np.zeros((count, planes, cfg.zlut_radialsteps), np.float32)
        c = size / 2
        for p in range(planes):
            posn = np.stack([np.full(count, c), np.full(count, c), np.full(count, 2.0 + 2.4 * p)], 1).astype(np.float32)
            imgs = api.generate_images_from_lut(lut_src[None], posn, np.zeros(count, np.int32), size, size,
                                                cfg.zlut_minradius, cfg.zlut_maxradius)
            imgs = (1000 + 800 * imgs).astype(np.float32)
            trk.build_lut(imgs, p)
            oracle.build_lut(o_cfg, imgs, p, ref, fourier=True)
        got = trk.get_radial_zlut()
        d = np.abs(got - ref)
        print(f"[fourier {size}] LUT identical {(d == 0).mean():.3f}, max|d| {d.max():.2e}", flush=True)
        assert d.max() <= 1e-5 * max(1.0, np.abs(ref).max())
        # localize with the oracle's LUT in both, so only the localization is compared
        trk.set_radial_zlut(ref)
        mode = api.LT_QI | api.LT_LocalizeZ | api.LT_FourierLUT
        trk.set_localization_mode(mode)
        jobs = w["jobs"].copy()
        jobs["zlutIndex"] = np.arange(len(jobs)) % count
        rois = w["rois"]
        got_r, best, prof = _batch(trk, rois, jobs)
        ref_r, rbest, rprof = _oracle(oracle, dict(w, luts=ref), mode=mode, jobs=jobs)
        _report(f"fourier {size}", got_r, ref_r)
        _assert_parity(got_r, ref_r)
        np.testing.assert_allclose(prof, rprof, rtol=1e-5, atol=1e-5)
        assert np.mean(best == rbest) >= 0.9  # a last-place logf difference can tip a near tie
    finally:
        trk.free()


def test_fourier_lut_validation(gpu):
    w = synth.make_workload("plain64", n=2)
    trk = api.QTrkInstance(w["settings"])
    p = api.Pipeline()
    try:
        trk.set_radial_zlut(w["luts"])
        trk.set_localization_mode(api.LT_QI | api.LT_LocalizeZ | api.LT_FourierLUT)
        with pytest.raises(api.QTrkError, match="LT_FourierLUT"):
            p.add(trk, 2, api.QTrkU16)
        # BUILDLUT_IMAGELUT with no image LUT set: the reference's loops over a 0 x 0 image LUT do
        # nothing (cputrack/QueuedCPUTracker.cpp:606-624), and the radial build goes on
        trk.begin_lut(api.BUILDLUT_IMAGELUT)
        trk.build_lut(w["rois"][:1], 0)
        assert trk.get_image_zlut().size == 0
    finally:
        p.free()
        trk.free()
