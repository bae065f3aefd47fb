"""The image LUT: QueuedTracker::SetImageZLUT / GetImageZLUT / GetImageZLUTSize and BuildLUT's
BUILDLUT_IMAGELUT accumulation (cputrack/QueuedCPUTracker.cpp:519-570, 606-624), exported as
QTrkSetImageZLUT / QTrkGetImageZLUT / QTrkGetImageZLUTSize.

CPU: the oracle's accumulation is pinned by the closed form at integer positions, where every
bilinear sample is a pixel (the image LUT plane is the ROI's crop around the position, zero
where the sample's 2 x 2 neighbourhood leaves the ROI) and by the half-pixel case, where each
sample is the float midpoint of two pixels.

GPU: the tracker's image LUT equals the oracle's bit for bit with known positions (no mean, the
same bilinear operations), and within float rounding of the mean with COM -> QI positions;
set/get round trips, the radial LUT set alongside, and the size checks.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth


def _images(count, seed, dtype):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:64, 0:64]
    out = []
    for i in range(count):
        cx, cy = 32 + rng.uniform(-2, 2), 32 + rng.uniform(-2, 2)
        r = np.hypot(x - cx, y - cy)
        im = 1000 + 600 * np.cos(r / 2.5) * np.exp(-(r / 12) ** 2) + rng.normal(0, 5, r.shape)
        out.append(im)
    out = np.array(out)
    return (np.round(out).astype(np.uint16) if dtype == np.uint16 else out.astype(np.float32))


def test_oracle_image_lut_integer_positions_are_crops(oracle):
    w = synth.make_workload("plain64", n=4)
    cfg = oracle.config_from_bytes(bytes(w["cfg"]))
    imgs = _images(3, 1, np.float32)
    ih, iw = 9, 12
    lut = np.zeros((3, 2, ih, iw), np.float32)
    kp = np.array([[30, 31], [5, 33], [60, 2]], np.float32)   # the last two reach the ROI's edges
    oracle.build_image_lut(cfg, imgs, 1, lut, known_pos=kp)
    assert np.all(lut[:, 0] == 0)
    for i, (px, py) in enumerate(kp.astype(int)):
        for yy in range(ih):
            for xx in range(iw):
                sx, sy = px - iw // 2 + xx, py - ih // 2 + yy
                inside = 0 <= sx and 0 <= sy and sx + 1 < 64 and sy + 1 < 64
                assert lut[i, 1, yy, xx] == (imgs[i, sy, sx] if inside else 0.0), (i, xx, yy)


def test_oracle_image_lut_half_pixel_is_the_midpoint(oracle):
    w = synth.make_workload("plain64", n=4)
    cfg = oracle.config_from_bytes(bytes(w["cfg"]))
    imgs = _images(1, 2, np.float32)
    lut = np.zeros((1, 1, 4, 4), np.float32)
    oracle.build_image_lut(cfg, imgs, 0, lut, known_pos=np.array([[20.5, 20]], np.float32))
    a = imgs[0, 18:22, 18:22]
    b = imgs[0, 18:22, 19:23]
    assert np.array_equal(lut[0, 0], a + (b - a) * np.float32(0.5))


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.uint16, np.float32], ids=["u16", "f32"])
@pytest.mark.parametrize("known", [True, False], ids=["known", "qi"])
def test_image_lut_build_matches_oracle(gpu, oracle, dtype, known):
    w = synth.make_workload("plain64", n=4)
    cfg = w["cfg"]
    o_cfg = oracle.config_from_bytes(bytes(cfg))
    count, planes, ih, iw = 4, 3, 40, 48
    trk = api.QTrkInstance(w["settings"])
    assert trk.set_image_zlut((count, planes, ih, iw))
    trk.begin_lut(api.BUILDLUT_IMAGELUT)
    ref = np.zeros((count, planes, ih, iw), np.float32)
    rng = np.random.default_rng(5)
    for p in (0, 2, 2):
        imgs = _images(count, 10 + p, dtype)
        kp = (np.stack([32 + rng.uniform(-3, 3, count), 32 + rng.uniform(-3, 3, count)], 1).astype(np.float32)
              if known else None)
        trk.build_lut(imgs, p, kp)
        oracle.build_image_lut(o_cfg, imgs, p, ref, known_pos=kp)
    got = trk.get_image_zlut()
    trk.free()
    assert got.shape == ref.shape
    print(f"{dtype.__name__} known={known}: max|d| {np.abs(got - ref).max():.3g}", flush=True)
    if known:
        assert np.array_equal(got, ref)
    else:
        assert np.abs(got - ref).max() <= 1e-3 * np.abs(ref).max()
    assert np.all(got[:, 1] == 0)


@pytest.mark.gpu
def test_image_lut_set_get_and_checks(gpu):
    w = synth.make_workload("plain64", n=4)
    cfg = w["cfg"]
    rng = np.random.default_rng(3)
    src = rng.normal(size=(2, 5, 6, 7)).astype(np.float32)
    radial = rng.normal(size=(2, 5, cfg.zlut_radialsteps)).astype(np.float32)
    trk = api.QTrkInstance(w["settings"])
    assert trk.set_image_zlut(src, radial)
    assert np.array_equal(trk.get_image_zlut(), src)
    assert np.array_equal(trk.get_radial_zlut(), radial)
    trk.begin_lut(api.BUILDLUT_IMAGELUT)
    with pytest.raises(api.QTrkError, match="fewer beads or planes"):
        trk.build_lut(_images(2, 1, np.float32), 5)               # plane 5 of 5 planes
    assert trk.set_image_zlut((0, 0, 0, 0), None)
    assert trk.get_image_zlut().size == 0
    trk.free()


@pytest.mark.gpu
def test_image_lut_derivatives_after_finalize(gpu, oracle):
    """QTrkFinalizeLUT computes the image LUT's first and second differences across planes, as the
    reference's FinalizeLUT does (cputrack/QueuedCPUTracker.cpp:661-696): equal to the oracle's bit
    for bit; none before FinalizeLUT, none after the image LUT is set again, none for fewer than 3
    planes."""
    w = synth.make_workload("plain64", n=4)
    rng = np.random.default_rng(9)
    lut = rng.normal(0, 100, (3, 7, 9, 11)).astype(np.float32)
    trk = api.QTrkInstance(w["settings"])
    try:
        trk.set_image_zlut(lut)
        assert trk.get_image_zlut_derivatives() is None
        trk.finalize_lut()
        got = trk.get_image_zlut_derivatives()
        assert got is not None
        ref = oracle.image_lut_diff(lut)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        trk.set_image_zlut(lut)
        assert trk.get_image_zlut_derivatives() is None
        trk.set_image_zlut(lut[:, :2])
        trk.finalize_lut()
        assert trk.get_image_zlut_derivatives() is None
    finally:
        trk.free()
