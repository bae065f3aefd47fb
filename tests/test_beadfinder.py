"""QTrkFindBeads (cputrack/qtrk_c_api.cpp:33-58, cputrack/BeadFinder.cpp:163-245).

CPU: the oracle's restatement is pinned by what fixes it independently of its code: its
correlation map equals a brute-force circular convolution of the padded frame with the
wrapped sample (times W*H, kissfft's inverse being unnormalised), and on frames with planted
beads it finds every bead within a pixel of where it was planted, none at acceptance 1.0 (no
value exceeds the maximum), and none whose ROI crosses the frame's edge.  The reference's own
BeadFinder.cpp is not built: it links utils.cpp's JPEG writer.

GPU: the library's positions equal the oracle's, in order, on frames of several sizes, ROI
sizes and acceptances; the sample block is copied out; bad arguments raise.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth


def _planted(w, h, n, roi, seed, noise=2.0, margin=None):
    """n bead centres on a jittered grid at least 2*roi apart, kept a ROI away from the edges."""
    rng = np.random.default_rng(seed)
    m = roi if margin is None else margin
    cols = max(1, int((w - 2 * m) // (2 * roi)))
    cent = []
    for i in range(n):
        r, c = divmod(i, cols)
        cent.append((m + roi // 2 + c * 2 * roi + rng.uniform(-3, 3), m + roi // 2 + r * 2 * roi + rng.uniform(-3, 3)))
    cent = [p for p in cent if p[0] < w - m and p[1] < h - m]
    return cent, synth.bead_frame(w, h, cent, noise=noise, seed=seed)


def _corner(c, roi):
    return int(c[0]) - roi // 2, int(c[1]) - roi // 2


def test_xcor_map_is_the_circular_convolution(oracle):
    rng = np.random.default_rng(0)
    h, w, roi, sx, sy = 19, 27, 5, 3, 4
    img = rng.normal(10, 3, (h, w)).astype(np.float32)
    got = oracle.bead_xcor_map(img, sx, sy, roi)
    H, W = got.shape
    assert (H, W) == (32, 32)
    P = np.zeros((H, W))
    P[:h, :w] = img.astype(np.float64) - img.astype(np.float64).mean()
    smp = img[sy:sy + roi, sx:sx + roi].astype(np.float64)
    smp -= smp.mean()
    ref = np.zeros((H, W))
    for y in range(roi):
        for x in range(roi):
            # the sample pixel sits at ((y - roi) mod H, (x - roi) mod W): add it times P shifted there
            ref += smp[y, x] * np.roll(P, ((y - roi) % H, (x - roi) % W), axis=(0, 1))
    ref *= W * H
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


def test_oracle_finds_planted_beads(oracle):
    roi = 32
    cent, img = _planted(300, 220, 6, roi, seed=1)
    sx, sy = _corner(cent[0], roi)
    xy = oracle.find_beads(img, sx, sy, roi, 0.8, 0.5)
    assert len(xy) == len(cent)
    found = xy + roi // 2
    for c in cent:
        d = np.hypot(found[:, 0] - c[0], found[:, 1] - c[1]).min()
        assert d <= 1.5, (c, found)


def test_oracle_threshold_and_edges(oracle):
    roi = 32
    cent, img = _planted(300, 220, 6, roi, seed=2)
    sx, sy = _corner(cent[0], roi)
    assert len(oracle.find_beads(img, sx, sy, roi, 0.8, 1.0)) == 0       # nothing exceeds the maximum
    assert 1 <= len(oracle.find_beads(img, sx, sy, roi, 0.8, 0.999)) <= 2
    edge = cent + [(6.0, 110.0), (296.0, 60.0)]                           # ROIs crossing the left/right edges
    img2 = synth.bead_frame(300, 220, edge, noise=2.0, seed=2)
    xy = oracle.find_beads(img2, sx, sy, roi, 0.8, 0.5)
    assert len(xy) == len(cent)
    assert np.all(xy[:, 0] >= 0) and np.all(xy[:, 0] <= 300 - 1 - roi)


# ------------------------------------------------------------------ GPU

CASES = [  # (w, h, beads, roi, img_rel_dist, acceptance, noise)
    (300, 220, 6, 32, 0.8, 0.5, 2.0),
    (513, 257, 14, 33, 0.8, 0.6, 4.0),
    (640, 480, 40, 24, 1.0, 0.4, 6.0),
    (1600, 1200, 120, 40, 0.8, 0.45, 3.0),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}_roi{c[3]}" for c in CASES])
def test_find_beads_matches_oracle(gpu, oracle, case):
    w, h, n, roi, dist, acc, noise = case
    cent, img = _planted(w, h, n, roi, seed=w, noise=noise)
    sx, sy = _corner(cent[0], roi)
    ref = oracle.find_beads(img, sx, sy, roi, dist, acc)
    got = api.find_beads(img, (sx, sy), roi, dist, acc)
    print(f"{w}x{h} roi {roi}: {len(got)} found, oracle {len(ref)}, planted {len(cent)}", flush=True)
    assert np.array_equal(got, ref)
    assert len(got) >= len(cent) // 2


@pytest.mark.gpu
def test_find_beads_sample_and_errors(gpu):
    roi = 32
    cent, img = _planted(300, 220, 6, roi, seed=3)
    sx, sy = _corner(cent[1], roi)
    got, smp = api.find_beads(img, (sx, sy), roi, 0.8, 0.5, want_sample=True)
    assert np.array_equal(smp, img[sy:sy + roi, sx:sx + roi])
    assert len(got) == len(cent)
    with pytest.raises(api.QTrkError, match="sample ROI"):
        api.find_beads(img, (300 - roi + 1, 0), roi, 0.8, 0.5)
    with pytest.raises(api.QTrkError, match="larger than"):
        api.find_beads(np.zeros((1, 8200), np.float32), (0, 0), 1, 0.8, 0.5)
    assert len(api.find_beads(img, (sx, sy), roi, 0.8, 1.0)) == 0
