"""Grouped radial Z-LUT match (QTrkLUTMatch*) against the oracle's LUTProfileCompare.

The reference scores each profile against every plane in float difference /
double accumulation and fits the best plane's neighbourhood in FP64
(cputrack/cpu_tracker.cpp:791-868).  Every precision of the GPU match must give
the oracle's best plane bit-exactly and z within 1e-3 planes; the exact path and
the BF16/FP16 paths (approximate GEMM scores, top-8 rescored in the reference
arithmetic) give the oracle's z bit-exactly whenever the best plane agrees.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth

PRECS = {"fp32": api.MATCH_FP32, "bf16": api.MATCH_BF16, "fp16": api.MATCH_FP16}


def make_case(beads, planes, res, frames, seed, noise=0.1, span=None, zmax=None):
    """Per-bead analytic LUTs (normalised rows) and profiles interpolated between
    planes at random z (below zmax if given) with Gaussian noise, normalised like measured
    profiles.  The LUT family's ring period and width change by as much over `span` planes
    (default: all of them) as realtime100's over its 200."""
    rng = np.random.default_rng(seed)
    scale = 200.0 / max(span or planes, 2)
    luts = np.stack([synth.normalize_profiles(synth.analytic_lut(planes, res, 1.0, 50.0, rng.uniform(7, 10),
                                                                 dperiod=0.04 * scale, dwidth=0.05 * scale))
                     for _ in range(beads)])
    lo = min(2.0, (planes - 1) / 2)
    z = rng.uniform(lo, (zmax or planes) - 1 - lo, size=(beads, frames))
    i = np.clip(np.floor(z).astype(int), 0, planes - 2) if planes > 1 else np.zeros_like(z, int)
    t = (z - i)[..., None] if planes > 1 else np.zeros(z.shape + (1,))
    rows = luts[np.arange(beads)[:, None], i]
    nxt = luts[np.arange(beads)[:, None], np.minimum(i + 1, planes - 1)]
    prof = (1 - t) * rows + t * nxt + noise * rng.standard_normal((beads, frames, res))
    return luts, synth.normalize_profiles(prof), z


def oracle_match(o, luts, prof, weights=None):
    beads, frames, _ = prof.shape
    z = np.zeros((beads, frames), np.float32)
    best = np.zeros((beads, frames), np.int32)
    for b in range(beads):
        for f in range(frames):
            zz, bb, _ = o.lut_compare(prof[b, f], luts[b], weights)
            z[b, f], best[b, f] = zz, bb
    return z, best


SHAPES = [
    (8, 500, 128, 64),   # the precision workload's LUT and profile tile at 8 beads
    (4, 200, 64, 16),    # realtime100's LUT, 16 frames per bead
    (3, 37, 50, 33),     # ragged: planes, radial steps and profiles off every tile size
    (2, 130, 100, 150),  # more than one 64-profile tile per bead
    (2, 5, 16, 7),       # fewer planes than candidates
    (2, 700, 160, 40),   # 22 plane tiles (two tile-mean tiles), 160 radial steps (16 k-steps)
    # 66 plane tiles: five tile-mean jobs, one wavefront takes two.  Neighbouring planes as far
    # apart as at 500 planes (the family spread over 2,100 differs less per plane than BF16
    # resolves) and beads below plane 1,000 (the rings of the upper planes grow past the radial
    # range and their profiles flatten); every candidate must then lie near the best plane, which
    # a wrong tile-mean term in the later tiles would break.
    (1, 2100, 48, 20, 500, 1000),
]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("weighted", [False, True], ids=["unweighted", "weighted"])
def test_match_parity(oracle, oracle64, gpu, shape, weighted):
    beads, planes, res, frames = shape[:4]
    local = len(shape) > 4
    luts, prof, ztrue = make_case(beads, planes, res, frames, seed=sum(shape) + weighted,
                                  span=shape[4] if local else None, zmax=shape[5] if local else None)
    w = np.linspace(0.5, 1.5, res).astype(np.float32) if weighted else None
    zo, bo = oracle_match(oracle, luts, prof, w)
    z64, _ = oracle_match(oracle64, luts, prof, w)
    m = api.LUTMatch(luts, w)
    try:
        for name, prec in PRECS.items():
            print(f"case {shape} weighted={weighted} precision={name}", flush=True)
            z, best, cand = m.match(prof, prec)
            assert np.array_equal(best, bo), f"{name}: {np.sum(best != bo)} best planes differ"
            assert np.array_equal(z, zo), f"{name}: max |dz| {np.max(np.abs(z - zo))}"
            if prec != api.MATCH_FP32:
                valid = cand >= 0
                assert np.all(valid.sum(-1) == min(8, planes))
                assert np.all(cand[valid] < planes)
                srt = np.sort(np.where(valid, cand, -1 - np.arange(8)), axis=-1)
                assert np.all(np.diff(srt, axis=-1) != 0), "duplicate candidates"
                assert np.all(np.any(cand == bo[..., None], axis=-1))
                if local:
                    assert np.all(np.abs(cand - bo[..., None]) <= 16), "candidates far from the best plane"
            print(f"  {name}: |z - FP64 oracle| max {np.max(np.abs(z - z64)):.3g} planes, "
                  f"rms vs truth {np.sqrt(np.mean((z - ztrue) ** 2)):.3g}")
    finally:
        m.free()


@pytest.mark.gpu
def test_match_nan_profile(oracle, gpu):
    luts, prof, _ = make_case(2, 60, 32, 5, seed=7)
    prof[1, 2, 10] = np.nan
    zo, bo = oracle_match(oracle, luts, prof)
    assert bo[1, 2] == 0 and np.isnan(zo[1, 2])
    m = api.LUTMatch(luts)
    try:
        for name, prec in PRECS.items():
            z, best, _ = m.match(prof, prec)
            assert np.array_equal(best, bo), name
            assert np.array_equal(z, zo, equal_nan=True), name
    finally:
        m.free()


@pytest.mark.gpu
def test_match_rejects_bad_input(gpu):
    luts, prof, _ = make_case(1, 20, 300, 4, seed=3)
    m = api.LUTMatch(luts)
    try:
        m.upload(prof)
        with pytest.raises(api.QTrkError, match="256 radial steps"):
            m.run(api.MATCH_BF16)
        with pytest.raises(api.QTrkError, match="unknown precision"):
            m.run(7)
        m.run(api.MATCH_FP32)
        m.sync()
    finally:
        m.free()
    with pytest.raises(api.QTrkError):
        api.LUTMatch(np.zeros((0, 4, 4), np.float32))
