"""Committed fixtures: frozen oracle outputs and the reference's measured LUT (tests/golden/)."""
import os

import numpy as np
import pytest

from qtrk_amd import api

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    return np.load(os.path.join(HERE, "golden_plain64.npz"))


def test_oracle_reproduces_golden(oracle):
    g = _golden()
    cfg = oracle.config_from_bytes(g["cfg"].tobytes())
    res, best, _ = oracle.localize_trace(cfg, int(g["mode"]), g["luts"], g["rois_u16"].astype(np.float32), g["jobs"])
    for k in ("x", "y", "z", "comx", "comy", "imageMean", "error"):
        assert np.array_equal(res[k], g[k]), k
    assert np.array_equal(best, g["best"])


def _measured_lut_case(n=24, seed=8):
    """Beads rendered from the reference's measured LUT (10x.radialzlut, bead 0; 150 planes x 116 steps)."""
    lut = np.load(os.path.join(HERE, "ref10x_bead0.npy"))
    st = api.default_settings(width=80, height=80, zlut_minradius=1.0, zlut_radial_coverage=2.975,
                              qi_iterations=5)
    cfg = api.computed_config(st)
    assert cfg.zlut_radialsteps == 116
    rng = np.random.default_rng(seed)
    pos = np.stack([40 + rng.uniform(-3, 3, n), 40 + rng.uniform(-3, 3, n), rng.uniform(10, 140, n)], 1)
    pos = pos.astype(np.float32)
    img = api.generate_images_from_lut(lut[None], pos, np.zeros(n, np.int32), 80, 80, cfg.zlut_minradius,
                                       cfg.zlut_maxradius)
    rois = (500 + 300 * img).astype(np.float32)
    luts = lut[None].copy()
    mode = api.LT_QI | api.LT_LocalizeZ | api.LT_NormalizeProfile
    return st, cfg, luts, rois, api.make_jobs(n), pos, mode


def test_oracle_on_measured_reference_lut(oracle):
    st, cfg, luts, rois, jobs, pos, mode = _measured_lut_case()
    r, best, _ = oracle.localize_trace(oracle.config_from_bytes(bytes(cfg)), mode, luts, rois, jobs)
    assert np.abs(r["z"] - pos[:, 2]).max() < 0.5  # spline-smoothed rendering of a measured LUT: ~0.3 plane bias
    assert np.abs(r["x"] - pos[:, 0]).max() < 0.1 and np.abs(r["y"] - pos[:, 1]).max() < 0.1


@pytest.mark.gpu
def test_gpu_matches_golden(gpu):
    g = _golden()
    st = api.QTrkSettings.from_buffer_copy(g["cfg"].tobytes()[:84])
    trk = api.QTrkInstance(st)
    trk.set_radial_zlut(g["luts"])
    trk.set_localization_mode(int(g["mode"]))
    b = api.DeviceBatch(trk, len(g["jobs"]), api.QTrkFloat)
    b.enable_diagnostics(True)
    b.upload(g["rois_u16"].astype(np.float32), g["jobs"])
    b.run(len(g["jobs"]))
    b.sync()
    r = b.download(len(g["jobs"]))
    best, _ = b.diagnostics(len(g["jobs"]))
    b.free()
    trk.free()
    assert np.array_equal(best, g["best"])
    assert np.array_equal(r["error"], g["error"])
    for k in ("x", "y", "comx", "comy"):
        assert np.abs(r[k] - g[k]).max() <= 1e-3, k
    assert np.abs(r["z"] - g["z"]).max() <= 1e-3
    print("golden: identical x %d/24, z %d/24" % ((r["x"] == g["x"]).sum(), (r["z"] == g["z"]).sum()))


@pytest.mark.gpu
def test_gpu_on_measured_reference_lut(gpu, oracle):
    st, cfg, luts, rois, jobs, pos, mode = _measured_lut_case(n=64, seed=9)
    ref, rbest, _ = oracle.localize_trace(oracle.config_from_bytes(bytes(cfg)), mode, luts, rois, jobs)
    trk = api.QTrkInstance(st)
    trk.set_radial_zlut(luts)
    trk.set_localization_mode(mode)
    b = api.DeviceBatch(trk, len(jobs), api.QTrkFloat)
    b.enable_diagnostics(True)
    b.upload(rois, jobs)
    b.run(len(jobs))
    b.sync()
    r = b.download(len(jobs))
    best, _ = b.diagnostics(len(jobs))
    b.free()
    trk.free()
    assert np.array_equal(best, rbest)
    for k in ("x", "y"):
        assert np.abs(r[k] - ref[k]).max() <= 1e-3
    assert np.abs(r["z"] - ref["z"]).max() <= 1e-3
