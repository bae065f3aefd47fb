"""Seeded synthetic bead workloads (the benchmark and parity-test inputs).

Beads are rendered from analytic radial LUTs with the reference's image model
(GenerateImageFromLUT: cubic B-spline in z, linear in r, 4x4 sub-pixel
samples, pixel centres at integer coordinates -- cputrack/utils.cpp:325-390)
and given Poisson noise.  The reference's own simulations use the same model
(cputrack-test/SharedTests.h:342-416, 495-569) with rand()-based noise, which
cannot be reproduced across C runtimes; here every random draw comes from a
numpy PCG64 generator seeded per workload (seed = 0x51AB0000 + index).

Expected photon counts per pixel: lambda = bg + peak * (1 + p) / 2 where p in
[-1, 1] is the rendered LUT value, so lambda >= bg.  Integer formats round and
clip to their range.
"""
import numpy as np

from . import api

SEED_BASE = 0x51AB0000

# QTrkSettings that reproduce each workload's sampling sizes through
# QTrkComputedConfig::Update (cputrack/QueuedTracker.cpp:6-21).
SETTINGS = {
    # 64x64: R=28, qi 32 radial x 64 angular/quadrant, zlut 32 x 123
    "plain64": dict(width=64, height=64, qi_minradius=2.0, qi_roi_coverage=0.875, qi_radial_coverage=1.25,
                    qi_angular_coverage=1.46, qi_angstep_factor=1.0, qi_iterations=5, zlut_minradius=2.0,
                    zlut_roi_coverage=0.875, zlut_radial_coverage=1.25, zlut_angular_coverage=0.7),
    # 100x100: R=50, qi 64 radial x 128 angular/quadrant, zlut 64 x 219
    "realtime100": dict(width=100, height=100, qi_minradius=1.0, qi_roi_coverage=1.0, qi_radial_coverage=1.5,
                        qi_angular_coverage=1.635, qi_angstep_factor=1.0, qi_iterations=5, zlut_minradius=1.0,
                        zlut_roi_coverage=1.0, zlut_radial_coverage=1.31, zlut_angular_coverage=0.7),
}

MODE_QI_Z = api.LT_QI | api.LT_LocalizeZ | api.LT_NormalizeProfile


def analytic_lut(planes, steps, minr, maxr, period, dperiod=0.04, width0=6.0, dwidth=0.05):
    """profile(r, z) = exp(-r^2 / (2 (width0 + dwidth z)^2)) cos(2 pi r / (period + dperiod z))
    at bin radii r_i = minr + i (maxr - minr) / steps and planes z = 0..planes-1; [planes][steps] float32."""
    r = minr + np.arange(steps, dtype=np.float64) * (maxr - minr) / steps
    z = np.arange(planes, dtype=np.float64)[:, None]
    w = width0 + dwidth * z
    p = np.exp(-r[None, :] ** 2 / (2 * w * w)) * np.cos(2 * np.pi * r[None, :] / (period + dperiod * z))
    return p.astype(np.float32)


def model_lut(planes=200, steps=48, maxr=24.0):
    """Nonnegative radial LUT for the bead-image model (bins r_i = i * maxr / steps, minr = 0): the
    realtime100 ring family made nonnegative, exp(-r^2 / (2 w(z)^2)) (1 + cos(2 pi r / T(z))) / 2 with
    w = 6 + 0.05 z px and T = 8 + 0.04 z px.  Every plane peaks at 1 (r = 0), so I in mu = bg + I g is
    the peak signal in photons per pixel, as the reference's rendered LUT images are scaled.
    [planes][steps] float32."""
    r = np.arange(steps, dtype=np.float64) * maxr / steps
    z = np.arange(planes, dtype=np.float64)[:, None]
    w = 6.0 + 0.05 * z
    T = 8.0 + 0.04 * z
    p = np.exp(-r[None, :] ** 2 / (2 * w * w)) * (1 + np.cos(2 * np.pi * r[None, :] / T)) / 2
    return p.astype(np.float32)


def model_params(n, rng, size=48, planes=200):
    """theta = (x, y, z, I, bg) of the differentiation workload: centre +-3 px, z in [20, 180],
    I in [2,000, 20,000], bg in [5, 50].  [n][5] float32."""
    c = size / 2
    return np.stack([c + rng.uniform(-3, 3, n), c + rng.uniform(-3, 3, n),
                     rng.uniform(20, min(180, planes - 20), n), rng.uniform(2000, 20000, n),
                     rng.uniform(5, 50, n)], axis=1).astype(np.float32)


# QTrkSettings of the differentiation workload's tracker (48x48 ROIs): QI 5 iterations, 32 radial
# steps, 48 Z-LUT radial steps from 0 to 24 px (the model LUT's bins)
MLE_SETTINGS = dict(width=48, height=48, qi_minradius=1.0, qi_roi_coverage=1.0, qi_radial_coverage=1.4,
                    qi_angular_coverage=1.5, qi_angstep_factor=1.0, qi_iterations=5, zlut_minradius=0.0,
                    zlut_roi_coverage=1.0, zlut_radial_coverage=2.0, zlut_angular_coverage=0.7)


def mle_start(model, lut, data):
    """The LUT-model fit's start for photon-count ROIs data [n][size][size] (a torch tensor on the
    GPU): x, y, z from the tracker's COM + QI + Z-LUT localization (MLE_SETTINGS, the model LUT's
    normalised profiles), bg from the ROI's 10th percentile (the rings reach the border, their dark
    minima do not), I from the photons above bg over the model's sum at that position.
    Returns theta0 [n][5] (torch, GPU)."""
    import torch

    n, size = data.shape[0], data.shape[1]
    planes = lut.shape[0]
    trk = api.QTrkInstance(api.default_settings(**MLE_SETTINGS))
    try:
        trk.set_radial_zlut(normalize_profiles(lut)[None])
        trk.set_localization_mode(MODE_QI_Z)
        batch = api.DeviceBatch(trk, n, api.QTrkFloat)
        try:
            batch.upload(data.cpu().numpy(), api.make_jobs(n))
            batch.run(n)
            batch.sync()
            res = batch.download(n)
        finally:
            batch.free()
    finally:
        trk.free()
    bg0 = torch.kthvalue(data.flatten(1), max(1, size * size // 10), dim=1).values.clamp_min(1.0)
    t0 = torch.zeros((n, 5), device=data.device)
    t0[:, 0] = torch.from_numpy(res["x"]).to(data.device)
    t0[:, 1] = torch.from_numpy(res["y"]).to(data.device)
    t0[:, 2] = torch.from_numpy(res["z"]).to(data.device).clamp(0, planes - 1)
    t0[:, 3] = 1.0
    gsum = model.forward(t0).sum((1, 2))
    t0[:, 3] = ((data.sum((1, 2)) - bg0 * size * size) / gsum).clamp_min(100.0)
    t0[:, 4] = bg0
    return t0


def scaled_model_lut(planes, steps, size):
    """model_lut's family for a size x size ROI: widths and ring periods scaled by size/48 and the
    family's z range spread over `planes` planes; bins at i * (size/2) / steps."""
    k = size / 48.0
    r = np.arange(steps, dtype=np.float64) * (size / 2) / steps
    z = np.arange(planes, dtype=np.float64)[:, None] * 200.0 / planes
    w = k * (6.0 + 0.05 * z)
    T = k * (8.0 + 0.04 * z)
    return (np.exp(-r[None, :] ** 2 / (2 * w * w)) * (1 + np.cos(2 * np.pi * r[None, :] / T)) / 2).astype(np.float32)


# The batching workload's trackers: ROI size, ROIs per batch, beads, LUT planes, radial steps
BATCHING = [(32, 2048, 64, 100, 32), (64, 4096, 256, 150, 64), (96, 1024, 64, 200, 96), (128, 512, 32, 300, 96)]


def batching_tracker(size, n, beads, planes, steps, rng, render=True):
    """One tracker of the batching workload: size x size uint16 ROIs (n = beads x frames, frame-major),
    per-tracker image LUT of `planes` x `steps` (scaled_model_lut), bead centre +-size/16 px, z in the
    middle 80% of the planes, I (peak photons/px) 2,000-20,000, bg 5-50, Poisson noise.  Returns
    dict(settings, model_lut, lut (normalised profiles for the tracker's compare), rois, jobs, truth)."""
    R = size / 2
    st = api.default_settings(width=size, height=size, qi_minradius=1.0, qi_roi_coverage=1.0,
                              qi_radial_coverage=1.4, qi_angular_coverage=1.5, qi_angstep_factor=1.0,
                              qi_iterations=5, zlut_minradius=0.0, zlut_roi_coverage=1.0,
                              zlut_radial_coverage=steps / R, zlut_angular_coverage=0.7)
    cfg = api.computed_config(st)
    assert cfg.zlut_radialsteps == steps, (cfg.zlut_radialsteps, steps)
    model = scaled_model_lut(planes, steps, size)
    bead = (np.arange(n) % beads).astype(np.int32)
    off = size / 16
    truth = np.stack([R + rng.uniform(-off, off, n), R + rng.uniform(-off, off, n),
                      rng.uniform(0.1 * planes, 0.9 * planes, n), rng.uniform(2000, 20000, n),
                      rng.uniform(5, 50, n)], axis=1).astype(np.float32)
    rois = None
    if render:
        g = api.generate_images_from_lut(model[None], truth[:, :3], np.zeros(n, np.int32), size, size, 0.0, R)
        lam = truth[:, 4, None, None] + truth[:, 3, None, None] * g.astype(np.float64)
        rois = np.clip(rng.poisson(lam), 0, 65535).astype(np.uint16)
    jobs = api.make_jobs(n, frame=(np.arange(n) // beads).astype(np.uint32), zlut_index=bead)
    lut = np.repeat(normalize_profiles(model)[None], beads, axis=0)
    return dict(settings=st, cfg=cfg, model_lut=model, lut=lut, rois=rois, jobs=jobs, truth=truth)


def normalize_profiles(lut):
    """Mean 0 / RMS 1 per profile (the normalisation the tracker applies to measured profiles)."""
    x = lut.astype(np.float64)
    x = x - x.mean(axis=-1, keepdims=True)
    x = x / np.sqrt((x * x).mean(axis=-1, keepdims=True))
    return x.astype(np.float32)


def _render(luts, lut_index, pos, cfg, peak, bg, rng, dtype):
    img = api.generate_images_from_lut(luts, pos, lut_index, cfg.width, cfg.height, cfg.zlut_minradius,
                                       cfg.zlut_maxradius, spline=True, ovs=4)
    lam = bg[:, None, None] + peak[:, None, None] * (1.0 + img.astype(np.float64)) * 0.5
    cnt = rng.poisson(np.maximum(lam, 0.0))
    if dtype == np.float32:
        return cnt.astype(np.float32)
    info = np.iinfo(dtype)
    return np.clip(cnt, 0, info.max).astype(dtype)


def make_workload(name, n=None, seed=None, planes=None, **overrides):
    """Returns dict(settings, mode, luts, rois, jobs, truth) for a named workload; `overrides`
    replace QTrkSettings fields (the LUTs follow the resulting radial sampling).  realtime100 with
    `planes` other than 200 spreads the same LUT family (and the z range) over that many planes.

    plain64:      1 bead, 100 planes x 32 steps, n (default 1000) float32 64x64 ROIs,
                  centre +-4 px, z in [5, 95], peak 10,000, background 20.
    realtime100:  256 beads x 16 frames (n default 4096) uint16 100x100 ROIs, per-bead LUT of
                  200 planes x 64 steps with ring period U[7, 10] px, offset +-5 px,
                  z in [10, 190], peak U[5000, 20000], background 50.
    """
    if name == "gauss16":
        return _gauss_workload(n, seed)
    idx = {"plain64": 0, "realtime100": 1}[name]
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + idx if seed is None else seed))
    st = api.default_settings(**dict(SETTINGS[name], **overrides))
    cfg = api.computed_config(st)
    if name == "plain64":
        n = 1000 if n is None else n
        nbeads, planes = 1, 100
        periods = np.array([8.0])
        render = analytic_lut(planes, cfg.zlut_radialsteps, cfg.zlut_minradius, cfg.zlut_maxradius, 8.0)[None]
        bead = np.zeros(n, np.int32)
        frame = np.arange(n, dtype=np.uint32)
        off = rng.uniform(-4, 4, size=(n, 2))
        z = rng.uniform(5, 95, size=n)
        peak = np.full(n, 10000.0)
        bg = np.full(n, 20.0)
        dtype = np.float32
    else:
        nbeads, nframes = 256, 16
        planes = 200 if planes is None else planes
        sc = 200.0 / planes
        n = nbeads * nframes if n is None else n
        periods = rng.uniform(7, 10, size=nbeads)
        render = np.stack([analytic_lut(planes, cfg.zlut_radialsteps, cfg.zlut_minradius, cfg.zlut_maxradius, p,
                                        dperiod=0.04 * sc, dwidth=0.05 * sc) for p in periods])
        k = np.arange(n)
        bead = (k % nbeads).astype(np.int32)
        frame = (k // nbeads).astype(np.uint32)
        off = rng.uniform(-5, 5, size=(n, 2))
        z = rng.uniform(10 / sc, 190 / sc, size=n)
        peak = rng.uniform(5000, 20000, size=n)
        bg = np.full(n, 50.0)
        dtype = np.uint16
    pos = np.stack([cfg.width / 2 + off[:, 0], cfg.height / 2 + off[:, 1], z], axis=1).astype(np.float32)
    rois = _render(render, bead, pos, cfg, peak, bg, rng, dtype)
    jobs = api.make_jobs(n, frame=frame, zlut_index=bead)
    return dict(name=name, settings=st, cfg=cfg, mode=MODE_QI_Z, luts=normalize_profiles(render),
                render_luts=render, periods=periods, rois=rois, jobs=jobs, truth=pos)


def bead_frame(w, h, centers, period=6.0, sigma=8.0, bg=100.0, amp=60.0, noise=0.0, seed=0):
    """A w x h float32 camera frame with radially symmetric beads (a cosine of the radius, of the
    given period, under a Gaussian envelope) centred at `centers` (float x, y), on a flat
    background, plus Gaussian noise of the given sigma."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.full((h, w), bg, np.float64)
    for cx, cy in centers:
        r = np.hypot(xx - cx, yy - cy)
        img += amp * np.cos(2 * np.pi * r / period) * np.exp(-(r / sigma) ** 2)
    if noise:
        img += np.random.default_rng(seed).normal(0, noise, img.shape)
    return img.astype(np.float32)


def gaussian_spots(pos, sigma, I0, bg, w, h):
    """Pixel-integrated Gaussian spots, GenerateGaussianSpotImage (cputrack/utils.cpp:393-402), float64."""
    from scipy.special import erf
    ed = 1.0 / np.sqrt(2 * sigma * sigma)
    x = np.arange(w)[None, :] - pos[:, 0:1]
    y = np.arange(h)[None, :] - pos[:, 1:2]
    dx = 0.5 * erf((x + 0.5) * ed) - 0.5 * erf((x - 0.5) * ed)
    dy = 0.5 * erf((y + 0.5) * ed) - 0.5 * erf((y - 0.5) * ed)
    return bg[:, None, None] + I0[:, None, None] * dy[:, :, None] * dx[:, None, :]


def _gauss_workload(n, seed):
    """gauss16: n (default 65,536) float32 16x16 ROIs of a sigma=1.5 px Gaussian spot, centre +-2 px,
    I0 U[2000, 20000], background U[5, 50], Poisson noise; LT_Gaussian2D with 10 Newton iterations."""
    n = 65536 if n is None else n
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 3 if seed is None else seed))
    st = api.default_settings(width=16, height=16, gauss2D_sigma=1.5, gauss2D_iterations=10)
    cfg = api.computed_config(st)
    pos = np.stack([8 + rng.uniform(-2, 2, n), 8 + rng.uniform(-2, 2, n)], 1)
    I0 = rng.uniform(2000, 20000, n)
    bg = rng.uniform(5, 50, n)
    lam = gaussian_spots(pos, 1.5, I0, bg, 16, 16)
    rois = rng.poisson(lam).astype(np.float32)
    jobs = api.make_jobs(n)
    truth = np.concatenate([pos, I0[:, None], bg[:, None]], 1).astype(np.float32)
    return dict(name="gauss16", settings=st, cfg=cfg, mode=api.LT_Gaussian2D, luts=None, rois=rois, jobs=jobs,
                truth=truth)
