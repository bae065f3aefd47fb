"""Throughput of the bead-localization hot path on one MI355X.

One step = one pass of the fused COM -> QI (5 iterations) -> Z-LUT chain over
one batch of 4,096 ROIs of 100x100 uint16 (256 beads x 16 frames, per-bead LUT
of 200 planes x 64 radial steps), with the ROIs already resident in GPU memory
(the device-resident batch of include/qtrk_amd.h).  The CPU baseline leg times
the oracle's QueuedCPUTracker restatement on the host cores on a bounded
sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
    python bench.py --copy-probe     # plain 256 MiB device copy, to calibrate HBM counters
    python bench.py --workload precision   # grouped Z-LUT match, exact vs BF16/FP16 matrix cores
    python bench.py --workload mle         # LUT-model MLE + Fisher/CRLB, and the 2D Gaussian MLE
    python bench.py --workload batching    # 4 trackers in one hipGraph per batch, p50/p99 latency
    python bench.py --workload mixed       # fused FP32 Z-LUT compare vs grouped BF16/FP16 match (device batch)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "localizations/sec (COM+QI+Z-LUT, 100×100 ROI) vs plain FP32; px/plane error"
PUBLISHED_CPU_LOCS = 4600.0  # reference CPU figure at ROI 100 (cudatrack-test/plot.png)
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per localization: 20,000 B ROI (100x100 u16) + 3,200 B of LUT
# (51,200 B per bead shared by its 16 frames) + 24 B job + 52 B result.
BYTES_PER_LOC = 20000 + 3200 + 24 + 52
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE (calibrated against a plain copy)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "localize_realtime100_hbm_traffic.txt")


SQ_PROFILE = os.path.join(ROOT, "profiles", "localize_realtime100_sq_counters.txt")
MATCH_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "lutmatch_precision_hbm_traffic.txt")
# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction
VALU_PEAK_WAVE_INSTR = 256 * 4 * 2.4e9 / 4


def _profile_value(path, key):
    try:
        for line in open(path):
            parts = line.lstrip("#").split()
            if parts and parts[0].rstrip(":") == key:
                return int(parts[1].replace(",", ""))
    except (OSError, IndexError, ValueError):
        pass
    return None


def measured_traffic():
    """Bytes per launch recorded by the committed counter summary, or None."""
    return _profile_value(TRAFFIC_PROFILE, "traffic_bytes_per_launch")


# untimed load before the warm-up steps, in seconds of consecutive batches
PREWARM_S = 0.2


def prewarm(b, n, seconds=PREWARM_S):
    """Consecutive batches of n ROIs on device batch b for at least `seconds` (untimed)."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        b.run(n, 8)
        b.sync()


def copy_probe(mib=256, reps=3):
    """A plain device-to-device copy of known size: the counter calibration kernel."""
    import torch

    a = torch.ones(mib * 2**18, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": a.numel() * 4, "copy_probe_bytes_written": b.numel() * 4,
                      "launches": reps}), flush=True)


# Worker threads of the CPU leg: the host cores this process may run on, up to the per-GPU CPU
# share of the measurement hosts (16 cores per GPU; their schedulers cap a GPU job's threads there)
CPU_SHARE = 16


def host_info():
    """nproc, the cores this process may use, and the CPU model (lscpu's "Model name")."""
    model = None
    try:
        import subprocess
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "affinity": aff, "model": model}


def cpu_baseline(w, seconds_target=20.0):
    """The oracle's QueuedCPUTracker restatement (job queue, worker threads, per-ROI ProcessJob) on
    the host cores, on a bounded sample of the same ROIs: ~seconds_target of single-core CPU work
    spread over the worker threads; the clock runs from the first job queued until every result is in (the reference's
    speed-test clock).  A single-thread run of 256 ROIs sizes the sample and is reported too."""
    from oracle.bind import Oracle

    o = Oracle()
    cfg = o.config_from_bytes(bytes(w["cfg"]))
    host = host_info()
    threads = max(1, min(host["affinity"], CPU_SHARE))
    n = len(w["jobs"])
    k1 = 256
    t1, _ = o.queued_run(cfg, w["mode"], w["luts"], w["rois"][:k1], w["jobs"][:k1], 1)
    per_loc = t1 / k1
    reps = max(1, int(seconds_target / (per_loc * n)))
    rois = np.concatenate([w["rois"]] * reps)
    jobs = np.concatenate([w["jobs"]] * reps)
    jobs["frame"] = np.arange(len(jobs))
    t, _ = o.queued_run(cfg, w["mode"], w["luts"], rois, jobs, threads)
    return {"value": round(len(jobs) / t, 1), "unit": "localizations/s", "cores": threads, "kind": "port",
            "one_thread": round(k1 / t1, 1), "host": host,
            "sample": f"{len(jobs)} ROIs ({reps}x the {n}-ROI batch) on {threads} worker threads (the host's "
                      f"cores available to this process, up to its {CPU_SHARE}-core per-GPU share), oracle "
                      f"QueuedCPUTracker restatement; one thread: 256 ROIs"}


def precision_profiles(beads, frames, planes, seed):
    """LUTs and measured profiles of the precision workload: per-bead analytic LUTs of `planes` x 128
    radial steps (ring period U[7, 10] px), and one 100x100 uint16 ROI per profile rendered from the
    bead's LUT at 10,000 photons over a background of 50 with Poisson noise, centre +-5 px, z in
    [10, planes - 10]; the profiles are the tracker's normalised radial profiles of those ROIs."""
    import torch
    from qtrk_amd import api, synth

    rng = np.random.Generator(np.random.PCG64(seed))
    st = api.default_settings(**dict(synth.SETTINGS["realtime100"], zlut_radial_coverage=2.6224))
    cfg = api.computed_config(st)
    assert cfg.zlut_radialsteps == 128, cfg.zlut_radialsteps
    scale = 200.0 / planes  # the realtime100 family's z range spread over `planes` planes
    render = np.stack([synth.analytic_lut(planes, 128, cfg.zlut_minradius, cfg.zlut_maxradius, p,
                                          dperiod=0.04 * scale, dwidth=0.05 * scale)
                       for p in rng.uniform(7, 10, size=beads)])
    luts = synth.normalize_profiles(render)
    n = beads * frames
    bead = (np.arange(n) // frames).astype(np.int32)
    off = rng.uniform(-5, 5, size=(n, 2))
    z = rng.uniform(10, planes - 10, size=n)
    pos = np.stack([50 + off[:, 0], 50 + off[:, 1], z], axis=1).astype(np.float32)
    trk = api.QTrkInstance(st)
    trk.set_radial_zlut(luts)
    trk.set_localization_mode(synth.MODE_QI_Z)
    chunk = 8192
    batch = api.DeviceBatch(trk, chunk, api.QTrkU16)
    batch.enable_diagnostics()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    prof = np.zeros((n, 128), np.float32)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        img = api.generate_images_from_lut(render, pos[c0:c1], bead[c0:c1], 100, 100, cfg.zlut_minradius,
                                           cfg.zlut_maxradius)
        lam = 50.0 + 10000.0 * (1.0 + torch.from_numpy(img).cuda()) * 0.5
        cnt = torch.poisson(lam.clamp_min(0), generator=gen).clamp_(0, 65535).to(torch.int32).cpu().numpy()
        batch.upload(cnt.astype(np.uint16), api.make_jobs(c1 - c0, zlut_index=bead[c0:c1]))
        batch.run(c1 - c0)
        batch.sync()
        prof[c0:c1] = batch.diagnostics(c1 - c0)[1]
    batch.free()
    trk.free()
    return luts, prof.reshape(beads, frames, 128), z.reshape(beads, frames)


# Bytes of one precision-workload match counted without the exact rescoring's FP32 rows: BF16 LUT
# 131.1 MB + BF16 profiles 16.8 MB + top-8 output 4.2 MB
MATCH_BYTES_NO_RESCORE = 1024 * 500 * 128 * 2 + 65536 * 128 * 2 + 65536 * 8 * 8


def precision_bench(a):
    """Grouped Z-LUT match, 1,024 beads x 500 planes x 128 radial steps, 64 profiles per bead:
    the reference arithmetic over every plane (fp32) against the BF16 and FP16 matrix-core paths."""
    from oracle.bind import Oracle
    from qtrk_amd import api

    beads, frames, planes = 1024, 64, 500
    luts, prof, ztrue = precision_profiles(beads, frames, planes, 0x51AB0002)
    n = beads * frames
    m = api.LUTMatch(luts)
    m.upload(prof)
    ptiles, ksteps = (planes + 31) // 32, 128 // 16
    out = {}
    for name, prec in (("fp32", api.MATCH_FP32), ("bf16", api.MATCH_BF16), ("fp16", api.MATCH_FP16)):
        m.run(prec, a.warmup)
        m.sync()
        t0 = time.perf_counter()
        m.run(prec, a.steps)
        ev_ms = m.sync()
        wall = time.perf_counter() - t0
        z, best, cand = m.download()
        kern_s = ev_ms / a.steps / 1e3
        r = {"profiles_per_s": round(n * a.steps / wall, 1), "kernel_ms": round(kern_s * 1e3, 4),
             "rms_z_vs_truth": round(float(np.sqrt(np.mean((z - ztrue) ** 2))), 4)}
        if prec == api.MATCH_FP32:
            bytes_ = luts.nbytes + prof.nbytes + n * 8
            z32, best32 = z, best
        else:
            # distinct (bead, plane) rows rescored in FP32: the candidates and the fit window
            win = best[..., None] + np.arange(-3, 4)
            rows = np.concatenate([cand, np.clip(win, 0, planes - 1)], axis=-1).reshape(beads, -1)
            distinct = sum(len(np.unique(rr[rr >= 0])) for rr in rows)
            mtiles = (ptiles + 15) // 16  # tile-mean fragments: 16 tiles' high and low parts per 32 rows
            bytes_ = (beads * ptiles * ksteps * 64 * 16 + beads * mtiles * ksteps * 64 * 16 + beads * ptiles * 32 * 4
                      + prof.nbytes + n * (4 + 4 + 32) + distinct * 128 * 4)
            r["mfma_tflops"] = round(2.0 * n * planes * 128 / kern_s / 1e12, 2)
            r["mfma_frac_of_dense_peak"] = round(2.0 * n * planes * 128 / kern_s / 2.5e15, 5)
            r["best_plane_equal_to_fp32_path"] = int(np.sum(best == best32))
            r["max_abs_dz_vs_fp32_path"] = float(np.max(np.abs(z - z32)))
        r["hbm_bytes_per_launch"] = int(bytes_)
        r["hbm_gbs"] = round(bytes_ / kern_s / 1e9, 1)
        out[name] = r
        if prec == api.MATCH_BF16:
            zb = z
    m.free()
    # parity sample against the oracle (FP32 restatement) and the FP64 oracle: 16 beads x 64 profiles
    o, o64 = Oracle(), Oracle(fp64=True)
    nb = 16
    t0 = time.perf_counter()
    zo = np.array([[o.lut_compare(prof[b, f], luts[b])[0] for f in range(frames)] for b in range(nb)])
    t_cpu = time.perf_counter() - t0
    z64 = np.array([[o64.lut_compare(prof[b, f], luts[b])[0] for f in range(frames)] for b in range(nb)])
    res = {
        "metric": "profiles/s (grouped Z-LUT match, BF16 MFMA + FP32 top-8 rescore + FP64 fit)",
        "value": out["bf16"]["profiles_per_s"], "unit": "profiles/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"precision: {beads} beads x {planes} planes x 128 radial steps, {frames} profiles "
                               "per bead from 100x100 u16 ROIs at 10,000 photons with Poisson noise",
                   "paths": out,
                   "oracle_sample": {"profiles": nb * frames,
                                     "bf16_equal_to_oracle": bool(np.array_equal(zb[:nb], zo.astype(np.float32))),
                                     "max_abs_dz_bf16_vs_fp64_oracle": float(np.max(np.abs(zb[:nb] - z64)))}},
        "roofline": {"bound": "hbm", "achieved": out["bf16"]["hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(out["bf16"]["hbm_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": _profile_value(MATCH_TRAFFIC_PROFILE, "traffic_bytes_per_launch"),
                     "kernel_ms": out["bf16"]["kernel_ms"],
                     # against the bytes without the exact rescoring's FP32 rows
                     "frac_without_rescore_bytes": round(MATCH_BYTES_NO_RESCORE / (out["bf16"]["kernel_ms"] * 1e-3) / 1e9
                                                 / HBM_PEAK_GBS, 4),
                     "mfma_frac_of_dense_bf16_peak": out["bf16"]["mfma_frac_of_dense_peak"]},
        "cpu_baseline": {"value": round(nb * frames / t_cpu, 1), "unit": "profiles/s", "cores": 1, "kind": "port",
                         "sample": f"{nb * frames} profiles through the oracle's LUTProfileCompare + FP64 fit"},
    }
    print(json.dumps(res), flush=True)


# FLOPs of one model pass per pixel, counted from the fit's pixel model in lutmodel.hip (a fused
# multiply-add counts 2): per sub-sample 20 (r^2 and its floor 2, reciprocal square root 1, scaled
# radius 3, clamp 1, fraction 1, value 3, z-slope 3, slope 2, x/y slope 4), ovs^2 = 16 sub-samples,
# the per-pixel sub-sample offsets 24, then mu and the Jacobian 7, the weights 3, the gradient 10,
# the Fisher matrix 35 and the final scaling 6.
MLE_FLOP_PER_PX_PASS = 16 * 20 + 24 + 7 + 3 + 10 + 35 + 6


def mle_bench(a):
    """Differentiation workload: 65,536 ROIs of 48x48 rendered from a 200-plane x 48-step B-spline
    LUT (x, y +-3 px, z 20-180, I 2,000-20,000, bg 5-50, Poisson noise), QI + Z-LUT start, 10 Newton
    (Fisher-scoring) iterations of the LUT-model MLE with FP64 5x5 Fisher/CRLB; and the 2D Gaussian
    MLE (sigma 1.5 px, 10 iterations) on 65,536 ROIs of 16x16."""
    import torch
    from qtrk_amd import api, synth

    n, size, iters = 65536, 48, 10
    rng = np.random.Generator(np.random.PCG64(0x51AB0003))
    lut = synth.model_lut(200, 48, 24.0)
    truth = synth.model_params(n, rng, size)
    model = api.LUTModel(lut, 0.0, 24.0, size, size)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x51AB0003)
    data = torch.poisson(model.forward(torch.from_numpy(truth)), generator=gen)
    # starting point: the tracker's COM + QI + Z-LUT position, a background estimate, and the
    # intensity from the photons above it over the model's sum at that position
    d = data
    t0 = synth.mle_start(model, lut, d)
    start_rms = np.sqrt(np.mean((t0[:, :3].cpu().numpy() - truth[:, :3]) ** 2, axis=0))

    model.fit(d, t0, iters)  # warm-up
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    ev0.record()
    for _ in range(a.steps):
        th, F, crlb, nll = model.fit(d, t0, iters)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    fit_ms = ev0.elapsed_time(ev1) / a.steps
    th, crlb = th.cpu().numpy(), crlb.cpu().numpy()
    ok = np.all(np.isfinite(crlb), axis=1)
    e = (th[ok] - truth[ok]) / crlb[ok]
    # a fit whose start lies outside the likelihood's capture range ends far from the truth: count
    # those (|error| > 6 CRLB in any parameter) and give the precision of the others
    inl = np.all(np.abs(e) <= 6, axis=1)
    flop = n * size * size * MLE_FLOP_PER_PX_PASS * (iters + 1)

    # the differentiable renderer on the same ROIs: forward (mu) and adjoint (d loss / d theta from
    # d loss / d mu), the QTrkLUTModelForward / QTrkLUTModelBackward pair
    gmu = torch.ones_like(d)
    model.forward(t0)
    model.backward(t0, gmu)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(a.steps):
        model.forward(t0)
    ev1.record()
    torch.cuda.synchronize()
    render_ms = ev0.elapsed_time(ev1) / a.steps
    ev0.record()
    for _ in range(a.steps):
        model.backward(t0, gmu)
    ev1.record()
    torch.cuda.synchronize()
    adjoint_ms = ev0.elapsed_time(ev1) / a.steps
    model.free()

    # 2D Gaussian MLE through the tracker (LT_Gaussian2D), 65,536 ROIs of 16x16
    g = synth.make_workload("gauss16")
    gtrk = api.QTrkInstance(g["settings"])
    gtrk.set_localization_mode(g["mode"])
    gb = api.DeviceBatch(gtrk, len(g["jobs"]), api.QTrkFloat)
    gb.upload(g["rois"], g["jobs"])
    gb.run(len(g["jobs"]), a.warmup)
    gb.sync()
    tg = time.perf_counter()
    gb.run(len(g["jobs"]), a.steps)
    g_ms = gb.sync() / a.steps
    g_wall = time.perf_counter() - tg
    gres = gb.download(len(g["jobs"]))
    gb.free()
    gtrk.free()
    gfin = np.isfinite(gres["x"]) & np.isfinite(gres["y"])
    g_rms = float(np.sqrt(np.mean((gres["x"][gfin] - g["truth"][gfin, 0]) ** 2
                                  + (gres["y"][gfin] - g["truth"][gfin, 1]) ** 2) / 2))

    out = {
        "metric": "ROIs/s (LUT-model MLE, 10 Newton iterations + FP64 5x5 Fisher/CRLB, 48x48 ROI)",
        "value": round(n * a.steps / wall, 1), "unit": "ROIs/s", "n_gpus": 1, "steps": a.steps,
        "warmup": 1, "ms_per_step": round(wall / a.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"differentiation: {n} ROIs of 48x48 f32 from a 200x48 B-spline LUT (I = peak "
                               "photons/px 2,000-20,000, bg 5-50), Poisson "
                               "noise, start from COM + QI + Z-LUT, 10 Fisher-scoring iterations",
                   "fit_kernel_ms": round(fit_ms, 3),
                   "start_rms_xyz": [round(float(v), 4) for v in start_rms],
                   "fit_rms_xyz": [round(float(v), 4) for v in np.sqrt(np.mean((th[:, :3] - truth[:, :3]) ** 2, 0))],
                   "std_error_over_crlb": dict(zip("x y z I bg".split(), np.round(e[inl].std(0), 4).tolist())),
                   "mean_error_over_crlb": dict(zip("x y z I bg".split(), np.round(e[inl].mean(0), 4).tolist())),
                   "rois_with_crlb": int(ok.sum()), "outliers_beyond_6_crlb": int((~inl).sum()),
                   "std_error_over_crlb_all": dict(zip("x y z I bg".split(), np.round(e.std(0), 3).tolist())),
                   "renderer": {"forward_ms": round(render_ms, 4), "adjoint_ms": round(adjoint_ms, 4),
                                "note": f"{n} ROIs of 48x48, 4x4 sub-samples per pixel, reference rounding"},
                   "gaussian2d_16x16": {"rois": len(g["jobs"]), "rois_per_s": round(len(g["jobs"]) * a.steps / g_wall, 1),
                                        "kernel_ms": round(g_ms, 4), "rms_xy_px_finite": round(g_rms, 4),
                                        "nan_results": int((~gfin).sum()),
                                        "note": "the oracle's restatement of the reference's diagonal Newton "
                                                "gives NaN on the same ROIs"}},
        "roofline": {"bound": "fp32", "achieved": round(flop / (fit_ms * 1e-3) / 1e12, 2), "peak": 157.3,
                     "unit": "TFLOP/s", "frac": round(flop / (fit_ms * 1e-3) / 157.3e12, 4), "traffic": None,
                     "flop_per_launch": flop},
    }
    print(json.dumps(out), flush=True)


def _batching_pipeline(frames_div, rng):
    """The batching workload's four trackers in one pipeline, each with 1/frames_div of its frames per
    batch (frames_div = 1: the full 2,048/4,096/1,024/512 ROIs).  Returns (pipeline, trackers, works)."""
    from qtrk_amd import api, synth

    p = api.Pipeline()
    trks, works = [], []
    for size, n, beads, planes, steps in synth.BATCHING:
        frames = max(1, n // beads // frames_div)
        w = synth.batching_tracker(size, beads * frames, beads, planes, steps, rng)
        trk = api.QTrkInstance(w["settings"])
        trk.set_radial_zlut(w["lut"])
        trk.set_localization_mode(synth.MODE_QI_Z)
        s = p.add(trk, len(w["jobs"]), api.QTrkU16, api.MATCH_BF16, w["model_lut"], 0.0, w["cfg"].zlut_maxradius, 5)
        p.set_jobs(s, w["jobs"])
        p.upload(s, w["rois"])
        trks.append(trk)
        works.append(w)
    p.finalize(True)
    return p, trks, works


def batching_bench(a):
    """4 trackers (ROI 32/64/96/128 px; 2,048/4,096/1,024/512 uint16 ROIs per batch; own LUTs of
    100-300 planes x 32-96 steps) served by one hipGraph launch per batch: COM -> QI (5) -> BF16
    grouped Z-LUT match -> 5-iteration LUT-model MLE; launch shapes autotuned; camera-paced batches
    with p50/p99 latency: the full batch at a 1 ms period (more work than 1 ms holds: the queue
    grows) and at a period it sustains, and smaller batches at a 1 ms period -- a quarter of each
    tracker's frames, and one frame of every tracker (a 1 kHz camera per tracker)."""
    from qtrk_amd import synth

    BATCHING = synth.BATCHING
    rng = np.random.default_rng(0x51AB0004)
    p, trks, works = _batching_pipeline(1, rng)
    report = p.report()
    nroi = sum(b[1] for b in BATCHING)
    # back-to-back graph launches: the sustained rate
    for _ in range(a.warmup):
        p.launch()
        p.sync()
    t0 = time.perf_counter()
    gms = []
    for _ in range(a.steps):
        p.launch()
        gms.append(p.sync())
    wall = (time.perf_counter() - t0) / a.steps
    per_batch = float(np.median(gms))
    # camera-paced: 1 ms batches, and batches at a period the GPU sustains
    lat1, _ = p.stream(200, 1.0)
    period = max(1.0, round(per_batch * 1.25, 2))
    latp, gpup = p.stream(200, period)
    acc = {}
    for (size, n, *_), w, s in zip(BATCHING, works, range(len(BATCHING))):
        res, th, cr = p.download(s)
        t = w["truth"]
        e = (th - t) / cr
        inl = np.all(np.abs(e) <= 6, axis=1)
        qi_err = np.abs(res["x"] - t[:, 0])
        acc[f"{size}px"] = {"median_abs_x_qi_px": round(float(np.median(qi_err)), 4),
                            "qi_beyond_1px": int(np.sum(qi_err > 1)),
                            "median_abs_z_match_planes": round(float(np.nanmedian(np.abs(res["z"] - t[:, 2]))), 4),
                            "median_abs_x_mle_px": round(float(np.median(np.abs(th[:, 0] - t[:, 0]))), 4),
                            "std_error_over_crlb_xyz": np.round(e[inl, :3].std(0), 3).tolist(),
                            "outliers_beyond_6_crlb": int((~inl).sum())}
    p.free()
    for trk in trks:
        trk.free()
    q = lambda x, v: round(float(np.percentile(x, v)), 3)
    # smaller batches at the 1 ms camera period
    paced = {}
    for label, div in (("quarter_of_the_frames", 4), ("one_frame_per_tracker", 10 ** 6)):
        sp, strks, sworks = _batching_pipeline(div, rng)
        k = sum(len(w["jobs"]) for w in sworks)
        for _ in range(a.warmup):
            sp.launch()
            sp.sync()
        lat, gpu = sp.stream(1000, 1.0)
        paced[label] = {"rois_per_batch": k, "rois_per_s": round(k / 1e-3, 1), "graph_gpu_ms_p50": q(gpu, 50),
                        "latency_ms": {"p50": q(lat, 50), "p99": q(lat, 99), "max": q(lat, 100)}}
        sp.free()
        for trk in strks:
            trk.free()
    out = {
        "metric": "ROIs/s through one hipGraph launch serving 4 trackers (COM->QI5->BF16 Z-LUT->MLE5)",
        "value": round(nroi / wall, 1), "unit": "ROIs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+bf16", "data": "synthetic",
        "config": {"workload": "batching: trackers " + ", ".join(f"{s}px x {n} ROIs ({pl} planes x {st} steps)"
                                                                   for s, n, _, pl, st in BATCHING)
                               + "; uint16 Poisson ROIs resident in HBM, one graph launch per batch",
                   "rois_per_batch": nroi, "graph_gpu_ms_p50": round(per_batch, 3),
                   "latency_ms_at_1ms_period": {"p50": q(lat1, 50), "p99": q(lat1, 99), "max": q(lat1, 100)},
                   f"latency_ms_at_{period}ms_period": {"p50": q(latp, 50), "p99": q(latp, 99),
                                                         "gpu_p50": q(gpup, 50)},
                   "smaller_batches_at_1ms_period": paced,
                   "accuracy": acc, "autotuner": report.strip().split("\n")},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rois", type=int, default=None, help="ROIs per step (realtime100: 4096, plain64: 1000)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-frames", action="store_true", help="skip the PCIe-inclusive frame-path rate")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--workload", default="realtime100",
                    choices=["realtime100", "plain64", "queue", "precision", "mle", "batching", "findbeads", "mixed"])
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the path shards by bead group across "
              "GPUs with no exchange step, so the whole-job rate is N x the one-GPU rate", file=sys.stderr)
        sys.exit(2)
    if a.workload == "precision":
        precision_bench(a)
        return
    if a.workload == "mle":
        mle_bench(a)
        return
    if a.workload == "batching":
        batching_bench(a)
        return
    if a.workload == "queue":
        queue_bench(a)
        return
    if a.workload == "findbeads":
        findbeads_bench(a)
        return
    if a.workload == "mixed":
        mixed_bench(a)
        return

    localize_bench(a, a.workload)


def findbeads_bench(a):
    """QTrkFindBeads on a 1,600 x 1,200 float frame with 120 beads and a 40 x 40 sample, padded
    to 2,048 x 2,048: GPU transforms and re-centring plus the host scan, timed
    per call from host memory; the oracle's restatement (one core) is the CPU baseline."""
    from qtrk_amd import api, synth

    w, h, roi = 1600, 1200, 40
    rng = np.random.default_rng(1600)
    cent = [(roi + roi // 2 + c * 2 * roi + rng.uniform(-3, 3), roi + roi // 2 + r * 2 * roi + rng.uniform(-3, 3))
            for r in range(6) for c in range(20)]
    img = synth.bead_frame(w, h, cent, noise=3.0, seed=1600)
    sx, sy = int(cent[0][0]) - roi // 2, int(cent[0][1]) - roi // 2
    for _ in range(max(1, a.warmup)):
        got = api.find_beads(img, (sx, sy), roi, 0.8, 0.45)
    steps = max(1, min(a.steps, 20))
    t0 = time.perf_counter()
    for _ in range(steps):
        got = api.find_beads(img, (sx, sy), roi, 0.8, 0.45)
    t = (time.perf_counter() - t0) / steps
    out = {"metric": "QTrkFindBeads calls/sec (1600x1200 float frame, 120 beads)", "value": round(1 / t, 2),
           "unit": "calls/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": round(t * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "findbeads: 1600x1200 f32 frame padded to 2048x2048, roi 40, "
                                  "img_rel_dist 0.8, acceptance 0.45", "beads_found": int(len(got)),
                      "beads_planted": len(cent)}}
    if not a.no_cpu:
        from oracle.bind import Oracle
        o = Oracle()
        t0 = time.perf_counter()
        ref = o.find_beads(img, sx, sy, roi, 0.8, 0.45)
        tc = time.perf_counter() - t0
        out["config"]["equal_to_oracle"] = bool(np.array_equal(ref, got))
        out["cpu_baseline"] = {"value": round(1 / tc, 3), "unit": "calls/s", "cores": 1, "kind": "port",
                               "sample": "one call of the same frame through the oracle's restatement"}
    print(json.dumps(out), flush=True)


def _camera_frames(w, per=256, side=16):
    """The realtime100 batch as 16 camera frames of 1,600 x 1,600 u16, each holding a 16 x 16 grid of
    100 x 100 ROIs (one per bead): (frames, ROI positions, one job per frame)."""
    from qtrk_amd import api

    n = len(w["jobs"])
    nf = n // per
    frames = np.zeros((nf, side * 100, side * 100), np.uint16)
    for k in range(n):
        f, i = divmod(k, per)
        r, c = divmod(i, side)
        frames[f, r * 100:(r + 1) * 100, c * 100:(c + 1) * 100] = w["rois"][k]
    pos = np.array([[(i % side) * 100, (i // side) * 100] for i in range(per)], np.int32)
    jobs = [api.make_jobs(1, frame=np.array([f], np.uint32), zlut_index=np.array([0], np.int32))[0]
            for f in range(nf)]
    return frames, pos, jobs


def frame_path_rates(trk, frames, pos, jobs, steps, warmup):
    """Localizations/s through QTrkScheduleFrame from host frames (copies into pinned memory and over
    PCIe included), two ways:
      * drained: the clock of the reference's speed tests (cputrack-test/SharedTests.h:300-314,
        cudatrack-test/test.cu:310-371) -- from the first schedule call until GetResultCount() == N
        after Flush, N = one pass of all frames; min/median/max over `steps` passes;
      * stream: the frames of `steps` passes scheduled back to back with results fetched as they
        arrive, then Flush and the last results (a camera stream's sustained rate).
    Returns (dict, the last drained pass's results)."""
    nf = len(frames)
    n = nf * len(pos)

    def one_pass():
        for f in range(nf):
            trk.schedule_frame(frames[f], pos, jobs[f])
        trk.wait_for(n)
        return trk.fetch_results(n)

    def stream(k):
        got = 0
        for _ in range(k):
            for f in range(nf):
                trk.schedule_frame(frames[f], pos, jobs[f])
            got += len(trk.fetch_results())
        trk.wait_for(n * k - got)
        got += len(trk.fetch_results())
        return got

    for _ in range(max(1, warmup)):
        one_pass()
    # the first streamed passes of a process run at about half the steady host-to-device rate
    # (measured: ~20 GB/s for the first 20 passes, then 40-47 GB/s), so the warm-up streams for at
    # least one second as well as `warmup` passes
    t0 = time.perf_counter()
    stream(max(1, warmup))
    while time.perf_counter() - t0 < 1.0:
        stream(4)
    t0 = time.perf_counter()
    got = stream(steps)
    wall = time.perf_counter() - t0
    assert got == n * steps, got
    passes = []
    for _ in range(steps):
        t = time.perf_counter()
        res = one_pass()
        passes.append(time.perf_counter() - t)
    med = float(np.median(passes))
    return {"drained_locs_per_s": round(n / med, 1),
            "drained_pass_ms_min_median_max": [round(v * 1e3, 3) for v in (min(passes), med, max(passes))],
            "stream_locs_per_s": round(n * steps / wall, 1),
            "pcie_gbs_stream": round(frames.nbytes * steps / wall / 1e9, 2),
            "host_bytes_per_pass": int(frames.nbytes)}, res


def c_abi_speed_test(w, mode, passes):
    """The reference's speed-test clock through the C ABI from a C program (qtrk_amd/speed_test.c):
    every ROI of the workload scheduled from host memory by one call (mode "frame1": QTrkScheduleFrame
    with the ROI as a one-ROI frame, cudatrack-test/test.cu:343-347; "roi": QTrkScheduleLocalization),
    then Flush and wait until the result count is N.  Returns (dict, results ordered by frame, bead)."""
    import shutil
    import subprocess
    import tempfile

    from qtrk_amd import api
    from qtrk_amd.build import SPEED_BIN

    if not os.path.exists(SPEED_BIN):
        raise RuntimeError(f"{SPEED_BIN} is not built (qtrk_amd/build.py build_speed_test)")
    n, h, wd = w["rois"].shape
    d = tempfile.mkdtemp(prefix="qtrk_speed_")
    try:
        luts = np.ascontiguousarray(w["luts"], np.float32)
        with open(os.path.join(d, "params.txt"), "w") as f:
            f.write(f"{n} {wd} {h} {luts.shape[0]} {luts.shape[1]} {luts.shape[2]} {int(w['mode'])}\n")
        with open(os.path.join(d, "settings.bin"), "wb") as f:
            f.write(bytes(w["settings"]))
        luts.tofile(os.path.join(d, "lut.bin"))
        np.ascontiguousarray(w["rois"], np.uint16).tofile(os.path.join(d, "rois.bin"))
        np.ascontiguousarray(w["jobs"]).tofile(os.path.join(d, "jobs.bin"))
        r = subprocess.run([SPEED_BIN, d, mode, str(passes)], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"qtrk_speed_test failed ({r.returncode}): {r.stderr[-2000:]}")
        out = json.loads(r.stdout.strip().splitlines()[-1])
        res = np.fromfile(os.path.join(d, "results.bin"), api.RESULT_DTYPE)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    rates = out["locs_per_s"]
    return {"call": {"frame1": "QTrkScheduleFrame, one ROI per call (the reference's SpeedTest)",
                     "roi": "QTrkScheduleLocalization per ROI"}[mode],
            "locs_per_s_median": round(float(np.median(rates)), 1),
            "locs_per_s_min_max": [round(min(rates), 1), round(max(rates), 1)], "passes": passes,
            "schedule_call_us_median": round(float(np.median(out["schedule_call_us"])), 3),
            "note": "median over passes; the first pass also grows the slots' pinned buffers"}, \
        res[np.lexsort((res["zlutIndex"], res["frame"]))]


def queue_bench(a):
    """The realtime100 ROIs through the drop-in queue API from host memory: 16 camera frames of
    1,600 x 1,600 u16 per step, each holding a 16 x 16 grid of the batch's 100 x 100 ROIs (one per
    bead), handed to QTrkScheduleFrame with the 256 ROI positions (frame_path_rates)."""
    from qtrk_amd import api, synth

    w = synth.make_workload("realtime100", n=4096)
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["luts"])
    trk.set_localization_mode(w["mode"])
    frames, pos, jobs = _camera_frames(w)
    rates, res = frame_path_rates(trk, frames, pos, jobs, a.steps, a.warmup)
    # the same localizations as the device-resident batch (results are LIFO: order by frame, bead)
    b = api.DeviceBatch(trk, 4096, api.QTrkU16)
    b.upload(w["rois"], w["jobs"])
    b.run(4096, 1)
    b.sync()
    ref = b.download(4096)
    b.free()
    trk.free()
    res = res[np.lexsort((res["zlutIndex"], res["frame"]))]
    same = bool(np.array_equal(res["x"], ref["x"]) and np.array_equal(res["z"], ref["z"]))
    # one ROI per call from a C client, the reference's own speed-test pattern
    per_roi = {}
    for mode in ("frame1", "roi"):
        st, sres = c_abi_speed_test(w, mode, max(3, a.steps // 4))
        st["results_equal_device_batch"] = bool(np.array_equal(sres["x"], ref["x"]) and np.array_equal(sres["z"], ref["z"]))
        per_roi[mode] = st
    rates["one_roi_per_call_c_client"] = per_roi
    value = rates["stream_locs_per_s"]
    out = {"metric": "localizations/sec through QTrkScheduleFrame from host frames (PCIe included)",
           "value": value, "unit": "localizations/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(4096 / value * 1e3, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": dict({"workload": "realtime100 ROIs as 16 frames of 1600x1600 u16 (256 ROIs each) per step, "
                                       "QTrkScheduleFrame streamed, results fetched as they arrive",
                           "results_equal_device_batch": same}, **rates)}
    print(json.dumps(out), flush=True)


def mixed_bench(a):
    """The realtime100 batch (4,096 ROIs of 100x100 u16, 256 beads x 16 frames) with the Z-LUT match
    done two ways, on the default 200 x 64 LUTs and on 500 x 128 LUTs (the precision workload's size):
      * fused FP32: one localize_kernel per batch, each ROI's LUT scores in the reference
        arithmetic (FP64 sums) inside its workgroup -- the headline path;
      * split FP32: the kernel in profile-only form, the ROIs grouped by bead on the GPU, then each
        bead's 16 profiles scored against every plane together in the reference arithmetic
        (QTRK_MATCH_FP32_GROUPED; results identical to the fused form's);
      * split BF16 / FP16: the same grouping, each bead's profiles matched together on the matrix
        cores, the 8 best planes and the fit window rescored in the reference arithmetic, FP64 fit.
    Per form: loc/s over consecutive batches, and the error against the truth (bias = mean,
    sigma = standard deviation; x, y in px, z in planes) with the count of results identical to
    the fused form's."""
    from qtrk_amd import api, synth

    cases = {}
    for label, over, planes in (("lut_200x64", {}, 200), ("lut_500x128", {"zlut_radial_coverage": 2.6224}, 500)):
        w = synth.make_workload("realtime100", n=4096, planes=planes, **over)
        n = len(w["jobs"])
        trk = api.QTrkInstance(w["settings"])
        trk.set_radial_zlut(w["luts"])
        trk.set_localization_mode(w["mode"])
        b = api.DeviceBatch(trk, n, api.QTrkU16)
        b.upload(w["rois"], w["jobs"])
        prewarm(b, n)
        forms, fused = {}, None
        for name, prec in (("fused_fp32", api.MATCH_FP32), ("split_fp32", api.MATCH_FP32_GROUPED),
                           ("split_bf16", api.MATCH_BF16), ("split_fp16", api.MATCH_FP16)):
            b.set_match_precision(prec)
            b.run(n, a.warmup)
            b.sync()
            t0 = time.perf_counter()
            b.run(n, a.steps)
            b.sync()
            sec = (time.perf_counter() - t0) / a.steps
            res = b.download(n)
            if fused is None:
                fused = res
            t = w["truth"]
            e = np.stack([res["x"] - t[:, 0], res["y"] - t[:, 1], res["z"] - t[:, 2]], 1).astype(np.float64)
            forms[name] = {"locs_per_s": round(n / sec, 1), "ms_per_batch": round(sec * 1e3, 4),
                           "bias_xyz": np.round(e.mean(0), 5).tolist(), "sigma_xyz": np.round(e.std(0), 5).tolist(),
                           "identical_to_fused": int(np.sum((res["x"] == fused["x"]) & (res["y"] == fused["y"])
                                                            & (res["z"] == fused["z"])))}
        b.free()
        trk.free()
        cases[label] = {"luts": f"{w['luts'].shape[0]} beads x {w['luts'].shape[1]} planes x {w['luts'].shape[2]} steps",
                        "forms": forms}
    big = cases["lut_500x128"]["forms"]
    out = {"metric": "localizations/sec, realtime100 ROIs: fused FP32 Z-LUT compare vs grouped BF16/FP16 match",
           "value": big["split_bf16"]["locs_per_s"], "unit": "localizations/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": big["split_bf16"]["ms_per_batch"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
           "config": {"workload": "realtime100: 4096 ROIs of 100x100 u16, 256 beads x 16 frames, COM + QI(5) + Z "
                                  "(value: the 500 x 128 LUT case, split BF16)", "cases": cases}}
    print(json.dumps(out), flush=True)


def localize_bench(a, name):
    """The localization chain through the tracker's device batch: realtime100 (the headline:
    4,096 ROIs of 100x100 u16, 256 beads with 200 x 64 LUTs) or plain64 (the CPU-runnable
    case: 1,000 ROIs of 64x64 f32, one bead with a 100 x 32 LUT)."""
    from qtrk_amd import api, synth

    realtime = name == "realtime100"
    n_req = a.rois if a.rois else (4096 if realtime else 1000)
    w = synth.make_workload(name, n=n_req)
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["luts"])
    trk.set_localization_mode(w["mode"])
    n = len(w["jobs"])
    b = api.DeviceBatch(trk, n, api.PDT_OF_DTYPE[w["rois"].dtype])
    b.upload(w["rois"], w["jobs"])
    # the GPU's clocks settle during the first tens of milliseconds of load: a streaming tracker runs
    # at the settled rate, so batches run for PREWARM_S before the warm-up steps and the timed ones
    prewarm(b, n)
    b.run(n, a.warmup)
    b.sync()

    # steady state: the steps are consecutive batches, launched on two alternating streams so that
    # one batch's first workgroups fill the CUs the previous batch's last ones leave idle
    t0 = time.perf_counter()
    b.run(n, a.steps)
    b.sync()
    t1 = time.perf_counter()
    res = b.download(n)
    wall = t1 - t0
    ms_per_step = wall / a.steps * 1e3
    value = n * a.steps / wall
    # one launch's duration, for the rooflines (what a kernel trace reports per dispatch)
    iso = []
    for _ in range(min(a.steps, 20)):
        b.run(n, 1)
        iso.append(b.sync())
    kern_s = float(np.mean(iso)) / 1e3
    H, W = w["rois"].shape[1:]
    # ROI + the bead's LUT shared by the bead's ROIs + job + result
    lut_bytes = w["luts"][0].size * 4
    bytes_per_loc = BYTES_PER_LOC if realtime else W * H * w["rois"].itemsize + lut_bytes / n + 24 + 52
    achieved = n * bytes_per_loc / kern_s / 1e9
    err = {k: float(np.sqrt(np.mean((res[k] - w["truth"][:, i]) ** 2))) for i, k in enumerate("xyz")}
    cfg = w["cfg"]
    workload = (f"{name}: {n} ROIs of {W}x{H} {np.dtype(w['rois'].dtype).name}; COM + QI ({cfg.qi_iterations} "
                f"iterations, {cfg.qi_radialsteps} radial x {cfg.qi_angstepspq} angular/quadrant) + "
                f"{'per-bead ' if realtime else ''}Z-LUT ({w['luts'].shape[1]} planes x {w['luts'].shape[2]} steps) "
                "with 7-point FP64 fit; ROIs resident in HBM")
    out = {
        "metric": METRIC if realtime else "localizations/sec (COM+QI+Z-LUT, 64×64 f32 ROI); px/plane error",
        "value": round(value, 1),
        "unit": "localizations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / PUBLISHED_CPU_LOCS, 1) if realtime else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "rois_per_step": n, "roi": [W, H],
                   "pixel": np.dtype(w["rois"].dtype).name, "parallelism": "1 GPU",
                   "rms_error_vs_truth": {"x_px": round(err["x"], 4), "y_px": round(err["y"], 4),
                                          "z_planes": round(err["z"], 4)}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic() if realtime else None,
                     "kernel_ms": round(kern_s * 1e3, 4),
                     "note": "kernel_ms: one launch on its own; value: consecutive batches, whose launches overlap"},
    }
    valu = _profile_value(SQ_PROFILE, "SQ_INSTS_VALU")
    if realtime and valu and n == 4096:
        # the sampling stages are VALU-issue bound: instructions per launch from the counter summary
        out["roofline"]["valu_issue"] = {"wave_instr_per_launch": valu,
                                         "achieved": round(valu / kern_s / 1e9, 1),
                                         "peak": VALU_PEAK_WAVE_INSTR / 1e9, "unit": "G wave-instr/s",
                                         "frac": round(valu / kern_s / VALU_PEAK_WAVE_INSTR, 4)}
    if realtime:
        # the batched forms on the same batch (QTrkBatchSetMatchPrecision), consecutive batches as
        # above, against the per-ROI FP32 results: the profile kernel, the ROIs grouped by bead on
        # the GPU, then each bead's profiles matched together -- in the reference arithmetic
        # (FP32 grouped: the plain-FP32 results, bit for bit) or on the matrix cores in BF16 with
        # the FP32 rescoring of the top 8 (mixed precision)
        def batched(prec):
            b.set_match_precision(prec)
            b.run(n, a.warmup)
            b.sync()
            t0 = time.perf_counter()
            b.run(n, a.steps)
            b.sync()
            bw = time.perf_counter() - t0
            r = b.download(n)
            b.set_match_precision(api.MATCH_FP32)
            same = int(np.sum((r["x"] == res["x"]) & (r["y"] == res["y"]) & (r["z"] == res["z"])))
            return {"value": round(n * a.steps / bw, 1), "unit": "localizations/s",
                    "ms_per_step": round(bw / a.steps * 1e3, 4), "speedup_vs_fp32": round(wall / bw, 4),
                    "results_identical_to_fp32": same, "rois": n}

        out["fp32_grouped"] = dict({"form": "profile kernel + grouped Z-LUT match in the reference arithmetic "
                                            "(QTRK_MATCH_FP32_GROUPED)"}, **batched(api.MATCH_FP32_GROUPED))
        out["mixed_precision"] = dict({"form": "profile kernel + grouped BF16 matrix-core Z-LUT match, FP32 top-8 "
                                               "rescoring, FP64 fit"}, **batched(api.MATCH_BF16))
    b.free()
    if realtime and n == 4096 and not a.no_frames:
        # the same batch from host memory through the drop-in C ABI (PCIe included): reported
        # beside the device-resident value, never as it
        frames, pos, jobs = _camera_frames(w)
        rates, _ = frame_path_rates(trk, frames, pos, jobs, 10, 3)
        out["pcie_inclusive"] = dict({"path": "QTrkScheduleFrame: 16 host frames of 1600x1600 u16, 256 ROIs each",
                                      "unit": "localizations/s"}, **rates)
    trk.free()
    if realtime and n == 4096 and not a.no_frames:
        # the reference's speed-test pattern from a C client: one QTrkScheduleFrame call per ROI
        st, sres = c_abi_speed_test(w, "frame1", 5)
        st["results_equal_device_batch"] = bool(np.array_equal(sres["x"], res["x"]) and np.array_equal(sres["z"], res["z"]))
        out["pcie_inclusive"]["one_roi_per_call_c_client"] = st
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
