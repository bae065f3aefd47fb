"""Parity of the MI355X path with the CPU oracle, through the C ABI.

Bars (the reference's own tolerances for this path):
  * ROI extraction, indexing and the best LUT plane: bit-exact;
  * x/y within 1e-3 px and z within 1e-3 planes of the oracle (in practice the
    kernel reproduces the oracle's float rounding, so most values are identical);
  * every result's error against the FP64 oracle is printed.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth

pytestmark = pytest.mark.gpu

XY_TOL = 1e-3   # px
Z_TOL = 1e-3    # LUT planes


def _tracker(w, mode=None, luts=True):
    trk = api.QTrkInstance(w["settings"])
    if luts:
        trk.set_radial_zlut(w["luts"])
    trk.set_localization_mode(w["mode"] if mode is None else mode)
    return trk


def _batch(trk, rois, jobs, diag=True):
    b = api.DeviceBatch(trk, len(jobs), api.PDT_OF_DTYPE[rois.dtype])
    if diag:
        b.enable_diagnostics(True)
    b.upload(rois, jobs)
    b.run(len(jobs), 1)
    b.sync()
    res = b.download(len(jobs))
    best, prof = b.diagnostics(len(jobs)) if diag else (None, None)
    b.free()
    return res, best, prof


def _oracle(oracle, w, mode=None, rois=None, jobs=None, luts="default", zcmp=None):
    cfg = oracle.config_from_bytes(bytes(w["cfg"]))
    return oracle.localize_trace(cfg, w["mode"] if mode is None else mode, w["luts"] if luts == "default" else luts,
                                 w["rois"] if rois is None else rois, w["jobs"] if jobs is None else jobs, zcmp=zcmp)


def _report(tag, got, ref):
    n = len(got)
    parts = []
    for k in ("comx", "comy", "x", "y", "z"):
        same = int((got[k] == ref[k]).sum())
        d = float(np.abs(got[k].astype(np.float64) - ref[k]).max()) if n else 0.0
        parts.append(f"{k}: {same}/{n} identical, max|d|={d:.2e}")
    print(f"[{tag}] " + "; ".join(parts), flush=True)


def _close(a, b, tol):
    """|a - b| <= tol elementwise, with NaN matching NaN (a flat ROI normalises 0/0 in both)."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(both_nan | (np.abs(a - b) <= tol)))


def _assert_parity(got, ref, best=None, rbest=None, check_z=True):
    assert np.array_equal(got["frame"], ref["frame"]) and np.array_equal(got["zlutIndex"], ref["zlutIndex"])
    assert np.array_equal(got["timestamp"], ref["timestamp"])
    for k in ("x", "y", "comx", "comy"):
        assert _close(got[k], ref[k], XY_TOL), k
    assert np.array_equal(got["error"], ref["error"])
    np.testing.assert_allclose(got["imageMean"], ref["imageMean"], rtol=1e-6)
    if check_z:
        assert _close(got["z"], ref["z"], Z_TOL)
    if best is not None:
        assert np.array_equal(best, rbest)


@pytest.mark.parametrize("name,n", [("plain64", 1000), ("realtime100", 1024)])
def test_workload_parity(gpu, oracle, oracle64, name, n):
    w = synth.make_workload(name, n=n)
    ref, rbest, rprof = _oracle(oracle, w)
    trk = _tracker(w)
    got, best, prof = _batch(trk, w["rois"], w["jobs"])
    trk.free()
    _report(name, got, ref)
    _assert_parity(got, ref, best, rbest)
    np.testing.assert_allclose(prof, rprof, atol=1e-5)
    r64, _, _ = oracle64.localize_trace(oracle64.config_from_bytes(bytes(w["cfg"])), w["mode"], w["luts"],
                                        w["rois"], w["jobs"])
    e = {k: float(np.abs(got[k] - r64[k]).max()) for k in "xyz"}
    t = {k: float(np.sqrt(np.mean((got[k] - w["truth"][:, i]) ** 2))) for i, k in enumerate("xyz")}
    print(f"[{name}] max |GPU - FP64 oracle|: x {e['x']:.2e} px, y {e['y']:.2e} px, z {e['z']:.2e} planes; "
          f"rms vs truth x {t['x']:.4f} y {t['y']:.4f} z {t['z']:.4f}", flush=True)
    assert e["x"] < XY_TOL and e["y"] < XY_TOL and e["z"] < Z_TOL


def test_workload_parity_angstep_factor(gpu, oracle):
    """realtime100 at qi_angstep_factor = 1.3 (the reference's own benchmark setting,
    cputrack-test/Benchmark.cpp:25): per iteration int(A / 1.3^(iters - k)) angular samples
    (clamped to >= 4) at table stride (int)(A / samples) (cputrack/cpu_tracker.cpp:314, 356,
    642-657); bit-exact indexing and best plane, x/y/z within the bars."""
    w = synth.make_workload("realtime100", n=512, qi_angstep_factor=1.3)
    ref, rbest, rprof = _oracle(oracle, w)
    trk = _tracker(w)
    got, best, prof = _batch(trk, w["rois"], w["jobs"])
    trk.free()
    _report("realtime100 factor 1.3", got, ref)
    _assert_parity(got, ref, best, rbest)
    np.testing.assert_allclose(prof, rprof, atol=1e-5)


@pytest.mark.parametrize("cov,factor", [(0.5, 1.0), (2.0, 1.3)], ids=["16_radial", "128_radial_factor1.3"])
def test_qi_radial_steps_parity(gpu, oracle, cov, factor):
    """100x100 ROIs (the instance reading the QI sample offsets from the tracker's table) with 16
    QI radial steps (64 (quadrant, radius) slots: slots not in 32-radius segments, three wavefronts
    without one) and with 128 (512 slots: two passes over the workgroup) at qi_angstep_factor 1.3;
    bit-exact indexing and best plane, x/y/z within the bars, profiles equal."""
    w = synth.make_workload("realtime100", n=512, qi_radial_coverage=cov, qi_angstep_factor=factor)
    nr = api.computed_config(w["settings"]).qi_radialsteps
    assert nr == (16 if cov < 1 else 128)
    ref, rbest, rprof = _oracle(oracle, w)
    trk = _tracker(w)
    got, best, prof = _batch(trk, w["rois"], w["jobs"])
    trk.free()
    _report(f"realtime100 {nr} QI radial steps", got, ref)
    _assert_parity(got, ref, best, rbest)
    np.testing.assert_allclose(prof, rprof, atol=1e-5)


@pytest.mark.parametrize("spec", synth.BATCHING, ids=[f"{s[0]}px" for s in synth.BATCHING])
def test_batching_tracker_sizes_parity(gpu, oracle, spec):
    """The batching workload's four trackers (ROI 32/64/96/128 px, LUTs of 100/150/200/300 planes x
    32/64/96/96 steps, uint16 Poisson ROIs) through the tracker's per-ROI kernel, 256 ROIs each,
    against the oracle's ProcessJob (cputrack/QueuedCPUTracker.cpp:273-361)."""
    size, _, beads, planes, steps = spec
    w = synth.batching_tracker(size, 256, min(beads, 64), planes, steps, np.random.default_rng(0x51AB0004 + size))
    wl = dict(w, mode=synth.MODE_QI_Z, luts=w["lut"])
    ref, rbest, _ = _oracle(oracle, wl)
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["lut"])
    trk.set_localization_mode(synth.MODE_QI_Z)
    got, best, _ = _batch(trk, w["rois"], w["jobs"])
    trk.free()
    _report(f"batching {size}px", got, ref)
    _assert_parity(got, ref, best, rbest)


@pytest.mark.parametrize("prec", [api.MATCH_BF16, api.MATCH_FP16, api.MATCH_FP32_GROUPED],
                         ids=["bf16", "fp16", "fp32_grouped"])
def test_batch_grouped_match_parity(gpu, oracle, prec):
    """A device batch with the grouped Z-LUT match (QTrkBatchSetMatchPrecision): the localization
    kernel in profile-only form, the ROIs grouped by bead on the GPU, each bead's profiles matched
    together -- on the matrix cores with the 8 best planes and the fit window rescored in the
    reference arithmetic (bf16, fp16), or every plane in the reference arithmetic (fp32_grouped,
    whose results must also equal the fused FP32 form's byte for byte).  Against the oracle's
    ProcessJob: indexing and best plane bit-exact, x/y/z
    within the bars, profiles equal, with one bead holding 200 of the 1,024 ROIs (its workgroups
    loop over several profile tiles) and ROIs of a bead outside the LUT (z 0, best plane -1, as in
    the per-ROI form); after SetRadialWeights the next run matches with the new weights."""
    w = synth.make_workload("realtime100", n=1024)
    jobs = w["jobs"].copy()
    jobs["zlutIndex"][:200] = 3
    jobs["zlutIndex"][200:210] = 5000
    trk = _tracker(w)
    b = api.DeviceBatch(trk, 1024, api.QTrkU16)
    b.set_match_precision(prec)
    b.enable_diagnostics(True)
    b.upload(w["rois"], jobs)
    for weights in (None, np.linspace(0.5, 1.5, trk.cfg.zlut_radialsteps).astype(np.float32)):
        if weights is not None:
            trk.set_radial_weights(weights)
        b.run(1024, 2)
        b.sync()
        got = b.download(1024)
        best, prof = b.diagnostics(1024)
        ref, rbest, rprof = _oracle(oracle, w, jobs=jobs, zcmp=weights)
        _report(f"batch grouped match {'weighted' if weights is not None else 'plain'}", got, ref)
        _assert_parity(got, ref, best, rbest)
        np.testing.assert_allclose(prof, rprof, atol=1e-5)
        assert np.all(got["z"][200:210] == 0) and np.all(best[200:210] == -1)
        if prec == api.MATCH_FP32_GROUPED:
            b.set_match_precision(api.MATCH_FP32)
            b.run(1024, 2)
            b.sync()
            fused = b.download(1024)
            fbest, _ = b.diagnostics(1024)
            b.set_match_precision(prec)
            assert np.array_equal(got.view(np.uint8), fused.view(np.uint8))
            assert np.array_equal(best, fbest)
    b.free()
    trk.free()


def test_batch_grouped_match_many_rois(gpu):
    """The grouped forms on a batch of 10,240 ROIs (more than the grouping kernel's 8,192 rows per
    pass, the last pass partial), beads of 40 to 200 ROIs and ROIs of a bead outside the LUT: every
    record byte-equal to the fused FP32 form's."""
    w = synth.make_workload("realtime100", n=1024)
    rois = np.tile(w["rois"], (10, 1, 1))
    jobs = np.tile(w["jobs"], 10)
    jobs["frame"] = np.arange(len(jobs))
    jobs["zlutIndex"][::7] = 11
    jobs["zlutIndex"][5::301] = 9999
    trk = _tracker(w)
    b = api.DeviceBatch(trk, len(jobs), api.QTrkU16)
    b.enable_diagnostics(True)
    b.upload(rois, jobs)
    try:
        out = {}
        for prec in (api.MATCH_FP32, api.MATCH_BF16, api.MATCH_FP32_GROUPED):
            print(f"grouped match, {len(jobs)} ROIs, precision {prec}", flush=True)
            b.set_match_precision(prec)
            b.run(len(jobs), 1)
            b.sync()
            out[prec] = (b.download(len(jobs)), b.diagnostics(len(jobs))[0])
        ref, rbest = out[api.MATCH_FP32]
        assert np.all(rbest[5::301] == -1)
        for prec in (api.MATCH_BF16, api.MATCH_FP32_GROUPED):
            got, best = out[prec]
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), prec
            assert np.array_equal(best, rbest), prec
    finally:
        b.free()
        trk.free()


@pytest.mark.parametrize("mp", ["bf16", "fp32_grouped"])
def test_queue_match_precision_grouped(gpu, oracle, mp):
    """match_precision = bf16 / fp32_grouped: queued batches, from ScheduleLocalization and from
    ScheduleFrame's pitched frame slots, run the profile kernel and the grouped match (matrix-core
    or reference arithmetic); results equal the oracle's ProcessJob, also after the radial weights
    change between batches."""
    w = synth.make_workload("realtime100", n=300)
    trk = _tracker(w)
    trk.set_config_value("batch_size", "128")
    trk.set_config_value("match_precision", mp)
    assert trk.get_config_values()["match_precision"] == mp
    got = trk.localize_all(w["rois"], w["jobs"])
    ref, _, _ = _oracle(oracle, w)
    _report(f"queue {mp}", got, ref)
    _assert_parity(got, ref)
    wts = np.linspace(0.3, 1.7, trk.cfg.zlut_radialsteps).astype(np.float32)
    trk.set_radial_weights(wts)
    frame = np.zeros((230, 1030), np.uint16)
    big = np.zeros((230, 1100), np.uint16)
    frame = big[:, :1030]
    pos = [(3 + 102 * (i % 10), 10 + 110 * (i // 10)) for i in range(20)]
    for i, (x, y) in enumerate(pos):
        frame[y:y + 100, x:x + 100] = w["rois"][i]
    job = api.make_jobs(1, frame=np.array([900]), zlut_index=np.array([0]))[0]
    assert trk.schedule_frame(frame, pos, job) == 20
    trk.wait_for(20)
    r = _by_index(trk.fetch_results(20))
    trk.free()
    ref = _oracle_frames(oracle, w, w["luts"], [(frame, pos, job)], zcmp=wts)
    _report(f"queue {mp} frames weighted", r, ref)
    _assert_parity(r, ref)
    with pytest.raises(api.QTrkError, match="match_precision"):
        t2 = _tracker(w)
        try:
            t2.set_config_value("match_precision", "int8")
        finally:
            t2.free()


def test_queue_api_matches_batch_and_is_lifo(gpu, oracle):
    w = synth.make_workload("plain64", n=300)
    trk = _tracker(w)
    batch, _, _ = _batch(trk, w["rois"], w["jobs"], diag=False)
    got = trk.localize_all(w["rois"], w["jobs"])
    assert np.array_equal(got, batch)
    # LIFO fetch like the CPU tracker (cputrack/QueuedCPUTracker.cpp:458-469)
    for i in range(10):
        trk.schedule_localization(w["rois"][i], w["jobs"][i])
    trk.wait_for(10)
    r = trk.fetch_results(10)
    assert list(r["frame"]) == list(range(9, -1, -1))
    assert trk.is_idle() and trk.result_count() == 0
    trk.free()


def test_is_idle_fills_the_return_register(gpu):
    """QTrkIsIdle read as a 4-byte integer, as C#'s default bool marshalling reads it
    (QTrkDotNet/QTrkDLL.cs): exactly 1 for an idle tracker, 0 or 1 while busy."""
    import ctypes as C
    w = synth.make_workload("realtime100", n=256)
    trk = api.QTrkInstance(w["settings"])
    try:
        trk.set_radial_zlut(w["luts"])
        trk.set_localization_mode(w["mode"])
        as_int = C.CFUNCTYPE(C.c_int, C.c_void_p)(C.cast(api.lib().QTrkIsIdle, C.c_void_p).value)
        assert as_int(trk.h) == 1
        for i in range(len(w["jobs"])):
            trk.schedule_localization(w["rois"][i], w["jobs"][i])
        assert as_int(trk.h) in (0, 1)
        trk.wait_for(len(w["jobs"]))
        assert as_int(trk.h) == 1
    finally:
        trk.free()


def test_small_batches_many_streams_block_and_complete(gpu, oracle):
    """Batches of 8 on 2 streams: ScheduleLocalization blocks while 16 ROIs are queued."""
    w = synth.make_workload("plain64", n=200)
    trk = _tracker(w)
    trk.set_config_value("batch_size", "8")
    trk.set_config_value("num_streams", "2")
    assert trk.get_config_values()["batch_size"] == "8"
    for i in range(200):
        trk.schedule_localization(w["rois"][i], w["jobs"][i])
        q, qmax = trk.queue_length()
        assert qmax == 16 and q <= qmax
    trk.wait_for(200)
    r = trk.fetch_results(200)
    r = r[np.argsort(r["frame"])]
    assert list(r["frame"]) == list(range(200))
    ref, _, _ = _oracle(oracle, w)
    _assert_parity(r, ref)
    trk.free()


def test_concurrent_schedulers_share_a_small_queue(gpu, oracle):
    """Threads scheduling into one tracker at once (the reference guards its queue with a mutex,
    cputrack/QueuedCPUTracker.cpp:435-456, cudatrack/QueuedCUDATracker.cu:431-446): batches of 8
    on 2 streams, so the threads block on the full queue and find each other's filling slot when
    they wake; two threads schedule single ROIs and one whole frames, while the main thread
    changes num_streams and batch_size (the slots are remade with threads waiting).  Every ROI's
    result arrives exactly once, the tracker becomes idle, and every result matches the oracle."""
    import threading
    import time

    w = synth.make_workload("plain64", n=400)
    trk = _tracker(w)
    trk.set_config_value("batch_size", "8")
    trk.set_config_value("num_streams", "2")
    errors = []
    frames = []
    for k in range(25):  # frames of 4 ROIs: ROIs 300 + 4k .. 303 + 4k
        f = np.zeros((70, 280), np.float32)
        for i in range(4):
            f[3:67, 4 + 68 * i:68 + 68 * i] = w["rois"][300 + 4 * k + i]
        frames.append((f, [(4 + 68 * i, 3) for i in range(4)],
                       api.make_jobs(1, frame=np.array([1000 + k]), zlut_index=np.array([0]))[0]))

    def rois(lo, hi):
        try:
            for i in range(lo, hi):
                trk.schedule_localization(w["rois"][i], w["jobs"][i])
        except Exception as e:  # noqa: BLE001 - reported by the main thread
            errors.append(e)

    def frame_thread():
        try:
            for f, pos, job in frames:
                assert trk.schedule_frame(f, pos, job) == 4
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    ths = [threading.Thread(target=rois, args=(0, 150)), threading.Thread(target=rois, args=(150, 300)),
           threading.Thread(target=frame_thread)]
    for t in ths:
        t.start()
    time.sleep(0.002)
    trk.set_config_value("num_streams", "3")
    trk.set_config_value("batch_size", "5")
    for t in ths:
        t.join(60)
        assert not t.is_alive()
    assert not errors, errors
    trk.wait_for(400, timeout=60)
    t0 = time.time()
    while not trk.is_idle():
        assert time.time() - t0 < 10, "tracker never became idle"
        time.sleep(0.001)
    r = trk.fetch_results(1000)
    trk.free()
    assert len(r) == 400
    got = _by_index(r)
    keys = set(zip(got["frame"].tolist(), got["zlutIndex"].tolist()))
    assert len(keys) == 400
    calls = [(w["rois"][i], w["jobs"][i]) for i in range(300)] + frames
    ref = _oracle_frames(oracle, w, w["luts"], calls)
    _report("concurrent", got, ref)
    _assert_parity(got, ref)


def test_pixel_formats_and_pitch(gpu, oracle):
    w = synth.make_workload("plain64", n=64)
    scale = w["rois"] / w["rois"].max()
    for dt, vmax in ((np.uint8, 255), (np.uint16, 60000), (np.float32, 1.0)):
        rois = (scale * vmax).astype(dt)
        ref, rbest, _ = _oracle(oracle, w, rois=rois)
        trk = _tracker(w)
        got, best, _ = _batch(trk, rois, w["jobs"])
        _assert_parity(got, ref, best, rbest)
        # a padded source pitch through ScheduleLocalization gives the same answer
        padded = np.zeros((rois.shape[0], rois.shape[1], rois.shape[2] + 7), dt)
        padded[:, :, :rois.shape[2]] = rois
        for i in range(8):
            trk.schedule_localization(padded[i][:, :rois.shape[2]], w["jobs"][i])
        trk.wait_for(8)
        r = trk.fetch_results(8)
        r = r[np.argsort(r["frame"])]
        assert np.array_equal(r, got[:8])
        trk.free()


@pytest.mark.parametrize("size", [(45, 45), (50, 38)], ids=["45x45", "50x38"])
def test_odd_sizes_staging(gpu, oracle, size):
    """ROI byte counts that are not multiples of 16: in one contiguous batch some ROIs start
    16-byte aligned (read in 16-byte chunks plus a tail) and others do not (read pixel by
    pixel); every format, with ClearFirstFourPixels, against the oracle."""
    w = synth.make_workload("plain64", n=48, width=size[0], height=size[1])
    scale = w["rois"] / w["rois"].max()
    for dt, vmax in ((np.uint8, 255), (np.uint16, 60000), (np.float32, 1.0)):
        rois = np.ascontiguousarray((scale * vmax).astype(dt))
        for mode in (w["mode"], w["mode"] | api.LT_ClearFirstFourPixels):
            ref, rbest, _ = _oracle(oracle, w, mode=mode, rois=rois)
            trk = _tracker(w, mode=mode)
            got, best, _ = _batch(trk, rois, w["jobs"])
            _report(f"{size} {np.dtype(dt).name} mode {mode}", got, ref)
            _assert_parity(got, ref, best, rbest)
            trk.free()


def test_mode_variants(gpu, oracle):
    w = synth.make_workload("plain64", n=128)
    for mode in (api.LT_OnlyCOM, api.LT_QI, api.LT_QI | api.LT_LocalizeZ, api.LT_LocalizeZ | api.LT_NormalizeProfile,
                 w["mode"] | api.LT_ClearFirstFourPixels):
        ref, rbest, _ = _oracle(oracle, w, mode=mode)
        trk = _tracker(w, mode=mode)
        got, best, _ = _batch(trk, w["rois"], w["jobs"])
        trk.free()
        _report(f"mode {mode}", got, ref)
        _assert_parity(got, ref, best, rbest)


def test_radial_weights_and_missing_lut(gpu, oracle):
    w = synth.make_workload("realtime100", n=256)
    zcmp = np.linspace(0.2, 1.5, w["cfg"].zlut_radialsteps).astype(np.float32)
    ref, rbest, _ = _oracle(oracle, w, zcmp=zcmp)
    trk = _tracker(w)
    trk.set_radial_weights(zcmp)
    got, best, _ = _batch(trk, w["rois"], w["jobs"])
    _assert_parity(got, ref, best, rbest)
    trk.set_radial_weights(None)
    # bead indices beyond the LUT: z = 0, no compare (the reference would read out of bounds)
    jobs = w["jobs"].copy()
    jobs["zlutIndex"][:32] = 5000
    ref2, rb2, _ = _oracle(oracle, w, jobs=jobs)
    got2, b2, _ = _batch(trk, w["rois"], jobs)
    _assert_parity(got2, ref2, b2, rb2)
    assert np.all(got2["z"][:32] == 0) and np.all(b2[:32] == -1)
    trk.free()
    # no LUT at all: LocalizeZ still computes the profile, z = 0
    trk = _tracker(w, luts=False)
    got3, b3, _ = _batch(trk, w["rois"], w["jobs"])
    assert np.all(got3["z"] == 0) and np.all(b3 == -1)
    trk.free()


def test_edge_cases_constant_and_boundary(gpu, oracle):
    w = synth.make_workload("plain64", n=16)
    rois = w["rois"].copy()
    rois[0] = 123.0                     # constant ROI: COM 0/0 -> centre fallback
    rois[1] = 0.0                       # all zero
    rois[2][:, :] = np.arange(64, dtype=np.float32)[None, :]   # ramp: COM far off-centre -> boundary hits
    rois[3][30:34, 1:3] = 50000         # bright spot at the edge
    ref, rbest, _ = _oracle(oracle, w, rois=rois)
    trk = _tracker(w)
    got, best, _ = _batch(trk, rois, w["jobs"])
    trk.free()
    _report("edge", got, ref)
    _assert_parity(got, ref, best, rbest)
    assert (got["comx"][0], got["comy"][0]) == (32.0, 32.0)
    assert got["error"].sum() >= 1


def test_calibration_images(gpu, oracle):
    w = synth.make_workload("realtime100", n=64)
    cfg = w["cfg"]
    nb = w["luts"].shape[0]
    rng = np.random.default_rng(7)
    off = rng.normal(0, 5, (nb, cfg.height, cfg.width)).astype(np.float32)
    gain = rng.normal(1, 0.05, (nb, cfg.height, cfg.width)).astype(np.float32)
    o_cfg = oracle.config_from_bytes(bytes(cfg))
    # (v + offset) * gain with unit factors (cputrack/cpu_tracker.cpp:108-111), applied to the
    # float pixels before the pipeline -- the same operations in the same order as the oracle's
    rois_f = w["rois"].astype(np.float32)
    corr = (rois_f + off[w["jobs"]["zlutIndex"]]) * gain[w["jobs"]["zlutIndex"]]
    ref, rbest, _ = oracle.localize_trace(o_cfg, w["mode"], w["luts"], corr.astype(np.float32), w["jobs"])
    trk = _tracker(w)
    trk.set_pixel_calibration_images(off, gain)
    got, best, _ = _batch(trk, w["rois"], w["jobs"])
    trk.free()
    _assert_parity(got, ref, best, rbest)


def _by_index(r):
    return r[np.lexsort((r["zlutIndex"], r["frame"]))]


def _oracle_frames(oracle, w, luts, calls, mode=None, zcmp=None):
    """The oracle's QueuedTracker::ScheduleFrame on each call (frame, positions, job) or per-ROI call
    (roi array, job), then ProcessJob on the ROIs it queued; results sorted by (frame, zlutIndex)."""
    cfg = oracle.config_from_bytes(bytes(w["cfg"]))
    out = []
    for c in calls:
        if len(c) == 2:
            rois, jobs = np.asarray(c[0])[None], np.asarray(c[1], api.JOB_DTYPE).reshape(1)
        else:
            rois, jobs = oracle.schedule_frame(cfg, c[0], c[1], c[2])
        if len(jobs):
            out.append(oracle.localize_trace(cfg, w["mode"] if mode is None else mode, luts, rois, jobs,
                                             zcmp=zcmp)[0])
    return _by_index(np.concatenate(out))


def test_schedule_frame_extraction(gpu, oracle):
    """QTrkScheduleFrame against the oracle's ScheduleFrame (cputrack/QueuedTracker.cpp:91-111):
    ROIs cut from the frame at their positions, those outside skipped (including one a pixel
    past the right and bottom edges; ROIs flush with them are kept), zlutIndex = i + base."""
    w = synth.make_workload("plain64", n=14)
    H, W = 300, 410
    frame = np.zeros((H, W), np.float32)
    pos = []
    for i in range(12):
        x, y = 10 + (i % 4) * 95, 20 + (i // 4) * 90
        frame[y:y + 64, x:x + 64] = w["rois"][i]
        pos.append((x, y))
    frame[H - 64:, W - 64:] = w["rois"][12]
    frame[0:64, 0:64] = w["rois"][13]
    pos += [(-1, 5), (W - 63, 0), (0, H - 63), (W - 64, H - 64), (0, 0)]  # 3 skipped, 2 flush kept
    trk = _tracker(w)
    job = np.zeros(1, api.JOB_DTYPE)[0]
    job["frame"] = 77
    job["zlutIndex"] = 0
    n = trk.schedule_frame(frame, pos, job)
    assert n == 14
    trk.wait_for(14)
    r = _by_index(trk.fetch_results(14))
    trk.free()
    assert list(r["zlutIndex"]) == list(range(12)) + [15, 16] and np.all(r["frame"] == 77)
    ref = _oracle_frames(oracle, w, w["luts"], [(frame, pos, job)])
    _report("frame extraction", r, ref)
    _assert_parity(r, ref)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32], ids=["u8", "u16", "f32"])
def test_schedule_frame_chunks_pitch_and_formats(gpu, oracle, dtype):
    """The frame is uploaded once per batch_size ROIs and the kernel reads each ROI at its offset in
    it: every result matches the oracle's ScheduleFrame + ProcessJob, with a padded frame pitch,
    ROIs outside the frame skipped, zlutIndex = i + the job's zlutIndex, and more ROIs than one
    batch."""
    n, base = 40, 5
    w = synth.make_workload("realtime100", n=n)
    rois = w["rois"]
    if dtype == np.uint8:
        rois = np.clip(rois // 64, 0, 255).astype(np.uint8)
    rois = rois.astype(dtype)
    H, W = 720, 1100
    big = np.zeros((H, W + 37), dtype)  # rows padded past the frame's width
    frame = big[:, :W]
    pos = []
    for i in range(n):
        x, y = 3 + (i % 10) * 109, 7 + (i // 10) * 175
        frame[y:y + 100, x:x + 100] = rois[i]
        pos.append((x, y))
    pos = pos[:20] + [(W - 99, 0), (-2, 40)] + pos[20:] + [(0, H - 99)]  # 3 outside, interleaved
    keep = [i for i, p in enumerate(pos) if p[0] >= 0 and p[1] >= 0 and p[0] + 100 <= W and p[1] + 100 <= H]
    nb = base + len(pos)
    luts = np.concatenate([w["luts"]] * (nb // len(w["luts"]) + 1))[:nb]
    trk = api.QTrkInstance(w["settings"])
    trk.set_config_value("batch_size", "16")
    trk.set_radial_zlut(luts)
    trk.set_localization_mode(w["mode"])
    job = np.zeros(1, api.JOB_DTYPE)[0]
    job["frame"] = 9
    job["zlutIndex"] = base
    assert frame.strides[0] == (W + 37) * np.dtype(dtype).itemsize
    assert trk.schedule_frame(frame, pos, job) == n
    trk.flush()
    trk.wait_for(n)
    r = trk.fetch_results(n)
    r = r[np.argsort(r["zlutIndex"])]
    assert list(r["zlutIndex"]) == [base + i for i in keep]
    trk.free()
    ref = _oracle_frames(oracle, w, luts, [(frame, pos, job)])
    _report(f"frame chunks {np.dtype(dtype).name}", r, ref)
    _assert_parity(r, ref)


def test_schedule_frame_packs_frames_into_batches(gpu, oracle):
    """Frames of one pixel type and pitch share a batch slot, their row blocks one after another,
    until it holds batch_size ROIs (the rest goes on to the next slot) or a block would outgrow
    it; a per-ROI call, another pitch or another pixel type launches the filling slot first.
    Every result matches the oracle's ScheduleFrame + ProcessJob on the same calls."""
    w = synth.make_workload("realtime100", n=40)
    src = w["rois"]
    src8 = np.clip(src // 64, 0, 255).astype(np.uint8)
    grid = lambda n, dx, dy, x0=0, y0=0: [(x0 + dx * (i % 6), y0 + dy * (i // 6)) for i in range(n)]
    # (dtype, frame height, width, pitch padding in pixels, ROI positions) or ("roi", count)
    steps = [(np.uint16, 260, 700, 0, grid(12, 110, 130, 10, 5)),  # cap: 2 blocks of this size
             (np.uint16, 260, 700, 0, grid(7, 110, 120, 30, 20)),  # 4 fill the batch, 3 go on
             ("roi", 2),                                          # launches the 3 first
             (np.uint16, 900, 700, 0, [(5, 0), (200, 400), (400, 790)]),
             (np.uint16, 240, 900, 16, grid(8, 110, 120)),        # other pitch: launches
             (np.uint16, 900, 900, 16, [(0, 0), (300, 400), (600, 790)]),  # outgrows: launches
             (np.uint8, 300, 400, 0, [(0, 0), (150, 100), (290, 190)])]   # other type: launches
    luts = np.concatenate([w["luts"]] * 2)
    trk = api.QTrkInstance(w["settings"])
    trk.set_config_value("batch_size", "16")
    trk.set_radial_zlut(luts)
    trk.set_localization_mode(w["mode"])
    k, calls = 0, []
    for f, st in enumerate(steps):
        if st[0] == "roi":
            for _ in range(st[1]):
                job = api.make_jobs(1, frame=np.array([f]), zlut_index=np.array([k]))[0]
                trk.schedule_localization(src[k], job)
                calls.append((src[k], job))
                k += 1
            continue
        dtype, H, W, pad, pos = st
        big = np.zeros((H, W + pad), dtype)
        frame = big[:, :W]
        for i, (x, y) in enumerate(pos):
            assert x + 100 <= W and y + 100 <= H
            frame[y:y + 100, x:x + 100] = (src8 if dtype == np.uint8 else src)[k + i]
        job = api.make_jobs(1, frame=np.array([f]), zlut_index=np.array([k]))[0]
        assert trk.schedule_frame(frame, pos, job) == len(pos)
        calls.append((frame, pos, job))
        k += len(pos)
    trk.wait_for(k)
    r = _by_index(trk.fetch_results(k))
    trk.free()
    assert sorted(r["zlutIndex"]) == list(range(k))
    ref = _oracle_frames(oracle, w, luts, calls)
    _report("frames packed", r, ref)
    _assert_parity(r, ref)


def test_schedule_frame_one_roi_frames(gpu, oracle):
    """The reference's speed-test pattern (cudatrack-test/test.cu:343-347): every ROI is a frame of
    its own, scheduled by one call.  Small row blocks go to the GPU together once 1 MiB is waiting
    or when the slot launches: with batch_size 256, the first slot's 256 blocks of 20 KB go in
    several copies and the second slot's 44 only at Flush.  Every result matches the oracle."""
    w = synth.make_workload("realtime100", n=300)
    src = w["rois"]
    luts = np.concatenate([w["luts"]] * 2)
    trk = api.QTrkInstance(w["settings"])
    trk.set_config_value("batch_size", "256")
    trk.set_radial_zlut(luts)
    trk.set_localization_mode(w["mode"])
    calls = []
    for k in range(len(src)):
        job = api.make_jobs(1, frame=np.array([k // 16]), zlut_index=np.array([k]))[0]
        frame = np.ascontiguousarray(src[k])
        assert trk.schedule_frame(frame, [(0, 0)], job) == 1
        calls.append((frame, [(0, 0)], job))
    trk.wait_for(len(src))
    r = _by_index(trk.fetch_results(len(src)))
    trk.free()
    assert sorted(r["zlutIndex"]) == list(range(len(src)))
    ref = _oracle_frames(oracle, w, luts, calls)
    _report("one-ROI frames", r, ref)
    _assert_parity(r, ref)


def test_frame_copy_threads_config(gpu, oracle):
    """frame_copy_threads sets how many threads copy a frame into pinned memory: results do not
    depend on it (a 1.5 MB row block, copied by 1 and by 8 threads) and match the oracle's
    ScheduleFrame + ProcessJob, and it is range-checked."""
    w = synth.make_workload("realtime100", n=32)
    frame = np.zeros((1000, 1600), np.uint16)
    pos = [(50 * (i % 8) * 4, 120 * (i // 8)) for i in range(32)]
    for i, (x, y) in enumerate(pos):
        frame[y:y + 100, x:x + 100] = w["rois"][i]
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["luts"][:32])
    trk.set_localization_mode(w["mode"])
    assert trk.get_config_values()["frame_copy_threads"] == "5"
    job = api.make_jobs(1, frame=np.array([3]), zlut_index=np.array([0]))[0]
    out = []
    for th in ("1", "8"):
        trk.set_config_value("frame_copy_threads", th)
        assert trk.get_config_values()["frame_copy_threads"] == th
        assert trk.schedule_frame(frame, pos, job) == 32
        trk.wait_for(32)
        r = trk.fetch_results(32)
        out.append(r[np.argsort(r["zlutIndex"])])
    for bad in ("0", "65"):
        with pytest.raises(api.QTrkError, match="frame_copy_threads"):
            trk.set_config_value("frame_copy_threads", bad)
    trk.free()
    assert np.array_equal(out[0], out[1])
    _assert_parity(_by_index(out[0]), _oracle_frames(oracle, w, w["luts"][:32], [(frame, pos, job)]))


def test_build_lut_matches_oracle(gpu, oracle):
    w = synth.make_workload("plain64", n=8)
    cfg = w["cfg"]
    o_cfg = oracle.config_from_bytes(bytes(cfg))
    planes, count = 20, 3
    rng = np.random.default_rng(11)
    lut_src = w["render_luts"][0]
    for known, norm in ((False, True), (True, False)):
        trk = api.QTrkInstance(w["settings"])
        trk.set_radial_zlut((count, planes))
        trk.begin_lut(api.BUILDLUT_NORMALIZE if norm else 0)
        ref = np.zeros((count, planes, cfg.zlut_radialsteps), np.float32)
        for p in range(planes):
            posn = np.stack([32 + rng.uniform(-2, 2, count), 32 + rng.uniform(-2, 2, count),
                             np.full(count, 5 + 4.5 * p)], 1).astype(np.float32)
            imgs = api.generate_images_from_lut(lut_src[None], posn, np.zeros(count, np.int32), 64, 64,
                                                cfg.zlut_minradius, cfg.zlut_maxradius)
            imgs = (1000 + 800 * imgs).astype(np.float32)
            kp = posn[:, :2] if known else None
            trk.build_lut(imgs, p, kp)
            oracle.build_lut(o_cfg, imgs, p, ref, known_pos=kp, normalize=norm)
        pre = trk.get_radial_zlut()
        assert np.abs(pre - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())
        trk.finalize_lut()
        oracle.finalize_lut(ref)
        got = trk.get_radial_zlut()
        np.testing.assert_allclose(got, ref, atol=1e-4)
        trk.free()


def test_zeroed_lut_and_sizes(gpu):
    w = synth.make_workload("plain64", n=4)
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut((3, 17))
    assert trk.get_radial_zlut_size() == (3, 17, w["cfg"].zlut_radialsteps)
    assert np.all(trk.get_radial_zlut() == 0)
    trk.set_radial_zlut(w["luts"])
    assert np.array_equal(trk.get_radial_zlut(), w["luts"])
    trk.free()


def test_full_batch_properties(gpu, oracle):
    """The full 4,096-ROI batch: a seeded sample against the oracle, and size-independent checks for all."""
    w = synth.make_workload("realtime100")
    trk = _tracker(w)
    got, best, _ = _batch(trk, w["rois"], w["jobs"])
    # the same batch twice gives identical results (no cross-ROI interference, no races)
    again, best2, _ = _batch(trk, w["rois"], w["jobs"])
    trk.free()
    assert np.array_equal(got, again) and np.array_equal(best, best2)
    assert np.array_equal(got["frame"], w["jobs"]["frame"]) and np.array_equal(got["zlutIndex"], w["jobs"]["zlutIndex"])
    assert np.all(np.isfinite(got["x"])) and np.all(np.isfinite(got["z"]))
    assert np.all((best >= 0) & (best < 200))
    ex, ez = np.abs(got["x"] - w["truth"][:, 0]), np.abs(got["z"] - w["truth"][:, 2])
    # the oracle on the same batch: rms 0.056 px, 99.5th percentile 0.31 px (QI has rare outliers)
    assert np.sqrt(np.mean(ex ** 2)) < 0.08 and np.percentile(ex, 99.5) < 0.5
    assert np.sqrt(np.mean(ez ** 2)) < 1.0
    sel = np.random.default_rng(3).choice(len(got), 300, replace=False)
    ref, rbest, _ = _oracle(oracle, w, rois=w["rois"][sel], jobs=w["jobs"][sel])
    _assert_parity(got[sel], ref, best[sel], rbest)


def test_large_roi_default_settings(gpu, oracle):
    """150x150 ROIs with the reference defaults: QI 256 radial steps (512-point FFT)."""
    st = api.default_settings(width=150, height=150, qi_iterations=3)
    cfg = api.computed_config(st)
    assert cfg.qi_radialsteps == 256
    lut = synth.analytic_lut(40, cfg.zlut_radialsteps, cfg.zlut_minradius, cfg.zlut_maxradius, 9.0)
    n = 16
    rng = np.random.default_rng(4)
    pos = np.stack([75 + rng.uniform(-3, 3, n), 75 + rng.uniform(-3, 3, n), rng.uniform(5, 35, n)], 1)
    rois = (100 + 2000 * (1 + api.generate_images_from_lut(lut[None], pos, np.zeros(n, np.int32), 150, 150,
                                                           cfg.zlut_minradius, cfg.zlut_maxradius))).astype(np.float32)
    jobs = api.make_jobs(n)
    luts = synth.normalize_profiles(lut)[None]
    mode = synth.MODE_QI_Z
    ref, rbest, _ = oracle.localize_trace(oracle.config_from_bytes(bytes(cfg)), mode, luts, rois, jobs)
    trk = api.QTrkInstance(st)
    trk.set_radial_zlut(luts)
    trk.set_localization_mode(mode)
    got, best, _ = _batch(trk, rois, jobs)
    trk.free()
    _report("150x150", got, ref)
    _assert_parity(got, ref, best, rbest)


def test_unsupported_mode_raises(gpu):
    w = synth.make_workload("plain64", n=2)
    trk = api.QTrkInstance(w["settings"])
    for bits in (32, 1024, 1 << 20):  # bits LocalizeModeEnum does not define (cputrack/qtrk_c_api.h:10-24)
        with pytest.raises(api.QTrkError, match="not implemented"):
            trk.set_localization_mode(bits | api.LT_QI)
    trk.free()


def _gauss_compare(tag, got, ref, r64):
    """x/y of the GPU against the FP32 oracle: the same NaN set and every finite value within
    1e-3 px (measured: all bit-identical).  Both evaluate qmath.h's erf and exp, so the model
    terms round the same; the FP32-FP64 oracle spread (the reference's own sensitivity to
    rounding) is printed beside."""
    for k in "xy":
        nan = np.isnan(ref[k])
        assert np.array_equal(np.isnan(got[k]), nan), k
        d = np.abs(got[k][~nan] - ref[k][~nan])
        same = int((got[k][~nan] == ref[k][~nan]).sum())
        spread = np.nanmax(np.abs(ref[k] - r64[k]))
        print(f"[{tag}] {k}: {same}/{d.size} bit-identical, max |GPU-oracle| {d.max():.2e} px, {int(nan.sum())} NaN; "
              f"oracle FP32-FP64 spread {spread:.2e}", flush=True)
        assert d.max() <= XY_TOL, k


def test_gaussian2d_mle(gpu, oracle, oracle64):
    """LT_Gaussian2D (cputrack/cpu_tracker.cpp:514-589) on sigma=1.5 spots in 16x16 ROIs: COM
    bit-exact and x/y within 1e-3 px of the oracle on every ROI.  The reference's independent
    (diagonal) Newton steps do not converge on every ROI, so its result is sensitive to the
    rounding of its erf/exp (its FP32 and FP64 restatements differ by up to ~0.06 px here); the
    kernel and the oracle share qmath.h's erf/exp for that reason."""
    w = synth.make_workload("gauss16", n=2048)
    ref, _, _ = _oracle(oracle, w, luts=None)
    r64, _, _ = oracle64.localize_trace(oracle64.config_from_bytes(bytes(w["cfg"])), w["mode"], None,
                                        w["rois"], w["jobs"])
    trk = api.QTrkInstance(w["settings"])
    trk.set_localization_mode(api.LT_Gaussian2D)
    got, _, _ = _batch(trk, w["rois"], w["jobs"], diag=False)
    trk.free()
    _report("gauss16", got, ref)
    assert np.array_equal(got["frame"], ref["frame"])
    assert np.array_equal(got["comx"], ref["comx"]) and np.array_equal(got["comy"], ref["comy"])
    _gauss_compare("gauss16", got, ref, r64)
    e_gpu = np.hypot(got["x"] - w["truth"][:, 0], got["y"] - w["truth"][:, 1])
    e_ref = np.hypot(ref["x"] - w["truth"][:, 0], ref["y"] - w["truth"][:, 1])
    assert abs(np.sqrt(np.nanmean(e_gpu ** 2)) - np.sqrt(np.nanmean(e_ref ** 2))) < 1e-4


def _gauss_sizes_run(oracle, W, H, n=512):
    """Gaussian spots (sigma 1.5, centre +-2 px, I 2,000-20,000, bg 5-50, Poisson) in W x H ROIs,
    through the GPU twice: LT_Gaussian2D alone (up to 1,024 pixels: one wavefront per ROI) and
    with LT_LocalizeZ added and no LUT (the workgroup kernel; its x/y are the same computation)."""
    rng = np.random.default_rng(W * 100 + H)
    st = api.default_settings(width=W, height=H, gauss2D_sigma=1.5, gauss2D_iterations=10)
    cfg = api.computed_config(st)
    pos = np.stack([W / 2 + rng.uniform(-2, 2, n), H / 2 + rng.uniform(-2, 2, n)], 1)
    lam = synth.gaussian_spots(pos, 1.5, rng.uniform(2000, 20000, n), rng.uniform(5, 50, n), W, H)
    w = dict(settings=st, cfg=cfg, mode=api.LT_Gaussian2D, luts=None, rois=rng.poisson(lam).astype(np.float32),
             jobs=api.make_jobs(n))
    ref, _, _ = _oracle(oracle, w, luts=None)
    out = []
    for mode in (api.LT_Gaussian2D, api.LT_Gaussian2D | api.LT_LocalizeZ):
        trk = api.QTrkInstance(st)
        trk.set_localization_mode(mode)
        r, _, _ = _batch(trk, w["rois"], w["jobs"], diag=False)
        trk.free()
        out.append(r)
    return w, ref, out[0], out[1]


@pytest.mark.parametrize("size", [(20, 12), (12, 20)], ids=["20x12", "12x20"])
def test_gaussian2d_mle_sizes(gpu, oracle, oracle64, size):
    """LT_Gaussian2D on non-square ROIs: the one-wavefront kernel and the workgroup kernel, each
    against the oracle at the 16x16 case's bar (every finite x/y within 1e-3 px)."""
    W, H = size
    w, ref, wave, wg = _gauss_sizes_run(oracle, W, H)
    r64, _, _ = oracle64.localize_trace(oracle64.config_from_bytes(bytes(w["cfg"])), w["mode"], None, w["rois"],
                                        w["jobs"])
    for tag, got in (("wavefront", wave), ("workgroup", wg)):
        _report(f"gauss {W}x{H} {tag}", got, ref)
        for k in ("comx", "comy", "imageMean"):
            assert np.array_equal(got[k], ref[k]), k
        _gauss_compare(f"gauss {W}x{H} {tag}", got, ref, r64)
    assert np.all(wave["z"] == 0) and np.all(wave["error"] == 0)


@pytest.mark.parametrize("size", [(24, 24), (32, 32)], ids=["24x24", "32x32"])
def test_gaussian2d_wavefront_matches_workgroup(gpu, oracle, oracle64, size):
    """Larger ROIs up to the one-wavefront kernel's 1,024 pixels: its x/y against the workgroup
    kernel's (the same NaN set, within 1e-4 px), COM and image mean bit-exact, and both against
    the oracle at the 1e-3 px bar.  On these sizes the reference's undamped diagonal Newton has
    not converged after 10 iterations from its start (I0 = mean W H / 2): its own FP32 and FP64
    restatements differ by up to 0.07 px (32x32) and a few percent of ROIs diverge to NaN, so the
    agreement rests on the shared erf/exp and the same operation order."""
    W, H = size
    w, ref, wave, wg = _gauss_sizes_run(oracle, W, H)
    r64, _, _ = oracle64.localize_trace(oracle64.config_from_bytes(bytes(w["cfg"])), w["mode"], None, w["rois"],
                                        w["jobs"])
    _gauss_compare(f"gauss {W}x{H} wavefront", wave, ref, r64)
    _report(f"gauss {W}x{H} wavefront", wave, ref)
    for k in ("comx", "comy", "imageMean"):
        assert np.array_equal(wave[k], ref[k]) and np.array_equal(wave[k], wg[k]), k
    for k in "xy":
        assert np.array_equal(np.isnan(wave[k]), np.isnan(wg[k]))
        fin = ~np.isnan(wave[k])
        d = np.abs(wave[k][fin] - wg[k][fin])
        print(f"[gauss {W}x{H}] {k}: {int((~fin).sum())} NaN; max |wavefront - workgroup| {d.max():.2e}", flush=True)
        assert np.all(d <= 1e-4)
