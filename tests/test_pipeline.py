"""Multi-tracker pipeline (QTrkPipeline*): one graph launch runs several trackers' chains.

Each stage of the pipeline must equal its standalone counterpart on the same
ROIs: COM/QI positions equal the tracker's own device-resident batch bit for
bit; the grouped BF16 match gives the tracker's exact per-ROI z (the
reference's LUTProfileCompare + FP64 fit, itself pinned to the oracle); the
MLE stage equals QTrkLUTModel's fit from the same start.  Autotuned launch
shapes must not change the localization results.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth

SPECS = [(32, 96, 8, 60, 32), (64, 64, 4, 80, 64)]  # size, ROIs, beads, planes, steps


def make_trackers(seed):
    rng = np.random.default_rng(seed)
    out = []
    for size, n, beads, planes, steps in SPECS:
        w = synth.batching_tracker(size, n, beads, planes, steps, rng)
        trk = api.QTrkInstance(w["settings"])
        trk.set_radial_zlut(w["lut"])
        trk.set_localization_mode(synth.MODE_QI_Z)
        out.append((w, trk))
    return out


def reference_results(w, trk):
    b = api.DeviceBatch(trk, len(w["jobs"]), api.QTrkU16)
    try:
        b.upload(w["rois"], w["jobs"])
        b.run(len(w["jobs"]))
        b.sync()
        return b.download(len(w["jobs"]))
    finally:
        b.free()


@pytest.mark.gpu
@pytest.mark.parametrize("autotune", [False, True], ids=["defaults", "autotuned"])
def test_pipeline_stages_equal_standalone(gpu, autotune):
    import torch
    tr = make_trackers(21)
    p = api.Pipeline()
    try:
        slots = []
        for w, trk in tr:
            s = p.add(trk, len(w["jobs"]), api.QTrkU16, api.MATCH_BF16, w["model_lut"], 0.0, w["cfg"].zlut_maxradius, 5)
            p.set_jobs(s, w["jobs"])
            p.upload(s, w["rois"])
            slots.append(s)
        p.finalize(autotune)
        print(p.report(), flush=True)
        p.launch()
        print("graph ms", p.sync())
        for (w, trk), s in zip(tr, slots):
            res, th, cr = p.download(s)
            ref = reference_results(w, trk)
            for k in ("frame", "timestamp", "zlutIndex", "x", "y", "comx", "comy", "error", "imageMean"):
                assert np.array_equal(res[k], ref[k]), k
            assert np.array_equal(res["z"], ref["z"]), np.max(np.abs(res["z"] - ref["z"]))
            # the MLE stage from the same start as a standalone fit
            t0 = np.stack([res["x"], res["y"], np.clip(res["z"], 0, w["model_lut"].shape[0] - 1),
                           -np.ones(len(res)), -np.ones(len(res))], 1).astype(np.float32)
            m = api.LUTModel(w["model_lut"], 0.0, w["cfg"].zlut_maxradius, w["cfg"].width, w["cfg"].height)
            try:
                th_ref, _, cr_ref, _ = m.fit(torch.from_numpy(w["rois"].astype(np.float32)), torch.from_numpy(t0), 5)
                th_ref, cr_ref = th_ref.cpu().numpy(), cr_ref.cpu().numpy()
            finally:
                m.free()
            if not autotune:  # the same lanes per ROI: the same reduction order
                assert np.array_equal(th, th_ref)
                assert np.array_equal(cr, cr_ref)
            else:
                assert np.allclose(th, th_ref, rtol=1e-4, atol=1e-3 * cr_ref)
            e = (th - w["truth"]) / cr
            print(f"{w['cfg'].width}px: |error|/CRLB median {np.round(np.median(np.abs(e), 0), 3)}")
    finally:
        p.free()
        for _, trk in tr:
            trk.free()


@pytest.mark.gpu
def test_pipeline_stream_and_repeat(gpu):
    tr = make_trackers(22)
    p = api.Pipeline()
    try:
        for w, trk in tr:
            s = p.add(trk, len(w["jobs"]), api.QTrkU16, api.MATCH_FP16, w["model_lut"], 0.0, w["cfg"].zlut_maxradius, 3)
            p.set_jobs(s, w["jobs"])
            p.upload(s, w["rois"])
        p.finalize(False)
        p.launch()
        p.sync()
        first = [p.download(s) for s in range(len(tr))]
        lat, gpu_ms = p.stream(12, 2.0)
        assert np.all(np.isfinite(lat)) and np.all(lat > 0) and np.all(gpu_ms > 0)
        print("latency ms", np.round(lat, 3), "gpu ms", np.round(gpu_ms, 3))
        # every replay of the graph gives the same results
        for s in range(len(tr)):
            again = p.download(s)
            assert np.array_equal(again[0], first[s][0]) and np.array_equal(again[1], first[s][1])
    finally:
        p.free()
        for _, trk in tr:
            trk.free()


@pytest.mark.gpu
def test_pipeline_errors(gpu):
    w = synth.batching_tracker(32, 16, 4, 40, 32, np.random.default_rng(3), render=False)
    bare = api.QTrkInstance(w["settings"])
    p = api.Pipeline()
    try:
        with pytest.raises(api.QTrkError, match="no radial Z-LUT"):
            p.add(bare, 16, api.QTrkU16)
        bare.set_radial_zlut(w["lut"])
        s = p.add(bare, 16, api.QTrkU16)
        with pytest.raises(api.QTrkError, match="SetJobs"):
            p.finalize()
        with pytest.raises(api.QTrkError, match="exactly"):
            p.set_jobs(s, w["jobs"][:5])
        with pytest.raises(api.QTrkError, match="Finalize"):
            p.launch()
    finally:
        p.free()
        bare.free()


@pytest.mark.gpu
def test_pipeline_batching_workload_against_oracle(gpu, oracle):
    """The batching workload at its stated shape: 4 trackers (ROI 32/64/96/128 px with 2,048/4,096/
    1,024/512 uint16 ROIs, LUTs of 100/150/200/300 planes x 32/64/96/96 steps) in one autotuned graph
    (COM -> QI(5) -> BF16 grouped match -> 5-iteration MLE).  Per tracker, 160 ROIs spread over the
    batch against the oracle's ProcessJob (cputrack/QueuedCPUTracker.cpp:273-361): indexing and
    boundary flags bit-exact, x/y within 1e-3 px, z within 1e-3 planes.  All 7,680 records through
    properties: each job's record in job order, finite positions inside the ROI, finite z,
    finite positive CRLBs for the MLE, and a second launch giving identical output."""
    rng = np.random.default_rng(0x51AB0004)
    p = api.Pipeline()
    tr = []
    try:
        for size, n, beads, planes, steps in synth.BATCHING:
            w = synth.batching_tracker(size, n, beads, planes, steps, rng)
            trk = api.QTrkInstance(w["settings"])
            trk.set_radial_zlut(w["lut"])
            trk.set_localization_mode(synth.MODE_QI_Z)
            tr.append((w, trk))
            s = p.add(trk, n, api.QTrkU16, api.MATCH_BF16, w["model_lut"], 0.0, w["cfg"].zlut_maxradius, 5)
            p.set_jobs(s, w["jobs"])
            p.upload(s, w["rois"])
        p.finalize(True)
        print(p.report(), flush=True)
        p.launch()
        print("graph ms", p.sync(), flush=True)
        outs = [p.download(s) for s in range(len(tr))]
        p.launch()
        p.sync()
        for s, (res, th, cr) in enumerate(outs):
            again = p.download(s)
            assert np.array_equal(again[0], res) and np.array_equal(again[1], th) and np.array_equal(again[2], cr)
        for (w, _), (res, th, cr) in zip(tr, outs):
            size, planes = w["cfg"].width, w["model_lut"].shape[0]
            jobs = w["jobs"]
            assert np.array_equal(res["frame"], jobs["frame"]) and np.array_equal(res["zlutIndex"], jobs["zlutIndex"])
            for k in ("x", "y", "comx", "comy"):
                assert np.all(np.isfinite(res[k])) and np.all((res[k] >= 0) & (res[k] < size)), k
            # the 7-point fit's vertex can leave the plane range where the scores are flat (as the
            # reference's does): finite, with the count outside the range printed
            assert np.all(np.isfinite(res["z"]))
            print(f"{size}px: z in [{res['z'].min():.2f}, {res['z'].max():.2f}], "
                  f"{int(((res['z'] < 0) | (res['z'] > planes - 1)).sum())} outside [0, {planes - 1}]", flush=True)
            ok = np.all(np.isfinite(th), 1) & np.all(np.isfinite(cr) & (cr > 0), 1)
            print(f"{size}px: {int((~ok).sum())} of {len(ok)} MLE fits without a finite CRLB", flush=True)
            assert ok.mean() >= 0.99
            idx = np.linspace(0, len(jobs) - 1, 160).astype(int)
            ref, _, _ = oracle.localize_trace(oracle.config_from_bytes(bytes(w["cfg"])), synth.MODE_QI_Z, w["lut"],
                                              w["rois"][idx], jobs[idx])
            got = res[idx]
            for k in ("frame", "zlutIndex", "error"):
                assert np.array_equal(got[k], ref[k]), k
            d = {k: float(np.abs(got[k].astype(np.float64) - ref[k]).max()) for k in ("x", "y", "z", "comx", "comy")}
            same = {k: int((got[k] == ref[k]).sum()) for k in ("x", "y", "z")}
            print(f"{size}px vs oracle (160 ROIs): max |d| {d}, bit-identical {same}", flush=True)
            assert max(d["x"], d["y"], d["comx"], d["comy"]) <= 1e-3 and d["z"] <= 1e-3
    finally:
        p.free()
        for _, trk in tr:
            trk.free()


@pytest.mark.gpu
def test_pipeline_lut_change_needs_finalize(gpu, oracle):
    """A tracker's LUT changed after finalize (here its radial weights) makes the launch fail until
    the pipeline is finalized again; then the grouped match scores against the new weights (its z
    against the oracle's ProcessJob with those weights)."""
    w = synth.batching_tracker(64, 128, 8, 80, 64, np.random.default_rng(5))
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["lut"])
    trk.set_localization_mode(synth.MODE_QI_Z)
    p = api.Pipeline()
    try:
        s = p.add(trk, 128, api.QTrkU16, api.MATCH_BF16)
        p.set_jobs(s, w["jobs"])
        p.upload(s, w["rois"])
        p.finalize(False)
        p.launch()
        p.sync()
        wts = np.linspace(0.2, 2.0, trk.cfg.zlut_radialsteps).astype(np.float32)
        trk.set_radial_weights(wts)
        with pytest.raises(api.QTrkError, match="finalize again"):
            p.launch()
        p.finalize(False)
        p.launch()
        p.sync()
        res = p.download(s)[0]
        ref, _, _ = oracle.localize_trace(oracle.config_from_bytes(bytes(w["cfg"])), synth.MODE_QI_Z, w["lut"],
                                          w["rois"], w["jobs"], zcmp=wts)
        assert np.array_equal(res["x"], ref["x"])
        assert float(np.abs(res["z"] - ref["z"]).max()) <= 1e-3
    finally:
        p.free()
        trk.free()
