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
