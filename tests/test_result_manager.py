"""The result sink (RM* of cputrack/qtrk_c_api.h:233-242, cputrack/ResultManager.cpp).

CPU tests cover what needs no tracker: the binary v3 header (the layout the
reference's trace reader takes, TraceViewer/ViewTraceDlg.cs:195-240), the
frame-name split, configuration and counters.  The GPU tests attach a real
tracker and check the per-frame tables, the text and binary traces, the
frame-info file and memory trimming against the tracker's own results.
"""
import os
import struct
import time

import numpy as np
import pytest

from qtrk_amd import api, synth


def frame_dtype(num_beads, ncols):
    """One binary record: uint32 frame, float64 timestamp, float32 info[ncols], float32 pos[beads][3],
    int32 error[beads], float32 imageMean[beads] (packed)."""
    return np.dtype([("frame", "<u4"), ("timestamp", "<f8"), ("info", "<f4", (ncols,)),
                     ("pos", "<f4", (num_beads, 3)), ("error", "<i4", (num_beads,)),
                     ("mean", "<f4", (num_beads,))])


def read_binary(path):
    raw = open(path, "rb").read()
    version, nb, nc, off = struct.unpack_from("<4i", raw, 0)
    names = raw[16:off].split(b"\0")[:nc]
    recs = np.frombuffer(raw[off:], frame_dtype(nb, nc))
    return version, nb, nc, [n.decode() for n in names], recs


def test_binary_header_and_name_split(tmp_path):
    f = str(tmp_path / "trace.bin")
    with open(f, "w") as h:
        h.write("stale")  # replaced on creation
    rm = api.ResultManager(f, "", 7, info_columns=["motor", "piezo", "temp"], binary=True, write_interval=5)
    try:
        version, nb, nc, names, recs = read_binary(f)
        assert (version, nb, nc) == (3, 7, 3)
        # the reference's split appends the ';' to the following name (cputrack/qtrk_c_api.cpp:183-193)
        assert names == ["motor", ";piezo", ";temp"]
        assert len(recs) == 0 and os.path.getsize(f) == 16 + len("motor;piezo;temp") + 3
        cfg = rm.config()
        assert (cfg.numBeads, cfg.numFrameInfoColumns, cfg.writeInterval, cfg.binaryOutput) == (7, 3, 5, 1)
    finally:
        rm.free()


def test_frame_info_counters_without_tracker(tmp_path):
    out, info = str(tmp_path / "r.txt"), str(tmp_path / "info.txt")
    rm = api.ResultManager(out, info, 4, info_columns=["a", "b"], scaling=(2, 2, 1), write_interval=1)
    try:
        for fr in range(12):
            rm.store_frame_info(fr, fr * 0.5, [fr, -fr])
        rm.store_frame_info(3, 99.0, [0, 0])        # a frame's info is kept from its first store
        rm.store_frame_info(30000, 1.0, [0, 0])      # beyond processedFrames + 20000: ignored
        c = rm.counters()
        assert c["capturedFrames"] == 12 and c["processedFrames"] == 0 and c["localizationsDone"] == 0
        rm.flush()                                   # no tracker: nothing processed, files created empty
        assert open(out).read() == "" and open(info).read() == ""
        n, _ = rm.results(0, 1)
        assert n == 0                                # frame 0 is not processed yet
        assert len(rm.bead_results(0, 5, 1)) == 0
    finally:
        rm.free()


def test_create_errors(tmp_path):
    with pytest.raises(api.QTrkError, match="Unable to open"):
        api.ResultManager(str(tmp_path / "missing" / "x.bin"), "", 2, binary=True)


# ------------------------------------------------------------------ GPU

def _workload(beads=8, frames=48, seed=3):
    """plain64's bead rendered for `beads` x `frames` ROIs, frame-major, one LUT copy per bead."""
    w = synth.make_workload("plain64", n=beads * frames, seed=seed)
    w["luts"] = np.repeat(w["luts"], beads, axis=0)
    k = np.arange(beads * frames)
    w["jobs"]["frame"] = k // beads
    w["jobs"]["zlutIndex"] = k % beads
    return w


def _expected(w, trk, scaling, offset):
    ref = trk.localize_all(w["rois"], w["jobs"])
    W, H = np.float32(w["cfg"].width), np.float32(w["cfg"].height)
    sc, off = np.float32(scaling), np.float32(offset)
    pos = np.stack([((ref["x"] - W * np.float32(0.5)) + off[0]) * sc[0],
                    ((ref["y"] - H * np.float32(0.5)) + off[1]) * sc[1],
                    ((ref["z"] - np.float32(0)) + off[2]) * sc[2]], 1).astype(np.float32)
    return ref, pos


def _run(rm, trk, w, frames, beads, info=True):
    rm.set_tracker(trk)
    for fr in range(frames):
        if info:
            rm.store_frame_info(fr, 0.01 * fr, [fr, 2.0 * fr])
        for b in range(beads):
            i = fr * beads + b
            trk.schedule_localization(w["rois"][i], w["jobs"][i])
    t0 = time.time()
    rm.flush()
    while rm.counters()["processedFrames"] < frames and time.time() - t0 < 30:
        time.sleep(0.01)
        rm.flush()
    rm.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("binary", [False, True], ids=["text", "binary"])
def test_result_manager_traces(gpu, tmp_path, binary):
    beads, frames = 8, 48
    w = _workload(beads, frames)
    scaling, offset = (0.5, 2.0, 3.0), (1.0, -2.0, 0.25)
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["luts"])
    trk.set_localization_mode(w["mode"])
    ref, pos = _expected(w, trk, scaling, offset)
    out, info = str(tmp_path / ("t.bin" if binary else "t.txt")), str(tmp_path / "info.txt")
    rm = api.ResultManager(out, info, beads, info_columns=["motor", "piezo"], scaling=scaling, offset=offset,
                           write_interval=10, binary=binary)
    try:
        _run(rm, trk, w, frames, beads)
        c = rm.counters()
        print("counters", c, flush=True)
        assert c["processedFrames"] == frames and c["localizationsDone"] == frames * beads
        assert c["lostFrames"] == 0 and c["fileError"] == 0 and c["lastSaveFrame"] == frames
        # in-memory tables
        n, tab = rm.results(0, frames)
        assert n == frames
        got = np.stack([tab["x"].ravel(), tab["y"].ravel(), tab["z"].ravel()], 1)
        assert np.array_equal(got, pos)
        assert np.array_equal(tab["error"].ravel(), ref["error"])
        assert np.array_equal(tab["imageMean"].ravel(), ref["imageMean"])
        br = rm.bead_results(5, 10, 3)
        assert len(br) == 10 and np.array_equal(br["x"], pos[np.arange(5, 15) * beads + 3, 0])
        # the trace file
        if binary:
            version, nb, nc, names, recs = read_binary(out)
            assert (version, nb, nc, names) == (3, beads, 2, ["motor", ";piezo"])
            assert np.array_equal(recs["frame"], np.arange(frames))
            assert np.array_equal(recs["timestamp"], 0.01 * np.arange(frames))
            assert np.array_equal(recs["info"], np.stack([np.arange(frames), 2.0 * np.arange(frames)], 1))
            assert np.array_equal(recs["pos"].reshape(-1, 3), pos)
            assert np.array_equal(recs["error"].ravel(), ref["error"].astype(np.int32))
            assert np.array_equal(recs["mean"].ravel(), ref["imageMean"])
        else:
            lines = open(out).read().splitlines()
            assert len(lines) == frames
            for fr, line in enumerate(lines):
                exp = "%d\t%f\t" % (fr, 0.01 * fr) + "".join(
                    "%.7f\t%.7f\t%.7f\t" % tuple(float(v) for v in pos[fr * beads + b]) for b in range(beads))
                assert line == exp
            info_lines = open(info).read().splitlines()
            assert info_lines[7] == "7\t0.070000\t7.0000000\t14.0000000\t"
    finally:
        rm.free()
        trk.free()


@pytest.mark.gpu
def test_result_manager_trims_memory_and_removes_beads(gpu, tmp_path):
    """maxFramesInMemory drops written frames from the front (cputrack/ResultManager.cpp:254-275);
    RemoveBead drops a bead's slot from the frames held.  Without frame info the frames are allocated
    by the results, which arrive in one batch: all are processed and written before the trim (frame
    info stored ahead of the results would let the trim drop frames that never completed, as in the
    reference)."""
    beads, frames = 4, 40
    w = _workload(beads, frames, seed=5)
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["luts"])
    trk.set_localization_mode(w["mode"])
    ref, pos = _expected(w, trk, (1, 1, 1), (0, 0, 0))
    out = str(tmp_path / "t.bin")
    rm = api.ResultManager(out, "", beads, write_interval=2, max_frames_in_memory=8, binary=True)
    try:
        _run(rm, trk, w, frames, beads, info=False)
        c = rm.counters()
        print("counters", c, flush=True)
        assert c["processedFrames"] == frames and c["startFrame"] > 0 and c["lostFrames"] == 0
        assert rm.results(0, 4)[0] == 0                     # dropped from memory
        s = c["startFrame"]
        n, tab = rm.results(s, frames - s)
        assert n == frames - s
        assert np.array_equal(tab["x"].ravel(), pos[s * beads:, 0])
        _, _, _, _, recs = read_binary(out)                 # every frame reached the file before it was dropped
        assert np.array_equal(recs["frame"], np.arange(frames))
        assert np.array_equal(recs["pos"].reshape(-1, 3), pos)
        rm.remove_bead(1)
        n, tab = rm.results(s, frames - s)
        assert n == frames - s
        assert np.array_equal(tab["x"][:, 1], pos[s * beads:, 0].reshape(-1, beads)[:, 2])
    finally:
        rm.free()
        trk.free()


def test_null_handles_raise():
    L = api.lib()
    L.QTrkClearLastError()
    L.RMFlush(None)
    with pytest.raises(api.QTrkError, match="null result manager"):
        api.check()
    fc = api.RMFrameCounters()
    L.RMGetFrameCounters(None, api.C.byref(fc))
    with pytest.raises(api.QTrkError, match="null result manager"):
        api.check()
