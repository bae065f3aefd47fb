"""QTrkLocalizeDeviceNoAlloc: the allocation-free variant of the device batch's run.

Through the C ABI: its results on the caller's GPU buffers equal QTrkBatchRun's (the existing
entry point with the same match precision) bit for bit, on two calls in a row, the second with
fewer ROIs (the shared scratch keeps the first call's contents); the FP32 form (no scratch)
likewise; a call needing more scratch than the fixed array and a BF16 call before the tracker's
LUT was converted fail with an error before anything is launched.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth

pytestmark = pytest.mark.gpu


def _device_results(trk, w, rois_d, jobs_d, n, prec):
    import torch

    res_d = torch.zeros(n * api.RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream()
    api.localize_device_noalloc(trk, rois_d.data_ptr(), api.QTrkU16, jobs_d.data_ptr(), res_d.data_ptr(), n, prec,
                                s.cuda_stream)
    s.synchronize()
    return np.frombuffer(res_d.cpu().numpy().tobytes(), api.RESULT_DTYPE)


def _batch_results(b, n):
    b.run(n, 1)
    b.sync()
    return b.download(n)


@pytest.mark.parametrize("prec", [api.MATCH_BF16, api.MATCH_FP32, api.MATCH_FP32_GROUPED],
                         ids=["bf16", "fp32", "fp32_grouped"])
def test_noalloc_equals_batch_run(gpu, prec):
    import torch

    w = synth.make_workload("realtime100", n=1024)
    jobs = w["jobs"].copy()
    jobs["zlutIndex"][:100] = 7  # one bead with many ROIs
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["luts"])
    trk.set_localization_mode(w["mode"])
    b = api.DeviceBatch(trk, 1024, api.QTrkU16)
    b.set_match_precision(prec)
    rois_d = torch.from_numpy(w["rois"].view(np.int16).copy()).cuda()
    jobs_d = torch.from_numpy(np.frombuffer(jobs.tobytes(), np.uint8).copy()).cuda()
    try:
        for n in (1024, 300):  # two calls in a row, the second smaller
            print(f"noalloc precision {prec}: {n} ROIs", flush=True)
            b.upload(w["rois"][:n], jobs[:n])
            ref = _batch_results(b, n)
            got = _device_results(trk, w, rois_d, jobs_d, n, prec)
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), n
    finally:
        b.free()
        trk.free()


def test_noalloc_rejections(gpu):
    import torch

    w = synth.make_workload("realtime100", n=16)
    trk = api.QTrkInstance(w["settings"])
    trk.set_radial_zlut(w["luts"])
    trk.set_localization_mode(w["mode"])
    buf = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda")
    p = buf.data_ptr()
    try:
        # not yet converted for BF16: refused before any launch
        with pytest.raises(api.QTrkError, match="not converted"):
            api.localize_device_noalloc(trk, p, api.QTrkU16, p, p, 16, api.MATCH_BF16)
        b = api.DeviceBatch(trk, 16, api.QTrkU16)
        b.set_match_precision(api.MATCH_BF16)
        b.free()
        # more scratch than the fixed array holds: refused before any launch (the buffers are
        # nevertheless large enough for n_big ROIs, jobs and results)
        n_big = (api.NOALLOC_SCRATCH_MB << 20) // (4 * trk.cfg.zlut_radialsteps) + 1
        big = torch.zeros(n_big * (100 * 100 * 2 + 24 + 52), dtype=torch.uint8, device="cuda")
        q = big.data_ptr()
        with pytest.raises(api.QTrkError, match="bytes of scratch"):
            api.localize_device_noalloc(trk, q, api.QTrkU16, q + n_big * 20000, q + n_big * 20024, n_big,
                                        api.MATCH_BF16)
        del big
        # a LUT change makes the converted tables stale: refused until converted again
        trk.set_radial_weights(np.ones(trk.cfg.zlut_radialsteps, np.float32))
        with pytest.raises(api.QTrkError, match="not converted"):
            api.localize_device_noalloc(trk, p, api.QTrkU16, p, p, 16, api.MATCH_BF16)
    finally:
        trk.free()
