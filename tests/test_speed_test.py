"""qtrk_amd/speed_test.c, the C client that times the reference's speed-test pattern through the
C ABI (cudatrack-test/test.cu:310-371 SpeedTest: every ROI scheduled by its own call, then Flush
and a wait for N results), as bench.py runs it: each call form's results equal the device-resident
batch on the same ROIs."""
import os

import numpy as np
import pytest

from qtrk_amd import api, synth
from qtrk_amd.build import SPEED_BIN

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["frame1", "roi"])
def test_speed_test_results_equal_device_batch(gpu, mode):
    import bench

    if not os.path.exists(SPEED_BIN):
        pytest.fail(f"{SPEED_BIN} is not built")
    w = synth.make_workload("realtime100", n=768)
    st, res = bench.c_abi_speed_test(w, mode, 2)
    assert st["passes"] == 2 and st["locs_per_s_median"] > 0
    trk = api.QTrkInstance(w["settings"])
    try:
        trk.set_radial_zlut(w["luts"])
        trk.set_localization_mode(w["mode"])
        b = api.DeviceBatch(trk, len(w["jobs"]), api.QTrkU16)
        b.upload(w["rois"], w["jobs"])
        b.run(len(w["jobs"]), 1)
        b.sync()
        ref = b.download(len(w["jobs"]))
        b.free()
    finally:
        trk.free()
    ref = ref[np.lexsort((ref["zlutIndex"], ref["frame"]))]
    assert len(res) == len(ref)
    for k in ("frame", "zlutIndex", "x", "y", "z", "comx", "comy", "error", "imageMean"):
        assert np.array_equal(res[k], ref[k]), k
