"""Writes the committed fixtures of tests/golden/ (run once, where the reference is mounted).

ref10x_bead0.npy   bead 0 of the reference's measured radial LUT fixture
                   cputrack-test/10x.radialzlut (header int32 10,150,116; float32
                   [10][150][116]), read as data: float32 [150 planes][116 steps].
golden_plain64.npz 24 seeded plain64 ROIs (Poisson counts, stored exactly as uint16)
                   with the oracle's outputs for them: x, y, z, COM, image mean,
                   boundary flag and best LUT plane.  The reference itself cannot be
                   built here (its sources need Win32 headers), so these are the
                   oracle's values, frozen: a regression pin for the oracle and a
                   fixed target for the GPU path.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF_LUT = "/root/reference/cputrack-test/10x.radialzlut"


def main():
    hdr = np.fromfile(REF_LUT, dtype="<i4", count=3)
    assert list(hdr) == [10, 150, 116], hdr
    lut = np.fromfile(REF_LUT, dtype="<f4", offset=12).reshape(10, 150, 116)
    np.save(os.path.join(HERE, "ref10x_bead0.npy"), lut[0].copy())

    from oracle.bind import Oracle
    from qtrk_amd import synth

    w = synth.make_workload("plain64", n=24)
    rois = w["rois"]
    assert np.all(rois == np.round(rois)) and rois.max() < 65536
    o = Oracle()
    res, best, _ = o.localize_trace(o.config_from_bytes(bytes(w["cfg"])), w["mode"], w["luts"], rois, w["jobs"])
    np.savez_compressed(os.path.join(HERE, "golden_plain64.npz"), rois_u16=rois.astype(np.uint16),
                        jobs=w["jobs"], luts=w["luts"], cfg=np.frombuffer(bytes(w["cfg"]), np.uint8),
                        mode=np.int32(w["mode"]), x=res["x"], y=res["y"], z=res["z"], comx=res["comx"],
                        comy=res["comy"], imageMean=res["imageMean"], error=res["error"], best=best)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
