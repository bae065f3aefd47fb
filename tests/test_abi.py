"""The drop-in boundary: packed layouts, exported symbols, and the host-side
parts of the C ABI that need no GPU (config derivation, utility exports)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from qtrk_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
REF = "/root/reference"

# Byte offsets of the packed structs (cputrack/qtrk_c_api.h:32-140)
LAYOUT = {
    "LocalizationJob": (24, {"frame": 0, "timestamp": 4, "zlutIndex": 8, "initialPos": 12}),
    "LocalizationResult": (52, {"job": 0, "pos": 24, "firstGuess": 36, "error": 44, "imageMean": 48}),
    "QTrkSettings": (84, {"width": 0, "height": 4, "numThreads": 8, "cuda_device": 12, "com_bgcorrection": 16,
                          "zlut_minradius": 20, "qi_iterations": 36, "qi_angstep_factor": 56,
                          "xc1_profileLength": 60, "gauss2D_sigma": 76, "downsample": 80}),
    "QTrkComputedConfig": (108, {"zlut_radialsteps": 84, "zlut_angularsteps": 88, "zlut_maxradius": 92,
                                 "qi_radialsteps": 96, "qi_angstepspq": 100, "qi_maxradius": 104}),
    "ROIPosition": (8, {"x": 0, "y": 4}),
    "ResultManagerConfig": (41, {"numBeads": 0, "numFrameInfoColumns": 4, "scaling": 8, "offset": 20,
                                 "writeInterval": 32, "maxFramesInMemory": 36, "binaryOutput": 40}),
    "RMFrameCounters": (28, {"startFrame": 0, "processedFrames": 4, "lastSaveFrame": 8, "capturedFrames": 12,
                             "localizationsDone": 16, "lostFrames": 20, "fileError": 24}),
}


def test_ctypes_layouts():
    for name, (size, offs) in LAYOUT.items():
        cls = getattr(api, name)
        assert C.sizeof(cls) == size, name
        for f, o in offs.items():
            assert getattr(cls, f).offset == o, (name, f)
    assert api.RESULT_DTYPE.itemsize == 52 and api.JOB_DTYPE.itemsize == 24


def test_c_header_layouts(tmp_path):
    """Compiles include/qtrk_c_api.h as C and checks sizeof/offsetof from C itself."""
    checks = []
    for name, (size, offs) in LAYOUT.items():
        checks.append(f'printf("%d\\n", (int)sizeof({name}));')
        for f in offs:
            checks.append(f'printf("%d\\n", (int)offsetof({name}, {f}));')
    src = tmp_path / "l.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "qtrk_c_api.h"\n'
                   "int main(void){" + "".join(checks) + "return 0;}\n")
    exe = tmp_path / "l"
    subprocess.check_call(["gcc", "-std=c99", "-I" + INCLUDE, str(src), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    exp = []
    for name, (size, offs) in LAYOUT.items():
        exp.append(size)
        exp.extend(offs.values())
    assert got == exp
    # and the header is valid C++ as well
    subprocess.check_call(["g++", "-std=c++11", "-fsyntax-only", "-I" + INCLUDE, "-x", "c++",
                           os.path.join(INCLUDE, "QueuedTracker.h")])


def _declared_symbols():
    names = set()
    for h in os.listdir(INCLUDE):
        if h.endswith(".h"):
            txt = "\n".join(l for l in open(os.path.join(INCLUDE, h)) if not l.lstrip().startswith("#"))
            names |= set(re.findall(r"QTRK_API\s+[\w\s\*]+?\b(\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = api.lib()
    declared = _declared_symbols()
    assert len(declared) > 50
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert declared <= set(api.SIGNATURES), declared - set(api.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(REF + "/QTrkDotNet/QTrkDLL.cs"), reason="reference not mounted")
def test_csharp_binding_resolves():
    """Every native function the C# binding imports from qtrk.dll exists here (QTrkDotNet/QTrkDLL.cs)."""
    cs = open(REF + "/QTrkDotNet/QTrkDLL.cs").read()
    imported = re.findall(r'DllImport\(DllName[^\]]*\]\s*(?:\[[^\]]*\]\s*)*public static extern [\w\*\[\]]+ (\w+)\(', cs)
    assert len(imported) > 40
    L = api.lib()
    missing = [n for n in imported if not hasattr(L, n)]
    assert missing == []


def test_default_and_computed_config_match_oracle(oracle):
    L = api.lib()
    s = api.QTrkSettings()
    L.QTrkGetDefaultConfig(C.byref(s))
    assert (s.width, s.height, s.numThreads, s.cuda_device, s.qi_iterations) == (150, 150, -1, -2, 4)
    assert (s.zlut_radial_coverage, s.qi_angular_coverage, s.gauss2D_sigma) == (3.0, np.float32(0.7), 4.0)
    rng = np.random.default_rng(0)
    for _ in range(200):
        kw = dict(width=int(rng.integers(8, 200)), height=int(rng.integers(8, 200)),
                  zlut_minradius=float(rng.uniform(0, 4)), zlut_roi_coverage=float(rng.uniform(0.5, 1)),
                  zlut_radial_coverage=float(rng.uniform(0.5, 3)), zlut_angular_coverage=float(rng.uniform(0.3, 2)),
                  qi_minradius=float(rng.uniform(0, 4)), qi_roi_coverage=float(rng.uniform(0.5, 1)),
                  qi_radial_coverage=float(rng.uniform(0.5, 3)), qi_angular_coverage=float(rng.uniform(0.3, 2)))
        cc = api.computed_config(api.default_settings(**kw))
        oc = oracle.config(**kw)
        assert bytes(cc) == bytes(oc)


def _imagedata(a):
    return api.ImageData(a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[1], a.shape[0])


def test_utility_exports_match_oracle(oracle):
    L = api.lib()
    from qtrk_amd import synth
    lut = synth.analytic_lut(60, 40, 1.0, 30.0, 8.5)
    for (x, y, z, spline) in [(31.3, 30.8, 17.6, True), (29.0, 33.5, 40.2, False), (30.0, 30.0, 1.2, True)]:
        img = np.zeros((62, 62), np.float32)
        L.GenerateImageFromLUT(C.byref(_imagedata(img)), C.byref(_imagedata(lut)), 1.0, 30.0,
                               api.Vector3f(x, y, z), spline, 4)
        assert np.array_equal(img, oracle.generate_image_from_lut(62, 62, lut, 1.0, 30.0, x, y, z, spline, 4))
        for norm in (False, True):
            prof = np.zeros(40, np.float32)
            L.ComputeRadialProfile(prof.ctypes.data_as(C.POINTER(C.c_float)), 40, 200, 1.0, 30.0,
                                   api.Vector2f(x, y), C.byref(_imagedata(img)), 3.5, norm)
            assert np.array_equal(prof, oracle.radial_profile(img, 40, 200, 1.0, 30.0, x, y, 3.5, norm))
    p = np.random.default_rng(2).standard_normal(64).astype(np.float32)
    q = p.copy()
    L.NormalizeRadialProfile(q.ctypes.data_as(C.POINTER(C.c_float)), 64)
    assert np.array_equal(q, oracle.normalize_profile(p))
    batch = api.generate_images_from_lut(lut[None], np.array([[31.3, 30.8, 17.6]]), [0], 62, 62, 1.0, 30.0)
    assert np.array_equal(batch[0], oracle.generate_image_from_lut(62, 62, lut, 1.0, 30.0, 31.3, 30.8, 17.6))


def test_create_instance_fails_loudly_without_device():
    L = api.lib()
    n = C.c_int(0)
    try:
        hip = C.CDLL("libamdhip64.so")
        ok = hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        ok = False
    if ok:
        pytest.skip("a GPU is present")
    with pytest.raises(api.QTrkError):
        api.QTrkInstance(api.default_settings(width=64, height=64))
