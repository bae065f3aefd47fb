"""ctypes binding of libqtrk.so, mirroring the reference's C# binding.

Struct layouts follow QTrkDotNet/QTrkDLL.cs:38-215 (packed, sequential) and
the function table follows QTrkDotNet/QTrkDLL.cs:240-424; ``QTrkInstance``
mirrors QTrkDotNet/QTrkInstance.cs:31-189.  Every call checks the library's
last-error slot and raises ``QTrkError`` when a call failed, so a failing GPU
path is never silent.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqtrk.so")

# LocalizeModeEnum (cputrack/qtrk_c_api.h:5-20)
LT_OnlyCOM = 0
LT_XCor1D = 1
LT_QI = 2
LT_Gaussian2D = 4
LT_ZLUTAlign = 8
LT_LocalizeZ = 16
LT_NormalizeProfile = 64
LT_ClearFirstFourPixels = 128
LT_FourierLUT = 256
LT_LocalizeZWeighted = 512

# QTRK_PixelDataType (cputrack/qtrk_c_api.h:24-29)
QTrkU8, QTrkU16, QTrkFloat = 0, 1, 2
PDT_OF_DTYPE = {np.dtype(np.uint8): QTrkU8, np.dtype(np.uint16): QTrkU16, np.dtype(np.float32): QTrkFloat}

BUILDLUT_IMAGELUT, BUILDLUT_FOURIER, BUILDLUT_NORMALIZE, BUILDLUT_BIASCORRECT = 1, 2, 4, 8


class Vector2f(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Vector3f(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class LocalizationJob(C.Structure):
    _pack_ = 1
    _fields_ = [("frame", C.c_uint), ("timestamp", C.c_uint), ("zlutIndex", C.c_int), ("initialPos", Vector3f)]


class LocalizationResult(C.Structure):
    _pack_ = 1
    _fields_ = [("job", LocalizationJob), ("pos", Vector3f), ("firstGuess", Vector2f),
                ("error", C.c_uint), ("imageMean", C.c_float)]


_SETTINGS_FIELDS = [
    ("width", C.c_int), ("height", C.c_int), ("numThreads", C.c_int), ("cuda_device", C.c_int),
    ("com_bgcorrection", C.c_float),
    ("zlut_minradius", C.c_float), ("zlut_radial_coverage", C.c_float),
    ("zlut_angular_coverage", C.c_float), ("zlut_roi_coverage", C.c_float),
    ("qi_iterations", C.c_int), ("qi_minradius", C.c_float), ("qi_radial_coverage", C.c_float),
    ("qi_angular_coverage", C.c_float), ("qi_roi_coverage", C.c_float), ("qi_angstep_factor", C.c_float),
    ("xc1_profileLength", C.c_int), ("xc1_profileWidth", C.c_int), ("xc1_iterations", C.c_int),
    ("gauss2D_iterations", C.c_int), ("gauss2D_sigma", C.c_float), ("downsample", C.c_int),
]


class QTrkSettings(C.Structure):
    _pack_ = 1
    _fields_ = _SETTINGS_FIELDS


class QTrkComputedConfig(C.Structure):
    _pack_ = 1
    _fields_ = _SETTINGS_FIELDS + [
        ("zlut_radialsteps", C.c_int), ("zlut_angularsteps", C.c_int), ("zlut_maxradius", C.c_float),
        ("qi_radialsteps", C.c_int), ("qi_angstepspq", C.c_int), ("qi_maxradius", C.c_float),
    ]


class ROIPosition(C.Structure):
    _pack_ = 1
    _fields_ = [("x", C.c_int), ("y", C.c_int)]


class ImageData(C.Structure):
    _pack_ = 4
    _fields_ = [("data", C.POINTER(C.c_float)), ("w", C.c_int), ("h", C.c_int)]


class ResultManagerConfig(C.Structure):
    """cputrack/qtrk_c_api.h:212-220"""
    _pack_ = 1
    _fields_ = [("numBeads", C.c_int), ("numFrameInfoColumns", C.c_int), ("scaling", Vector3f),
                ("offset", Vector3f), ("writeInterval", C.c_int), ("maxFramesInMemory", C.c_uint),
                ("binaryOutput", C.c_uint8)]


class RMFrameCounters(C.Structure):
    """cputrack/qtrk_c_api.h:221-230"""
    _pack_ = 1
    _fields_ = [("startFrame", C.c_int), ("processedFrames", C.c_int), ("lastSaveFrame", C.c_int),
                ("capturedFrames", C.c_int), ("localizationsDone", C.c_int), ("lostFrames", C.c_int),
                ("fileError", C.c_int)]


# numpy views of the packed records
JOB_DTYPE = np.dtype([("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                      ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
RESULT_DTYPE = np.dtype([("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                         ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4"),
                         ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                         ("comx", "<f4"), ("comy", "<f4"), ("error", "<u4"), ("imageMean", "<f4")])
assert JOB_DTYPE.itemsize == 24 and RESULT_DTYPE.itemsize == 52


class QTrkError(RuntimeError):
    pass


_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)

# name: (restype, argtypes) -- QTrkDotNet/QTrkDLL.cs:240-367 plus include/qtrk_amd.h
SIGNATURES = {
    "TestDLLCallConv": (None, [C.c_int]),
    "QTrkGetDefaultConfig": (None, [C.POINTER(QTrkSettings)]),
    "QTrkGetComputedConfig": (None, [C.POINTER(QTrkSettings), C.POINTER(QTrkComputedConfig)]),
    "QTrkCreateInstance": (_P, [C.POINTER(QTrkSettings)]),
    "QTrkFreeInstance": (None, [_P]),
    "QTrkSetLocalizationMode": (None, [_P, C.c_int]),
    "QTrkSetPixelCalibrationImages": (None, [_P, _P, _P]),
    "QTrkSetPixelCalibrationFactors": (None, [C.c_float, C.c_float]),
    "QTrkScheduleLocalization": (None, [_P, _P, C.c_int, C.c_int, C.POINTER(LocalizationJob)]),
    "QTrkClearResults": (None, [_P]),
    "QTrkFlush": (None, [_P]),
    "QTrkScheduleFrame": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                    C.POINTER(LocalizationJob)]),
    "QTrkEnableRadialZLUTCompareProfile": (None, [C.c_bool]),
    "QTrkGetRadialZLUTCompareProfile": (None, [_P]),
    "QTrkSetRadialZLUT": (None, [_P, _P, C.c_int, C.c_int]),
    "QTrkGetRadialZLUT": (None, [_P, _P]),
    "QTrkGetRadialZLUTSize": (None, [_P, _IP, _IP, _IP]),
    "QTrkSetRadialWeights": (None, [_P, _P]),
    "QTrkBeginLUT": (None, [_P, C.c_uint]),
    "QTrkBuildLUT": (None, [_P, _P, C.c_int, C.c_int, C.c_int, _P]),
    "QTrkBuildLUTFromFrame": (None, [_P, C.POINTER(ImageData), C.c_int, _P, C.c_int]),
    "QTrkFinalizeLUT": (None, [_P]),
    "QTrkGetResultCount": (C.c_int, [_P]),
    "QTrkFetchResults": (C.c_int, [_P, _P, C.c_int]),
    "QTrkGetQueueLength": (C.c_int, [_P, _IP]),
    "QTrkIsIdle": (C.c_bool, [_P]),
    "QTrkGetProfileReport": (None, [_P, C.c_char_p, C.c_int]),
    "QTrkGetWarnings": (None, [_P, C.c_char_p, C.c_int]),
    "QTrkFreeROIPositions": (None, [_P]),
    "QTrkSetImageZLUT": (C.c_int, [_P, _FP, _FP, _IP]),
    "QTrkGetImageZLUT": (None, [_P, _FP]),
    "QTrkGetImageZLUTSize": (None, [_P, _IP]),
    "QTrkGetImageZLUTDerivatives": (C.c_int, [_P, _P, _P]),
    "QTrkFindBeads": (_P, [C.POINTER(ImageData), C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _IP,
                           C.POINTER(ImageData)]),
    "ComputeRadialProfile": (None, [_FP, C.c_int, C.c_int, C.c_float, C.c_float, Vector2f,
                                    C.POINTER(ImageData), C.c_float, C.c_bool]),
    "NormalizeRadialProfile": (None, [_FP, C.c_int]),
    "NormalizeImage": (None, [C.POINTER(ImageData)]),
    "GenerateImageFromLUT": (None, [C.POINTER(ImageData), C.POINTER(ImageData), C.c_float, C.c_float,
                                    Vector3f, C.c_bool, C.c_int]),
    "ApplyPoissonNoise": (None, [C.POINTER(ImageData), C.c_float, C.c_float]),
    "ApplyGaussianNoise": (None, [C.POINTER(ImageData), C.c_float]),
    "GenerateGaussianSpotImage": (None, [C.POINTER(ImageData), Vector2f, C.c_float, C.c_float, C.c_float]),
    "GetPreciseTime": (C.c_double, []),
    # result sink: cputrack/qtrk_c_api.h:233-242
    "RMCreate": (_P, [C.c_char_p, C.c_char_p, C.POINTER(ResultManagerConfig), C.c_char_p]),
    "RMSetTracker": (None, [_P, _P]),
    "RMDestroy": (None, [_P]),
    "RMStoreFrameInfo": (None, [_P, C.c_int, C.c_double, _P]),
    "RMGetBeadResults": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P]),
    "RMGetFrameCounters": (None, [_P, C.POINTER(RMFrameCounters)]),
    "RMFlush": (None, [_P]),
    "RMGetResults": (C.c_int, [_P, C.c_int, C.c_int, _P]),
    "RMRemoveBead": (None, [_P, C.c_int]),
    "RMGetConfig": (None, [_P, C.POINTER(ResultManagerConfig)]),
    # include/qtrk_amd.h
    "QTrkGetLastError": (C.c_char_p, []),
    "QTrkClearLastError": (None, []),
    "QTrkGetDeviceInfo": (C.c_int, [C.c_char_p, C.c_int]),
    "QTrkSetConfigValue": (None, [_P, C.c_char_p, C.c_char_p]),
    "QTrkComputeZLUTBiasTable": (C.c_int, [_P, C.c_int, _P, C.c_int, C.c_bool]),
    "QTrkSetZLUTBiasTable": (C.c_int, [_P, _P, C.c_int, C.c_int]),
    "QTrkGetZLUTBiasTable": (C.c_int, [_P, _P, _P, _P]),
    "QTrkGetConfigValues": (None, [_P, C.c_char_p, C.c_int]),
    "QTrkBatchCreate": (_P, [_P, C.c_int, C.c_int]),
    "QTrkBatchFree": (None, [_P]),
    "QTrkBatchUpload": (C.c_int, [_P, _P, _P, C.c_int]),
    "QTrkBatchDeviceROIs": (_P, [_P]),
    "QTrkBatchDeviceJobs": (_P, [_P]),
    "QTrkBatchDeviceResults": (_P, [_P]),
    "QTrkBatchRun": (C.c_int, [_P, C.c_int, C.c_int]),
    "QTrkBatchSetMatchPrecision": (C.c_int, [_P, C.c_int]),
    "QTrkLocalizeDeviceNoAlloc": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, C.c_int, _P]),
    "QTrkBatchSync": (C.c_float, [_P]),
    "QTrkBatchDownload": (C.c_int, [_P, _P, C.c_int]),
    "QTrkBatchEnableDiagnostics": (C.c_int, [_P, C.c_bool]),
    "QTrkBatchDownloadDiagnostics": (C.c_int, [_P, _P, _P, C.c_int]),
    "QTrkGenerateImagesFromLUT": (None, [_FP, C.c_int, C.c_int, C.c_int, _FP, C.c_int, C.c_int, C.c_int,
                                         C.c_float, C.c_float, _FP, _IP, C.c_bool, C.c_int]),
    "QTrkLUTMatchCreate": (_P, [C.c_int, C.c_int, C.c_int]),
    "QTrkLUTMatchFree": (None, [_P]),
    "QTrkLUTMatchSetLUT": (C.c_int, [_P, _FP, _FP]),
    "QTrkLUTMatchUploadProfiles": (C.c_int, [_P, _FP, C.c_int]),
    "QTrkLUTMatchDeviceProfiles": (_FP, [_P, C.c_int]),
    "QTrkLUTMatchRun": (C.c_int, [_P, C.c_int, C.c_int]),
    "QTrkLUTMatchSync": (C.c_float, [_P]),
    "QTrkLUTMatchDownload": (C.c_int, [_P, _FP, _IP, _IP, C.c_int]),
    "QTrkLUTModelCreate": (_P, [_FP, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                C.c_bool]),
    "QTrkLUTModelFree": (None, [_P]),
    "QTrkLUTModelForward": (C.c_int, [_P, _P, _P, C.c_int, _P, _P, _P]),
    "QTrkLUTModelBackward": (C.c_int, [_P, _P, _P, C.c_int, _P, _P, _P]),
    "QTrkLUTModelFitMLE": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "QTrkLUTModelSync": (C.c_int, [_P, _P]),
    "QTrkLUTModelSetFitLanes": (C.c_int, [_P, C.c_int]),
    "QTrkPipelineCreate": (_P, []),
    "QTrkPipelineFree": (None, [_P]),
    "QTrkPipelineAddTracker": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _FP, C.c_int, C.c_int, C.c_int,
                                         C.c_float, C.c_float, C.c_int]),
    "QTrkPipelineSetJobs": (C.c_int, [_P, C.c_int, _P, C.c_int]),
    "QTrkPipelineDeviceROIs": (_P, [_P, C.c_int]),
    "QTrkPipelineUploadROIs": (C.c_int, [_P, C.c_int, _P, C.c_int]),
    "QTrkPipelineFinalize": (C.c_int, [_P, C.c_bool]),
    "QTrkPipelineLaunch": (C.c_int, [_P]),
    "QTrkPipelineSync": (C.c_float, [_P]),
    "QTrkPipelineDownload": (C.c_int, [_P, C.c_int, _P, _FP, _FP, C.c_int]),
    "QTrkPipelineStream": (C.c_int, [_P, C.c_int, C.c_float, _FP, _FP]),
    "QTrkPipelineGetReport": (C.c_int, [_P, C.c_char_p, C.c_int]),
}

MATCH_FP32, MATCH_BF16, MATCH_FP16 = 0, 1, 2
MATCH_FP32_GROUPED = 3  # device batches: profile kernel + grouped match in the reference arithmetic
NOALLOC_SCRATCH_MB = 32  # QTRK_NOALLOC_SCRATCH_MB (include/qtrk_amd.h)

_lib = None


def lib(path=None):
    """Loads libqtrk.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("QTRK_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise QTrkError(f"{p} not found: run __graft_entry__.build() (hipcc) first")
    L = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        if not hasattr(L, name):  # an older build: calling it raises AttributeError
            continue
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = L
    return L


def check():
    """Raises QTrkError if the last library call on this thread failed."""
    L = lib()
    msg = L.QTrkGetLastError()
    if msg:
        L.QTrkClearLastError()
        raise QTrkError(msg.decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _fptr(a):
    return a.ctypes.data_as(_FP)


def default_settings(**kw):
    s = QTrkSettings()
    lib().QTrkGetDefaultConfig(C.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def computed_config(settings):
    cc = QTrkComputedConfig()
    lib().QTrkGetComputedConfig(C.byref(settings), C.byref(cc))
    return cc


def make_jobs(n, frame=None, zlut_index=None, timestamp=0):
    j = np.zeros(n, dtype=JOB_DTYPE)
    j["frame"] = np.arange(n) if frame is None else frame
    j["timestamp"] = timestamp
    j["zlutIndex"] = 0 if zlut_index is None else zlut_index
    return j


class QTrkInstance:
    """A tracker instance (QTrkDotNet/QTrkInstance.cs:31-189)."""

    def __init__(self, settings):
        L = lib()
        L.QTrkClearLastError()
        self.settings = settings
        self.h = L.QTrkCreateInstance(C.byref(settings))
        check()
        if not self.h:
            raise QTrkError("QTrkCreateInstance returned NULL")
        self.cfg = computed_config(settings)

    def _call(self, name, *args):
        r = getattr(lib(), name)(self.h, *args)
        check()
        return r

    def free(self):
        if self.h:
            lib().QTrkFreeInstance(self.h)
            self.h = None
            check()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def set_localization_mode(self, mode):
        self._call("QTrkSetLocalizationMode", int(mode))

    def set_radial_zlut(self, lut):
        """lut: float32 [count][planes][zlut_radialsteps], or (count, planes) to allocate a zeroed one."""
        if isinstance(lut, tuple):
            self._call("QTrkSetRadialZLUT", None, int(lut[0]), int(lut[1]))
            return
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        assert lut.ndim == 3 and lut.shape[2] == self.cfg.zlut_radialsteps, lut.shape
        self._call("QTrkSetRadialZLUT", _ptr(lut), lut.shape[0], lut.shape[1])

    def set_image_zlut(self, image_lut, radial=None):
        """image_lut: float32 [beads][planes][height][width], or its 4 dims to allocate a zeroed one;
        radial: float32 [beads][planes][zlut_radialsteps] or None (zeroed)."""
        if isinstance(image_lut, tuple):
            dims, src = np.array(image_lut, np.int32), None
        else:
            src = np.ascontiguousarray(image_lut, np.float32)
            dims = np.array(src.shape, np.int32)
        rad = None if radial is None else np.ascontiguousarray(radial, np.float32)
        return bool(self._call("QTrkSetImageZLUT", None if src is None else _fptr(src),
                               None if rad is None else _fptr(rad), dims.ctypes.data_as(_IP)))

    def get_image_zlut(self):
        dims = np.zeros(4, np.int32)
        self._call("QTrkGetImageZLUTSize", dims.ctypes.data_as(_IP))
        out = np.zeros(tuple(int(d) for d in dims), np.float32)
        if out.size:
            self._call("QTrkGetImageZLUT", _fptr(out))
        return out

    def get_image_zlut_derivatives(self):
        """(dz, dz2) of the image LUT after FinalizeLUT ([beads][planes][h][w] each), or None."""
        dims = np.zeros(4, np.int32)
        lib().QTrkGetImageZLUTSize(self.h, dims.ctypes.data_as(_IP))
        dz = np.zeros(tuple(int(d) for d in dims), np.float32)
        dz2 = np.zeros_like(dz)
        r = self._call("QTrkGetImageZLUTDerivatives", _ptr(dz), _ptr(dz2))
        return (dz, dz2) if r == 1 else None

    def get_radial_zlut_size(self):
        c, p, r = C.c_int(), C.c_int(), C.c_int()
        self._call("QTrkGetRadialZLUTSize", C.byref(c), C.byref(p), C.byref(r))
        return c.value, p.value, r.value

    def get_radial_zlut(self):
        c, p, r = self.get_radial_zlut_size()
        out = np.zeros((c, p, r), np.float32)
        if out.size:
            self._call("QTrkGetRadialZLUT", _ptr(out))
        return out

    def compute_zlut_bias_table(self, bias_planes, smp_per_pixel=4, spline=True):
        """QueuedTracker::ComputeZBiasCorrection on the GPU; the table becomes active.
        Returns it as float32 [beads][bias_planes]."""
        c, _, _ = self.get_radial_zlut_size()
        out = np.zeros((c, bias_planes), np.float32)
        self._call("QTrkComputeZLUTBiasTable", int(bias_planes), _ptr(out), int(smp_per_pixel), bool(spline))
        return out

    def set_zlut_bias_table(self, table):
        """Installs a [beads][bias_planes] Z-bias table (None clears it)."""
        if table is None:
            self._call("QTrkSetZLUTBiasTable", None, 0, 0)
            return
        t = np.ascontiguousarray(table, np.float32)
        self._call("QTrkSetZLUTBiasTable", _ptr(t), t.shape[1], t.shape[0])

    def get_zlut_bias_table(self):
        """The active Z-bias table [beads][bias_planes], or None."""
        w, h = C.c_int(), C.c_int()
        if not self._call("QTrkGetZLUTBiasTable", None, C.byref(w), C.byref(h)):
            return None
        out = np.zeros((h.value, w.value), np.float32)
        self._call("QTrkGetZLUTBiasTable", _ptr(out), C.byref(w), C.byref(h))
        return out

    def set_radial_weights(self, w):
        if w is None:
            self._call("QTrkSetRadialWeights", None)
        else:
            w = np.ascontiguousarray(w, np.float32)
            self._w = w
            self._call("QTrkSetRadialWeights", _ptr(w))

    def set_pixel_calibration_images(self, offset, gain):
        o = None if offset is None else np.ascontiguousarray(offset, np.float32)
        g = None if gain is None else np.ascontiguousarray(gain, np.float32)
        self._call("QTrkSetPixelCalibrationImages", None if o is None else _ptr(o), None if g is None else _ptr(g))

    def schedule_localization(self, roi, job):
        """roi: 2-D uint8/uint16/float32 array (any row pitch); job: LocalizationJob or JOB_DTYPE record."""
        roi = np.asarray(roi)
        pdt = PDT_OF_DTYPE[roi.dtype]
        if roi.strides[1] != roi.itemsize:
            roi = np.ascontiguousarray(roi)
        if not isinstance(job, LocalizationJob):
            j = LocalizationJob()
            C.memmove(C.byref(j), np.asarray(job, JOB_DTYPE).tobytes(), 24)
            job = j
        self._call("QTrkScheduleLocalization", _ptr(roi), roi.strides[0], pdt, C.byref(job))

    def schedule_frame(self, frame, positions, job):
        frame = np.asarray(frame)
        pdt = PDT_OF_DTYPE[frame.dtype]
        pos = np.ascontiguousarray(positions, dtype=np.int32)
        j = LocalizationJob()
        C.memmove(C.byref(j), np.asarray(job, JOB_DTYPE).tobytes(), 24)
        return self._call("QTrkScheduleFrame", _ptr(frame), frame.strides[0], frame.shape[1], frame.shape[0],
                          _ptr(pos), len(pos), pdt, C.byref(j))

    def flush(self):
        self._call("QTrkFlush")

    def clear_results(self):
        self._call("QTrkClearResults")

    def result_count(self):
        return self._call("QTrkGetResultCount")

    def fetch_results(self, max_results=None):
        n = self.result_count() if max_results is None else max_results
        out = np.zeros(max(n, 0), RESULT_DTYPE)
        got = self._call("QTrkFetchResults", _ptr(out), n) if n > 0 else 0
        return out[:got]

    def queue_length(self):
        m = C.c_int()
        n = self._call("QTrkGetQueueLength", C.byref(m))
        return n, m.value

    def is_idle(self):
        return bool(self._call("QTrkIsIdle"))

    def begin_lut(self, flags):
        self._call("QTrkBeginLUT", int(flags))

    def build_lut(self, images, plane, known_pos=None):
        """images: [count][h][w] stack, image i -> bead i."""
        images = np.ascontiguousarray(images)
        pdt = PDT_OF_DTYPE[images.dtype]
        kp = None
        if known_pos is not None:
            kp = np.ascontiguousarray(known_pos, np.float32)
        self._call("QTrkBuildLUT", _ptr(images), images.strides[1], pdt, int(plane),
                   None if kp is None else _ptr(kp))

    def finalize_lut(self):
        self._call("QTrkFinalizeLUT")

    def set_config_value(self, name, value):
        self._call("QTrkSetConfigValue", str(name).encode(), str(value).encode())

    def get_config_values(self):
        buf = C.create_string_buffer(4096)
        self._call("QTrkGetConfigValues", buf, 4096)
        return dict(l.split("=", 1) for l in buf.value.decode().splitlines() if "=" in l)

    def profile_report(self):
        buf = C.create_string_buffer(4096)
        self._call("QTrkGetProfileReport", buf, 4096)
        return buf.value.decode()

    def warnings(self):
        buf = C.create_string_buffer(8192)
        self._call("QTrkGetWarnings", buf, 8192)
        return buf.value.decode()

    def wait_for(self, n, timeout=120.0):
        """Flush, then poll until n results exist (cputrack-test/SharedTests.h:300-314)."""
        import time
        self.flush()
        t0 = time.time()
        while self.result_count() < n:
            if time.time() - t0 > timeout:
                raise QTrkError(f"timeout waiting for {n} results")
            time.sleep(0.0005)

    def localize_all(self, rois, jobs):
        """Schedules every ROI, waits, and returns results sorted back into job order."""
        for i in range(len(rois)):
            self.schedule_localization(rois[i], jobs[i])
        self.wait_for(len(rois))
        res = self.fetch_results(len(rois))
        key = {(int(j["frame"]), int(j["zlutIndex"])): i for i, j in enumerate(jobs)}
        order = np.array([key[(int(r["frame"]), int(r["zlutIndex"]))] for r in res])
        out = np.zeros(len(rois), RESULT_DTYPE)
        out[order] = res
        return out


def localize_device_noalloc(trk, rois_ptr, pdt, jobs_ptr, results_ptr, n, precision=MATCH_FP32, stream=None):
    """QTrkLocalizeDeviceNoAlloc on GPU buffers given as device addresses (ints); asynchronous on
    `stream` (a hipStream_t handle as an int, or None for the default stream)."""
    if lib().QTrkLocalizeDeviceNoAlloc(trk.h, C.c_void_p(rois_ptr), int(pdt), C.c_void_p(jobs_ptr),
                                       C.c_void_p(results_ptr), int(n), int(precision),
                                       None if stream is None else C.c_void_p(stream)) != 0:
        check()


class DeviceBatch:
    """ROIs resident in GPU memory, localized with a tracker's settings (include/qtrk_amd.h)."""

    def __init__(self, trk, max_rois, pdt):
        self.trk = trk
        self.h = lib().QTrkBatchCreate(trk.h, int(max_rois), int(pdt))
        check()
        self.max = max_rois

    def upload(self, rois, jobs):
        rois = np.ascontiguousarray(rois)
        jobs = np.ascontiguousarray(jobs, JOB_DTYPE)
        if lib().QTrkBatchUpload(self.h, _ptr(rois), _ptr(jobs), len(jobs)) != 0:
            check()

    def run(self, n, repeats=1):
        if lib().QTrkBatchRun(self.h, int(n), int(repeats)) != 0:
            check()

    def sync(self):
        ms = lib().QTrkBatchSync(self.h)
        if ms < 0:
            check()
        return ms

    def download(self, n):
        out = np.zeros(n, RESULT_DTYPE)
        if lib().QTrkBatchDownload(self.h, _ptr(out), int(n)) != 0:
            check()
        return out

    def set_match_precision(self, precision):
        """MATCH_FP32 (per-ROI compare in the localization kernel), MATCH_BF16 / MATCH_FP16
        (profile kernel + grouped matrix-core match) or MATCH_FP32_GROUPED (profile kernel + grouped
        match in the reference arithmetic)."""
        if lib().QTrkBatchSetMatchPrecision(self.h, int(precision)) != 0:
            check()

    def enable_diagnostics(self, on=True):
        if lib().QTrkBatchEnableDiagnostics(self.h, bool(on)) != 0:
            check()

    def diagnostics(self, n):
        """(best plane per ROI, radial profile per ROI) of the last run."""
        best = np.zeros(n, np.int32)
        prof = np.zeros((n, max(self.trk.cfg.zlut_radialsteps, 1)), np.float32)
        if lib().QTrkBatchDownloadDiagnostics(self.h, _ptr(best), _ptr(prof), int(n)) != 0:
            check()
        return best, prof

    def free(self):
        if self.h:
            lib().QTrkBatchFree(self.h)
            self.h = None


def generate_images_from_lut(luts, pos, lut_index, width, height, minr, maxr, spline=True, ovs=4):
    """Batch GenerateImageFromLUT on the host (include/qtrk_amd.h)."""
    luts = np.ascontiguousarray(luts, np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    li = np.ascontiguousarray(lut_index, np.int32)
    n = len(pos)
    out = np.zeros((n, height, width), np.float32)
    lib().QTrkGenerateImagesFromLUT(_fptr(out), n, width, height, _fptr(luts), luts.shape[0], luts.shape[1],
                                    luts.shape[2], minr, maxr, _fptr(pos), li.ctypes.data_as(_IP), spline, ovs)
    check()
    return out


def find_beads(img, sample_corner, roi, img_rel_dist, acceptance, want_sample=False):
    """QTrkFindBeads (cputrack/qtrk_c_api.cpp:33-58) on a 2-D float frame: the beads' ROI corners as
    an (n, 2) int32 array (x, y), in the library's order; with want_sample also the roi x roi sample
    block it copied out."""
    im = np.ascontiguousarray(img, np.float32)
    h, w = im.shape
    idata = ImageData(_fptr(im), w, h)
    smp = np.zeros((roi, roi), np.float32)
    sdata = ImageData(_fptr(smp), roi, roi)
    n = C.c_int(0)
    L = lib()
    p = L.QTrkFindBeads(C.byref(idata), int(sample_corner[0]), int(sample_corner[1]), int(roi),
                        float(img_rel_dist), float(acceptance), C.byref(n), C.byref(sdata) if want_sample else None)
    check()
    try:
        arr = C.cast(p, C.POINTER(ROIPosition * max(n.value, 1))).contents
        out = np.array([(arr[i].x, arr[i].y) for i in range(n.value)], np.int32).reshape(-1, 2)
    finally:
        L.QTrkFreeROIPositions(p)
    return (out, smp) if want_sample else out


class LUTMatch:
    """Grouped radial Z-LUT match of many profiles per bead (include/qtrk_amd.h)."""

    def __init__(self, luts, radial_weights=None):
        luts = np.ascontiguousarray(luts, np.float32)
        self.beads, self.planes, self.res = luts.shape
        self.h = lib().QTrkLUTMatchCreate(self.beads, self.planes, self.res)
        check()
        self.frames = 0
        self.set_lut(luts, radial_weights)

    def set_lut(self, luts, radial_weights=None):
        luts = np.ascontiguousarray(luts, np.float32)
        w = None if radial_weights is None else np.ascontiguousarray(radial_weights, np.float32)
        if lib().QTrkLUTMatchSetLUT(self.h, _fptr(luts), None if w is None else _fptr(w)) != 0:
            check()

    def upload(self, profiles):
        """profiles: [beads][frames][radialsteps] float32."""
        p = np.ascontiguousarray(profiles, np.float32)
        assert p.shape[0] == self.beads and p.shape[2] == self.res, p.shape
        if lib().QTrkLUTMatchUploadProfiles(self.h, _fptr(p), p.shape[1]) != 0:
            check()
        self.frames = p.shape[1]

    def run(self, precision=MATCH_FP32, repeats=1):
        if lib().QTrkLUTMatchRun(self.h, int(precision), int(repeats)) != 0:
            check()

    def sync(self):
        ms = lib().QTrkLUTMatchSync(self.h)
        if ms < 0:
            check()
        return ms

    def download(self):
        """(z [beads][frames], best plane, candidates [beads][frames][8])."""
        n = self.beads * self.frames
        z = np.zeros(n, np.float32)
        best = np.zeros(n, np.int32)
        cand = np.full((n, 8), -1, np.int32)
        if lib().QTrkLUTMatchDownload(self.h, _fptr(z), best.ctypes.data_as(_IP), cand.ctypes.data_as(_IP), n) != 0:
            check()
        shp = (self.beads, self.frames)
        return z.reshape(shp), best.reshape(shp), cand.reshape(shp + (8,))

    def match(self, profiles, precision=MATCH_FP32):
        self.upload(profiles)
        self.run(precision)
        self.sync()
        return self.download()

    def free(self):
        if self.h:
            lib().QTrkLUTMatchFree(self.h)
            self.h = None


def _dptr(t):
    """Device pointer of a torch CUDA tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


class LUTModel:
    """Bead-image model rendered from radial Z-LUTs: forward, adjoint and Poisson MLE with Fisher/CRLB
    (include/qtrk_amd.h).  Arrays are torch CUDA tensors; calls run on torch's current stream."""

    def __init__(self, luts, minr, maxr, width, height, ovs=4, spline=True):
        luts = np.ascontiguousarray(luts, np.float32)
        if luts.ndim == 2:
            luts = luts[None]
        self.count, self.planes, self.steps = luts.shape
        self.width, self.height = width, height
        self.h = lib().QTrkLUTModelCreate(_fptr(luts), self.count, self.planes, self.steps, minr, maxr, width,
                                          height, ovs, bool(spline))
        check()

    @staticmethod
    def _stream():
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _prep(self, theta, lut_index):
        import torch
        theta = theta.to(device="cuda", dtype=torch.float32).contiguous()
        idx = None if lut_index is None else torch.as_tensor(lut_index, dtype=torch.int32, device="cuda").contiguous()
        return theta, idx

    def forward(self, theta, lut_index=None, jacobian=False):
        """mu [n][H][W] (and d mu / d theta [n][5][H][W]) for theta [n][5] = (x, y, z, I, bg)."""
        import torch
        theta, idx = self._prep(theta, lut_index)
        n = theta.shape[0]
        mu = torch.empty((n, self.height, self.width), device="cuda")
        jac = torch.empty((n, 5, self.height, self.width), device="cuda") if jacobian else None
        if lib().QTrkLUTModelForward(self.h, _dptr(theta), _dptr(idx), n, _dptr(mu), _dptr(jac), self._stream()):
            check()
        return (mu, jac) if jacobian else mu

    def backward(self, theta, grad_mu, lut_index=None):
        """sum_px grad_mu * d mu / d theta: [n][5]."""
        import torch
        theta, idx = self._prep(theta, lut_index)
        g = grad_mu.to(device="cuda", dtype=torch.float32).contiguous()
        out = torch.empty((theta.shape[0], 5), device="cuda")
        if lib().QTrkLUTModelBackward(self.h, _dptr(theta), _dptr(idx), theta.shape[0], _dptr(g), _dptr(out),
                                      self._stream()):
            check()
        return out

    def fit(self, data, theta0, iterations=10, lut_index=None):
        """Poisson MLE from theta0: (theta [n][5], Fisher [n][5][5] float64, CRLB [n][5], NLL [n])."""
        import torch
        theta0, idx = self._prep(theta0, lut_index)
        d = data.to(device="cuda", dtype=torch.float32).contiguous()
        n = theta0.shape[0]
        th = torch.empty((n, 5), device="cuda")
        fisher = torch.empty((n, 5, 5), device="cuda", dtype=torch.float64)
        crlb = torch.empty((n, 5), device="cuda")
        nll = torch.empty(n, device="cuda")
        if lib().QTrkLUTModelFitMLE(self.h, _dptr(d), _dptr(theta0), _dptr(idx), n, int(iterations), _dptr(th),
                                    _dptr(fisher), _dptr(crlb), _dptr(nll), self._stream()):
            check()
        return th, fisher, crlb, nll

    def set_fit_lanes(self, lanes):
        """Lanes per ROI of fit(): 64, 128, 256 (default), 512 (models of <= 6x6 sub-samples), 1024 (<= 4x4)."""
        if lib().QTrkLUTModelSetFitLanes(self.h, int(lanes)):
            check()

    def render(self, theta, lut_index=None):
        """mu as a differentiable function of theta (torch autograd through the adjoint kernel)."""
        return _LUTRender.apply(theta, self, lut_index)

    def free(self):
        if self.h:
            lib().QTrkLUTModelFree(self.h)
            self.h = None


def _make_render_fn():
    try:
        import torch
    except ImportError:  # pragma: no cover
        return None

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model, lut_index):
            ctx.model, ctx.lut_index = model, lut_index
            ctx.save_for_backward(theta)
            return model.forward(theta, lut_index)

        @staticmethod
        def backward(ctx, grad_mu):
            (theta,) = ctx.saved_tensors
            return ctx.model.backward(theta, grad_mu, ctx.lut_index).to(theta.dtype), None, None

    return _Render


_LUTRender = _make_render_fn()


class Pipeline:
    """Several trackers' COM -> QI -> grouped Z-LUT match -> LUT-model MLE chains in one hipGraph
    (include/qtrk_amd.h)."""

    def __init__(self):
        self.h = lib().QTrkPipelineCreate()
        check()
        self.slots = []

    def add(self, trk, rois_per_batch, pdt, precision=MATCH_BF16, model_luts=None, model_minr=0.0,
            model_maxr=0.0, iterations=5):
        ml = None
        cnt = planes = steps = 0
        if model_luts is not None:
            ml = np.ascontiguousarray(model_luts, np.float32)
            if ml.ndim == 2:
                ml = ml[None]
            cnt, planes, steps = ml.shape
        slot = lib().QTrkPipelineAddTracker(self.h, trk.h, int(rois_per_batch), int(pdt), int(precision),
                                            None if ml is None else _fptr(ml), cnt, planes, steps, model_minr,
                                            model_maxr, int(iterations))
        if slot < 0:
            check()
        self.slots.append((trk, rois_per_batch, ml is not None))
        return slot

    def set_jobs(self, slot, jobs):
        jobs = np.ascontiguousarray(jobs, JOB_DTYPE)
        if lib().QTrkPipelineSetJobs(self.h, slot, _ptr(jobs), len(jobs)) != 0:
            check()

    def upload(self, slot, rois):
        rois = np.ascontiguousarray(rois)
        if lib().QTrkPipelineUploadROIs(self.h, slot, _ptr(rois), len(rois)) != 0:
            check()

    def finalize(self, autotune=True):
        if lib().QTrkPipelineFinalize(self.h, bool(autotune)) != 0:
            check()

    def launch(self):
        if lib().QTrkPipelineLaunch(self.h) != 0:
            check()

    def sync(self):
        ms = lib().QTrkPipelineSync(self.h)
        if ms < 0:
            check()
        return ms

    def download(self, slot):
        """(results with the matched z, MLE theta [n][5] or None, CRLB [n][5] or None)."""
        n, mle = self.slots[slot][1], self.slots[slot][2]
        res = np.zeros(n, RESULT_DTYPE)
        th = np.zeros((n, 5), np.float32) if mle else None
        cr = np.zeros((n, 5), np.float32) if mle else None
        if lib().QTrkPipelineDownload(self.h, slot, _ptr(res), None if th is None else _fptr(th),
                                      None if cr is None else _fptr(cr), n) != 0:
            check()
        return res, th, cr

    def stream(self, batches, period_ms):
        """Camera-paced launches: (latency ms per batch, GPU ms per batch)."""
        lat = np.zeros(batches, np.float32)
        gpu = np.zeros(batches, np.float32)
        if lib().QTrkPipelineStream(self.h, int(batches), float(period_ms), _fptr(lat), _fptr(gpu)) != 0:
            check()
        return lat, gpu

    def report(self):
        buf = C.create_string_buffer(8192)
        if lib().QTrkPipelineGetReport(self.h, buf, 8192) != 0:
            check()
        return buf.value.decode()

    def free(self):
        if self.h:
            lib().QTrkPipelineFree(self.h)
            self.h = None


class ResultManager:
    """The result sink after FetchResults (RM* of cputrack/qtrk_c_api.h:233-242): a native thread
    drains the attached tracker into per-frame tables and appends finished frames to a text or
    binary trace file."""

    def __init__(self, file, frameinfo, num_beads, info_columns=(), scaling=(1, 1, 1), offset=(0, 0, 0),
                 write_interval=1, max_frames_in_memory=0, binary=False):
        cfg = ResultManagerConfig()
        cfg.numBeads = num_beads
        cfg.numFrameInfoColumns = len(info_columns)
        cfg.scaling = Vector3f(*scaling)
        cfg.offset = Vector3f(*offset)
        cfg.writeInterval = write_interval
        cfg.maxFramesInMemory = max_frames_in_memory
        cfg.binaryOutput = 1 if binary else 0
        L = lib()
        L.QTrkClearLastError()
        self.num_beads, self.ncols = num_beads, len(info_columns)
        self.h = L.RMCreate((file or "").encode(), (frameinfo or "").encode(), C.byref(cfg),
                            ";".join(info_columns).encode())
        check()
        if not self.h:
            raise QTrkError("RMCreate returned NULL")

    def _call(self, name, *args):
        r = getattr(lib(), name)(self.h, *args)
        check()
        return r

    def set_tracker(self, trk):
        self._call("RMSetTracker", trk.h if trk is not None else None)

    def store_frame_info(self, frame, timestamp, cols=()):
        c = np.ascontiguousarray(cols, np.float32) if self.ncols else np.zeros(1, np.float32)
        self._call("RMStoreFrameInfo", int(frame), float(timestamp), _ptr(c))

    def counters(self):
        fc = RMFrameCounters()
        self._call("RMGetFrameCounters", C.byref(fc))
        return {k: getattr(fc, k) for k, _ in RMFrameCounters._fields_}

    def config(self):
        cfg = ResultManagerConfig()
        self._call("RMGetConfig", C.byref(cfg))
        return cfg

    def flush(self):
        self._call("RMFlush")

    def results(self, start_frame, num_frames):
        out = np.zeros(num_frames * self.num_beads, RESULT_DTYPE)
        n = self._call("RMGetResults", int(start_frame), int(num_frames), _ptr(out))
        return n, out.reshape(num_frames, self.num_beads)

    def bead_results(self, start, num_frames, bead):
        out = np.zeros(max(num_frames, 0), RESULT_DTYPE)
        n = self._call("RMGetBeadResults", int(start), int(num_frames), int(bead), _ptr(out))
        return out[:max(n, 0)]

    def remove_bead(self, bead):
        self._call("RMRemoveBead", int(bead))

    def free(self):
        if self.h:
            lib().RMDestroy(self.h)
            self.h = None
            check()

