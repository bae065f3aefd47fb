"""MI355X bead tracker: Python access to libqtrk.so.

The product is the C ABI in include/ implemented by qtrk_amd/libqtrk.so
(host runtime in C++, kernels in HIP for gfx950).  This package is a thin
ctypes layer over it (mirroring the reference's C# binding) plus the
synthetic-workload generator used by the benchmark.
"""
from .api import (  # noqa: F401
    BUILDLUT_NORMALIZE, JOB_DTYPE, LT_ClearFirstFourPixels, LT_Gaussian2D, LT_LocalizeZ, LT_NormalizeProfile,
    LT_OnlyCOM, LT_QI, LT_XCor1D, LT_ZLUTAlign, QTrkError, QTrkFloat, QTrkInstance, QTrkU8, QTrkU16,
    RESULT_DTYPE, DeviceBatch, computed_config, default_settings, lib, make_jobs,
)
