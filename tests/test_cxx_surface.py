"""The C++ surface of QueuedTracker that C++ callers (the reference's LabVIEW wrapper,
cputrack/lv_queuetrk_api.cpp) use besides the virtual interface: the non-virtual members
declared at cputrack/QueuedTracker.h:38, 87-92 and the factory (:99).

CPU: a program that declares those members exactly as the reference's header does (its own
stand-in declarations of the types: vector3<float>, TImageData<float>, CImageData) links against
libqtrk.so, so a caller built against the reference's headers finds every symbol.
GPU: tests/cxx_surface_check.cpp exercises them on the MI355X tracker.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "qtrk_amd")

# Declarations as in cputrack/QueuedTracker.h:12-99 and cputrack/qtrk_c_api.h (types declared only)
REF_DECLS = r"""
#include <map>
#include <string>
template <typename T> struct vector2;
template <typename T> struct vector3;
typedef vector2<float> vector2f;
typedef vector3<float> vector3f;
template <typename T> struct TImageData;
typedef TImageData<float> ImageData;
class CImageData;
typedef unsigned char uchar;
typedef unsigned int uint;
enum QTRK_PixelDataType { QTrkU8 = 0, QTrkU16 = 1, QTrkFloat = 2 };
struct LocalizationJob;
struct ROIPosition;
struct QTrkComputedConfig;
class QueuedTracker {
public:
	QueuedTracker();
	virtual ~QueuedTracker();
	virtual int ScheduleFrame(void* imgptr, int pitch, int width, int height, ROIPosition* positions, int numROI,
	                          QTRK_PixelDataType pdt, const LocalizationJob* jobInfo);
	void ScheduleImageData(ImageData* data, const LocalizationJob* jobInfo);
	ImageData DebugImage(int ID);
	void ScheduleLocalization(uchar* data, int pitch, QTRK_PixelDataType pdt, uint frame, uint timestamp,
	                          vector3f* initial, uint zlutIndex);
	void ComputeZBiasCorrection(int bias_planes, CImageData* result, int smpPerPixel, bool useSplineInterp);
	float ZLUTBiasCorrection(float z, int zlut_planes, int bead);
	void SetZLUTBiasCorrection(const CImageData& data);
	CImageData* GetZLUTBiasCorrection();
};
QueuedTracker* CreateQueuedTracker(const QTrkComputedConfig& cc);
"""

REF_USE = r"""
#include <cstdio>
int main(int argc, char**)
{
	void* p[] = {
		(void*)&CreateQueuedTracker,
	};
	auto a = &QueuedTracker::ScheduleImageData;
	auto b = &QueuedTracker::DebugImage;
	void (QueuedTracker::*c)(uchar*, int, QTRK_PixelDataType, uint, uint, vector3f*, uint) =
	    &QueuedTracker::ScheduleLocalization;
	auto d = &QueuedTracker::ComputeZBiasCorrection;
	auto e = &QueuedTracker::ZLUTBiasCorrection;
	auto f = &QueuedTracker::SetZLUTBiasCorrection;
	auto g = &QueuedTracker::GetZLUTBiasCorrection;
	auto h = &QueuedTracker::ScheduleFrame;
	if (argc > 99) std::printf("%p %d %d %d %d %d %d %d %d\n", p[0], a != nullptr, b != nullptr, c != nullptr,
	                           d != nullptr, e != nullptr, f != nullptr, g != nullptr, h != nullptr);
	return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_members_link_under_the_reference_declarations(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libqtrk.so")):
        pytest.skip("libqtrk.so not built")
    src = tmp_path / "refnames.cpp"
    src.write_text(REF_DECLS + REF_USE)
    exe = tmp_path / "refnames"
    r = subprocess.run(["g++", "-std=c++14", str(src), "-L" + LIBDIR, "-lqtrk", "-Wl,--no-undefined",
                        "-Wl,-rpath," + LIBDIR, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cxx_members_on_the_gpu_tracker(gpu):
    exe = os.path.join(ROOT, "tests", "_bin", "cxx_surface_check")
    assert os.path.exists(exe), "run the build step (builds tests/_bin/cxx_surface_check)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
