/*
 * QueuedTracker.h -- C++ surface of the tracker (for C++ callers such as the
 * LabVIEW wrappers, which reach the implementation through
 * CreateQueuedTracker and read the public `cfg` member).
 *
 * Mirrors the abstract interface of the reference, cputrack/QueuedTracker.h:22-99:
 * same virtual functions, same argument meaning, same public data member.
 * The only implementation shipped by libqtrk.so is the MI355X tracker
 * (QueuedHIPTracker); CreateQueuedTracker and QTrkCreateInstance return it.
 */
#ifndef QTRK_QUEUED_TRACKER_H
#define QTRK_QUEUED_TRACKER_H

#include "qtrk_c_api.h"

#include <map>
#include <string>
#include <vector>

/* Minimum samples per radial bin before the image mean is used instead
 * (cputrack/QueuedTracker.h:18). */
#define MIN_RADPROFILE_SMP_COUNT 4

/* Peak-fit weights (cputrack/QueuedTracker.h:106-116). */
#define QI_LSQFIT_WEIGHTS { 0.14f, 0.5f, 0.85f, 1.0f, 0.85f, 0.5f, 0.14f }
#define QI_LSQFIT_NWEIGHTS 7
#define ZLUT_LSQFIT_NWEIGHTS 7
#define ZLUT_LSQFIT_WEIGHTS { 0.15f, 0.5f, 0.85f, 1.0f, 0.85f, 0.5f, 0.15f }

/* cputrack/QueuedTracker.h:121-124 */
inline int PDT_BytesPerPixel(QTRK_PixelDataType pdt)
{
	static const int b[] = { 1, 2, 4 };
	return b[(int)pdt];
}

/* QTrkComputedConfig::Update (cputrack/QueuedTracker.cpp:6-21) as a free
 * function, since the struct is shared with C. */
void QTrkComputedConfigUpdate(QTrkComputedConfig* cfg);
QTrkComputedConfig QTrkComputeConfig(const QTrkSettings& s);
int NearestPowerOf2(int v); /* cputrack/utils.cpp:644-652 */

class CImageData; /* ImageData.h */
typedef unsigned char uchar;

class QueuedTracker {
public:
	QueuedTracker();
	virtual ~QueuedTracker();

	virtual void SetLocalizationMode(LocMode_t locType) = 0;
	virtual void SetPixelCalibrationImages(float* offset, float* gain) = 0;
	virtual void SetPixelCalibrationFactors(float offsetFactor, float gainFactor) = 0;

	virtual void ScheduleLocalization(void* data, int pitch, QTRK_PixelDataType pdt, const LocalizationJob* jobInfo) = 0;
	virtual void ClearResults() = 0;
	virtual void Flush() = 0;

	/* Extracts each ROI of a frame and schedules it (cputrack/QueuedTracker.cpp:91-111). */
	virtual int ScheduleFrame(void* imgptr, int pitch, int width, int height, ROIPosition* positions, int numROI,
	                          QTRK_PixelDataType pdt, const LocalizationJob* jobInfo);

	virtual void SetRadialZLUT(float* data, int count, int planes) = 0;
	virtual void GetRadialZLUT(float* dst) = 0;
	virtual void GetRadialZLUTSize(int& count, int& planes, int& radialsteps) = 0;
	virtual void SetRadialWeights(float* zcmp) = 0;

	virtual void EnableRadialZLUTCompareProfile(bool enabled) = 0;
	virtual void GetRadialZLUTCompareProfile(float* dst) = 0;

	virtual void GetImageZLUTSize(int* dims) {}
	virtual void GetImageZLUT(float* dst) {}
	virtual bool SetImageZLUT(float* dst, float* radial_zlut, int* dims) { return false; }

	virtual void BeginLUT(uint flags) = 0;
	virtual void BuildLUT(void* data, int pitch, QTRK_PixelDataType pdt, int plane, vector2f* known_pos = 0) = 0;
	virtual void FinalizeLUT() = 0;

	virtual int GetResultCount() = 0;
	virtual int FetchResults(LocalizationResult* results, int maxResults) = 0;

	virtual int GetQueueLength(int* maxQueueLen = 0) = 0;
	virtual bool IsIdle() = 0;

	virtual void SetConfigValue(std::string name, std::string value) = 0;
	typedef std::map<std::string, std::string> ConfigValueMap;
	virtual ConfigValueMap GetConfigValues() = 0;

	virtual std::string GetProfileReport() { return ""; }
	virtual std::string GetWarnings() { return ""; }
	virtual bool GetDebugImage(int ID, int* w, int* h, float** pData) { return false; }

	QTrkComputedConfig cfg;

	/* The reference's non-virtual members (cputrack/QueuedTracker.h:38, 87-92), which its
	 * LabVIEW wrapper calls (cputrack/lv_queuetrk_api.cpp:461, 478, 613):
	 *  - ScheduleImageData: a float ROI (cputrack/QueuedTracker.cpp:64-67);
	 *  - ScheduleLocalization with the job's fields as arguments (:69-77; initial may be NULL);
	 *  - ComputeZBiasCorrection: the table of QTrkComputeZLUTBiasTable, which becomes the active
	 *    correction, copied into *result (w = bias planes, h = beads) when result is not NULL;
	 *  - SetZLUTBiasCorrection / GetZLUTBiasCorrection: the table as an image, w = bias planes,
	 *    h = beads; Get returns a new copy (delete it) or NULL when none is set;
	 *  - DebugImage: GetDebugImage's image (none here: an empty image). */
	void ScheduleImageData(ImageData* data, const LocalizationJob* jobInfo);
	void ScheduleLocalization(uchar* data, int pitch, QTRK_PixelDataType pdt, uint frame, uint timestamp,
	                          vector3f* initial, uint zlutIndex);
	void ComputeZBiasCorrection(int bias_planes, CImageData* result, int smpPerPixel, bool useSplineInterp);
	void SetZLUTBiasCorrection(const CImageData& data);
	CImageData* GetZLUTBiasCorrection();
	ImageData DebugImage(int ID);

	/* Z-bias table: w = planes of the table, h = beads
	 * (cputrack/QueuedTracker.cpp:115-140, 192-203). */
	float ZLUTBiasCorrection(float z, int zlut_planes, int bead);
	void SetZLUTBiasCorrection(const float* table, int tablePlanes, int beads);
	bool HasZLUTBiasCorrection() const { return !zbias.empty(); }
	/* GetZLUTBiasCorrection (cputrack/QueuedTracker.cpp:199-203) as a copy into dst (if not
	 * NULL) of the [beads][tablePlanes] table; returns false when no table is set. */
	bool GetZLUTBiasCorrection(float* dst, int* tablePlanes, int* beads) const;

protected:
	std::vector<float> zbias;
	int zbias_w = 0, zbias_h = 0;
};

/* Factory used by C++ callers (cputrack/QueuedTracker.h:99). */
QueuedTracker* CreateQueuedTracker(const QTrkComputedConfig& cc);

#endif
