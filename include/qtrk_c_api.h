/*
 * qtrk_c_api.h -- C ABI of the MI355X bead tracker.
 *
 * Drop-in replacement for the reference's cputrack/qtrk_c_api.h: same exported
 * names, same argument meaning, byte-identical packed structs, so the C#
 * QTrkDotNet P/Invoke layer (QTrkDotNet/QTrkDLL.cs:240-424) and the LabVIEW
 * wrappers bind to libqtrk.so unchanged.  Each declaration cites the
 * reference declaration it replaces.
 *
 * Differences from the reference header, all ABI-neutral:
 *   - usable from C as well as C++ (no constructors in the structs; defaults
 *     come from QTrkGetDefaultConfig, as the C# caller already does);
 *   - no __declspec / __cdecl: on x86-64 Linux there is one C calling
 *     convention and symbols have default visibility;
 *   - exceptions never cross this boundary.  A failing call records a
 *     message readable with QTrkGetLastError() (include/qtrk_amd.h) and
 *     prints it to stderr; constructors that fail return NULL.
 */
#ifndef QTRK_C_API_H
#define QTRK_C_API_H

#include <stdint.h>

#ifdef __cplusplus
#include <cstdbool>
#define QTRK_EXTERN_C extern "C"
#else
#include <stdbool.h>
#define QTRK_EXTERN_C
#endif

#define QTRK_API QTRK_EXTERN_C __attribute__((visibility("default")))

typedef unsigned int uint;

/* Localization mode bit flags -- cputrack/qtrk_c_api.h:5-20 */
enum LocalizeModeEnum {
	LT_OnlyCOM = 0,
	LT_XCor1D = 1,
	LT_QI = 2,
	LT_Gaussian2D = 4,
	LT_ZLUTAlign = 8,
	LT_LocalizeZ = 16,
	LT_NormalizeProfile = 64,
	LT_ClearFirstFourPixels = 128,
	LT_FourierLUT = 256,
	LT_LocalizeZWeighted = 512,
	LT_Force32Bit = 0x7fffffff
};
typedef int LocMode_t;

/* Pixel formats -- cputrack/qtrk_c_api.h:24-29 */
enum QTRK_PixelDataType { QTrkU8 = 0, QTrkU16 = 1, QTrkFloat = 2 };
#ifndef __cplusplus
typedef enum QTRK_PixelDataType QTRK_PixelDataType;
#endif

/* BuildLUT flags -- cputrack/QueuedTracker.h:62-65 */
#define BUILDLUT_IMAGELUT 1
#define BUILDLUT_FOURIER 2
#define BUILDLUT_NORMALIZE 4
#define BUILDLUT_BIASCORRECT 8

/* cuda_device codes kept for struct compatibility -- cputrack/qtrk_c_api.h:85-87 */
#define QTrkCUDA_UseList -3
#define QTrkCUDA_UseAll -2
#define QTrkCUDA_UseBest -1

/* vector2f / vector3f: plain float pairs/triples, 4-byte aligned as under the
 * reference's pack(4) (cputrack/std_incl.h:20-116).  In C++ they carry the
 * reference's template names, vector2<float> and vector3<float>, so that C++
 * member functions taking them link under the names that callers built against
 * the reference's headers use. */
#ifdef __cplusplus
template <typename T> struct vector2 { T x, y; };
template <typename T> struct vector3 { T x, y, z; };
typedef vector2<float> vector2f;
typedef vector3<float> vector3f;
#else
typedef struct vector2f { float x, y; } vector2f;
typedef struct vector3f { float x, y, z; } vector3f;
#endif

#pragma pack(push, 1)

/* 24 bytes -- cputrack/qtrk_c_api.h:35-45 */
typedef struct LocalizationJob {
	uint frame, timestamp;
	int zlutIndex;          /* bead index: which radial LUT to use */
	vector3f initialPos;
} LocalizationJob;

/* 52 bytes -- cputrack/qtrk_c_api.h:50-57 */
typedef struct LocalizationResult {
	LocalizationJob job;
	vector3f pos;
	vector2f firstGuess;    /* COM position */
	uint error;             /* 1 when the sampling circle crossed the ROI edge */
	float imageMean;
} LocalizationResult;

/* Field list of QTrkSettings (84 bytes) -- cputrack/qtrk_c_api.h:82-113.
 * Written once as a macro so that QTrkComputedConfig can repeat it in C. */
#define QTRK_SETTINGS_FIELDS \
	int width, height; \
	int numThreads; \
	int cuda_device; \
	float com_bgcorrection; \
	float zlut_minradius; \
	float zlut_radial_coverage; \
	float zlut_angular_coverage; \
	float zlut_roi_coverage; \
	int qi_iterations; \
	float qi_minradius; \
	float qi_radial_coverage; \
	float qi_angular_coverage; \
	float qi_roi_coverage; \
	float qi_angstep_factor; \
	int xc1_profileLength; \
	int xc1_profileWidth; \
	int xc1_iterations; \
	int gauss2D_iterations; \
	float gauss2D_sigma; \
	int downsample;

typedef struct QTrkSettings { QTRK_SETTINGS_FIELDS } QTrkSettings;

/* 8 bytes -- cputrack/qtrk_c_api.h:116-119; top-left corner of an ROI */
typedef struct ROIPosition { int x, y; } ROIPosition;

/* 108 bytes -- cputrack/qtrk_c_api.h:123-138: settings + derived sizes */
#ifdef __cplusplus
struct QTrkComputedConfig : public QTrkSettings {
#else
typedef struct QTrkComputedConfig {
	QTRK_SETTINGS_FIELDS
#endif
	int zlut_radialsteps;
	int zlut_angularsteps;
	float zlut_maxradius;
	int qi_radialsteps;
	int qi_angstepspq;
	float qi_maxradius;
#ifdef __cplusplus
};
#else
} QTrkComputedConfig;
#endif

/* ResultManager structs -- cputrack/qtrk_c_api.h:210-231 */
typedef struct ResultManagerConfig {
	int numBeads, numFrameInfoColumns;
	vector3f scaling;
	vector3f offset;             /* output is (position - ROI centre + offset) * scaling */
	int writeInterval;           /* frames between file writes */
	uint maxFramesInMemory;      /* 0: unlimited */
	uint8_t binaryOutput;
} ResultManagerConfig;

typedef struct RMFrameCounters {
	int startFrame;              /* first frame held in memory */
	int processedFrames;         /* frames below this have every bead's result (or were given up) */
	int lastSaveFrame;
	int capturedFrames;          /* frames with frame info stored */
	int localizationsDone;
	int lostFrames;
	int fileError;
} RMFrameCounters;

#pragma pack(pop)

/* ImageData as marshalled by QTrkDotNet/QTrkDLL.cs (pack(4): pointer, w, h)
 * -- cputrack/utils.h:77-138.  In C++ it is the reference's TImageData<float>
 * (same layout, defined in ImageData.h). */
#ifdef __cplusplus
template <typename T> struct TImageData;
typedef TImageData<float> ImageData;
#else
#pragma pack(push, 4)
typedef struct ImageData { float* data; int w, h; } ImageData;
#pragma pack(pop)
#endif

#ifdef __cplusplus
class QueuedTracker;
#else
typedef struct QueuedTracker QueuedTracker;
#endif

/* ---- tracker C ABI: cputrack/qtrk_c_api.h:144-199, cputrack/qtrk_c_api.cpp:12-175 ---- */

QTRK_API void TestDLLCallConv(int a);                                       /* qtrk_c_api.cpp:12 */
QTRK_API void QTrkGetDefaultConfig(QTrkSettings* cfg);                      /* qtrk_c_api.cpp:17 */
QTRK_API void QTrkGetComputedConfig(QTrkSettings* base, QTrkComputedConfig* cfg); /* qtrk_c_api.cpp:22 */

QTRK_API QueuedTracker* QTrkCreateInstance(QTrkSettings* cfg);              /* qtrk_c_api.h:144 */
QTRK_API void QTrkFreeInstance(QueuedTracker* qtrk);                        /* qtrk_c_api.h:145 */
QTRK_API void QTrkSetLocalizationMode(QueuedTracker* qtrk, LocMode_t locType); /* qtrk_c_api.h:148 */

/* Per-bead calibration images [width*height*numbeads]; result = gain*(pixel+offset). */
QTRK_API void QTrkSetPixelCalibrationImages(QueuedTracker* qtrk, float* offset, float* gain); /* qtrk_c_api.h:152 */
QTRK_API void QTrkSetPixelCalibrationFactors(float offsetFactor, float gainFactor);          /* qtrk_c_api.h:153 */

/* Copies the ROI (pitch in bytes) before returning; blocks while the queue is full. */
QTRK_API void QTrkScheduleLocalization(QueuedTracker* qtrk, void* data, int pitch,
                                       QTRK_PixelDataType pdt, const LocalizationJob* jobInfo); /* qtrk_c_api.h:158 */
QTRK_API void QTrkClearResults(QueuedTracker* qtrk);                        /* qtrk_c_api.h:159 */
QTRK_API void QTrkFlush(QueuedTracker* qtrk);                               /* qtrk_c_api.h:160 */
/* Returns the number of ROIs queued; ROIs outside the frame are skipped. */
QTRK_API int QTrkScheduleFrame(QueuedTracker* qtrk, void* imgptr, int pitch, int width, int height,
                               ROIPosition* positions, int numROI, QTRK_PixelDataType pdt,
                               const LocalizationJob* jobInfo);             /* qtrk_c_api.h:163 */

QTRK_API void QTrkEnableRadialZLUTCompareProfile(bool enabled);             /* qtrk_c_api.h:165 (declared, undefined in the reference) */
QTRK_API void QTrkGetRadialZLUTCompareProfile(float* dst);                  /* qtrk_c_api.h:166 (declared, undefined in the reference) */

/* data may be NULL: allocates a zeroed LUT of [count][planes][zlut_radialsteps]. */
QTRK_API void QTrkSetRadialZLUT(QueuedTracker* qtrk, float* data, int count, int planes); /* qtrk_c_api.h:169 */
QTRK_API void QTrkGetRadialZLUT(QueuedTracker* qtrk, float* dst);           /* qtrk_c_api.h:170 */
QTRK_API void QTrkGetRadialZLUTSize(QueuedTracker* qtrk, int* count, int* planes, int* radialsteps); /* qtrk_c_api.h:171 */
QTRK_API void QTrkSetRadialWeights(QueuedTracker* qtrk, float* zcmp);       /* qtrk_c_api.h:174 */

QTRK_API void QTrkBeginLUT(QueuedTracker* qtrk, uint flags);                /* qtrk_c_api.h:178 */
QTRK_API void QTrkBuildLUT(QueuedTracker* qtrk, void* data, int pitch, QTRK_PixelDataType pdt,
                           int plane, vector2f* known_pos);                 /* qtrk_c_api.h:179 */
QTRK_API void QTrkBuildLUTFromFrame(QueuedTracker* qtrk, ImageData* frame, int plane,
                                    ROIPosition* roipos, int numroi);       /* qtrk_c_api.h:184 (declared, undefined in the reference) */
QTRK_API void QTrkFinalizeLUT(QueuedTracker* qtrk);                         /* qtrk_c_api.h:185 */

QTRK_API int QTrkGetResultCount(QueuedTracker* qtrk);                       /* qtrk_c_api.h:187 */
QTRK_API int QTrkFetchResults(QueuedTracker* qtrk, LocalizationResult* results, int maxResults); /* qtrk_c_api.h:188 */
QTRK_API int QTrkGetQueueLength(QueuedTracker* qtrk, int* maxQueueLen);     /* qtrk_c_api.h:190 */
QTRK_API bool QTrkIsIdle(QueuedTracker* qtrk);                              /* qtrk_c_api.h:191 */
QTRK_API void QTrkGetProfileReport(QueuedTracker* qtrk, char* dst, int maxStrLen); /* qtrk_c_api.h:198 */
QTRK_API void QTrkGetWarnings(QueuedTracker* qtrk, char* dst, int maxStrLen);      /* qtrk_c_api.h:199 */

/* Bead finding (cputrack/qtrk_c_api.cpp:33-58, cputrack/BeadFinder.cpp): the
 * frame is cross-correlated on the GPU with the roi x roi sample at
 * (smpCornerPosX, smpCornerPosY), which must lie inside the frame; the result
 * is a new[] array of top-left ROI corners for QTrkFreeROIPositions, equal to
 * the reference's positions in the same order.  Frames up to 8192 pixels a
 * side after padding to powers of two, roi up to 200.  On an error it returns
 * NULL with *beadcount = 0 and QTrkGetLastError() set. */
QTRK_API void QTrkFreeROIPositions(ROIPosition* data);                      /* qtrk_c_api.h:203 */
QTRK_API ROIPosition* QTrkFindBeads(ImageData* imgData, int smpCornerPosX, int smpCornerPosY, int roi,
                                    float imgRelDist, float acceptance, int* beadcount,
                                    ImageData* sample);                     /* qtrk_c_api.h:205 */

/* ---- result sink -- cputrack/qtrk_c_api.h:233-242, cputrack/qtrk_c_api.cpp:181-241 ---- */
#ifdef __cplusplus
class ResultManager;
#else
typedef struct ResultManager ResultManager;
#endif
QTRK_API ResultManager* RMCreate(const char* file, const char* frameinfo, ResultManagerConfig* cfg,
                                 const char* semicolonSeparatedNames);     /* qtrk_c_api.h:233 */
QTRK_API void RMSetTracker(ResultManager* rm, QueuedTracker* qtrk);        /* qtrk_c_api.h:234 */
QTRK_API void RMDestroy(ResultManager* rm);                                 /* qtrk_c_api.h:235 */
QTRK_API void RMStoreFrameInfo(ResultManager* rm, int frame, double timestamp, float* cols); /* qtrk_c_api.h:236 */
QTRK_API int RMGetBeadResults(ResultManager* rm, int start, int numFrames, int bead,
                              LocalizationResult* results);                /* qtrk_c_api.h:237 */
QTRK_API void RMGetFrameCounters(ResultManager* rm, RMFrameCounters* dst);  /* qtrk_c_api.h:238 */
QTRK_API void RMFlush(ResultManager* rm);                                   /* qtrk_c_api.h:239 */
QTRK_API int RMGetResults(ResultManager* rm, int startFrame, int numFrames,
                          LocalizationResult* results);                    /* qtrk_c_api.h:240 */
QTRK_API void RMRemoveBead(ResultManager* rm, int bead);                    /* qtrk_c_api.h:241 */
QTRK_API void RMGetConfig(ResultManager* rm, ResultManagerConfig* cfg);     /* qtrk_c_api.h:242 */

/* ---- utility exports bound by C# -- cputrack/utils.h:182-219, QTrkDotNet/QTrkDLL.cs:350-367 ---- */
QTRK_API void ComputeRadialProfile(float* dst, int radialSteps, int angularSteps, float minradius,
                                   float maxradius, vector2f center, ImageData* src, float mean,
                                   bool normalize);                         /* utils.h:182 */
QTRK_API void NormalizeRadialProfile(float* prof, int rsteps);              /* utils.h:183 */
QTRK_API void NormalizeImage(ImageData* image);                             /* utils.h:185 */
QTRK_API void GenerateImageFromLUT(ImageData* image, ImageData* zlut, float minradius, float maxradius,
                                   vector3f pos, bool useSplineInterp, int ovs); /* utils.h:186 */
QTRK_API void ApplyPoissonNoise(ImageData* img, float poissonMax, float maxValue); /* utils.h:187 (by reference in C++; a pointer at the ABI) */
QTRK_API void ApplyGaussianNoise(ImageData* img, float sigma);              /* utils.h:188 (same) */
QTRK_API void GenerateGaussianSpotImage(ImageData* img, vector2f pos, float sigma, float I0, float Ibg); /* utils.h:216 */
QTRK_API double GetPreciseTime(void);                                       /* utils.h:219 */

#endif /* QTRK_C_API_H */
