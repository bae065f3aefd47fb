/*
 * qtrk_amd.h -- entry points of libqtrk.so that have no counterpart in the
 * reference's C ABI.  They serve callers that keep their ROIs in GPU memory
 * (the device-resident batch) and callers that need the error text of a call
 * that failed (the reference threw C++ exceptions across its C ABI instead).
 */
#ifndef QTRK_AMD_H
#define QTRK_AMD_H

#include "qtrk_c_api.h"

/* Text of the last error raised on this thread, or "" (no reference
 * counterpart: the reference lets std::runtime_error escape its C ABI). */
QTRK_API const char* QTrkGetLastError(void);
QTRK_API void QTrkClearLastError(void);

/* QueuedTracker::SetConfigValue / GetConfigValues (cputrack/QueuedTracker.h:76-78),
 * which the reference's C ABI leaves commented out (cputrack/qtrk_c_api.h:193-196).
 * Keys: "batch_size" (ROIs per GPU launch), "num_streams" (batches in flight).
 * GetConfigValues writes "key=value" lines. */
QTRK_API void QTrkSetConfigValue(QueuedTracker* qtrk, const char* name, const char* value);
QTRK_API void QTrkGetConfigValues(QueuedTracker* qtrk, char* dst, int maxStrLen);

/* Name, compute units and clock of the GPU the library runs on. */
QTRK_API int QTrkGetDeviceInfo(char* dst, int maxStrLen);

/*
 * Device-resident batch.  Holds `maxROIs` ROIs of cfg.width x cfg.height
 * pixels of one pixel type in GPU memory, contiguous ([n][height][width]),
 * plus their jobs and results.  A run localizes them with the tracker's
 * current mode, LUT, radial weights and calibration, on the batch's own HIP
 * stream, with no host copies.  The tracker must outlive the batch.
 *
 * No counterpart in the reference: its queue always starts from host
 * buffers (cputrack/QueuedTracker.h:37).
 */
typedef struct QTrkBatch QTrkBatch;

QTRK_API QTrkBatch* QTrkBatchCreate(QueuedTracker* qtrk, int maxROIs, QTRK_PixelDataType pdt);
QTRK_API void QTrkBatchFree(QTrkBatch* b);
/* Copies n ROIs (contiguous, pitch = width*bytesPerPixel) and their jobs to the GPU. 0 on success. */
QTRK_API int QTrkBatchUpload(QTrkBatch* b, const void* rois, const LocalizationJob* jobs, int n);
/* Device pointers of the batch's ROI, job and result buffers, for callers that fill them on the GPU. */
QTRK_API void* QTrkBatchDeviceROIs(QTrkBatch* b);
QTRK_API void* QTrkBatchDeviceJobs(QTrkBatch* b);
QTRK_API void* QTrkBatchDeviceResults(QTrkBatch* b);
/* Launches `repeats` back-to-back localizations of the first n ROIs
 * (asynchronous).  HIP events on the batch stream bracket all of them. */
QTRK_API int QTrkBatchRun(QTrkBatch* b, int n, int repeats);
/* Waits for the batch stream; returns the event-timed milliseconds of the
 * last QTrkBatchRun (all repeats), or a negative value on error. */
QTRK_API float QTrkBatchSync(QTrkBatch* b);
/* Copies n results back to the host (synchronous). 0 on success. */
QTRK_API int QTrkBatchDownload(QTrkBatch* b, LocalizationResult* out, int n);
/* Diagnostics: when enabled, each run also records every ROI's best LUT plane
 * (-1 when no LUT compare ran) and its radial profile [n][zlut_radialsteps];
 * either output pointer may be NULL.  0 on success. */
QTRK_API int QTrkBatchEnableDiagnostics(QTrkBatch* b, bool enable);
QTRK_API int QTrkBatchDownloadDiagnostics(QTrkBatch* b, int* bestPlane, float* profiles, int n);

/*
 * Renders n bead images of width x height from radial LUTs, on the host
 * with all cores: image i uses LUT lutIndex[i] (luts is [nluts][planes][steps])
 * at pos[i] = (x, y, z) with GenerateImageFromLUT's model (cputrack/utils.cpp:325-390).
 * out is [n][height][width].  A batch form of the reference's
 * GenerateImageFromLUT export, used to synthesise benchmark workloads.
 */
QTRK_API void QTrkGenerateImagesFromLUT(float* out, int n, int width, int height, const float* luts, int nluts,
                                        int planes, int steps, float minradius, float maxradius,
                                        const float* pos, const int* lutIndex, bool useSplineInterp, int ovs);

#endif
