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

/*
 * Grouped radial Z-LUT match.  Holds `beads` LUTs of planes x radialsteps
 * (the reference's [bead][plane][r] layout, cputrack/cpu_tracker.h:37) in GPU
 * memory and matches, in one launch, every bead's profiles against its own
 * LUT.  Each profile gets the best plane and z of the reference's
 * LUTProfileCompare + 7-point FP64 fit (cputrack/cpu_tracker.cpp:791-868),
 * which replaces the per-profile compare inside the tracker for callers that
 * hold many profiles per bead.
 *
 * precision: QTRK_MATCH_FP32 scores every plane in the reference arithmetic
 * (float difference, double sum).  QTRK_MATCH_BF16 / QTRK_MATCH_FP16 compute
 * approximate scores for all planes as a BF16/FP16 matrix-core GEMM with FP32
 * accumulation, rescore the 8 best planes (and the fit window around the best
 * of them) in the reference arithmetic and fit in FP64; the result equals the
 * FP32 path's whenever the exact best plane is among those 8.  The first run in
 * a low precision after QTrkLUTMatchSetLUT converts the LUT (once).
 */
#define QTRK_MATCH_FP32 0
#define QTRK_MATCH_BF16 1
#define QTRK_MATCH_FP16 2

typedef struct QTrkLUTMatch QTrkLUTMatch;

/* On the current HIP device.  NULL on error (QTrkGetLastError). */
QTRK_API QTrkLUTMatch* QTrkLUTMatchCreate(int beads, int planes, int radialsteps);
QTRK_API void QTrkLUTMatchFree(QTrkLUTMatch* m);
/* Copies the LUTs ([beads][planes][radialsteps], host memory) and optional
 * radial weights ([radialsteps] or NULL), as SetRadialZLUT/SetRadialWeights do.  0 on success. */
QTRK_API int QTrkLUTMatchSetLUT(QTrkLUTMatch* m, const float* luts, const float* radialweights);
/* Copies framesPerBead profiles per bead ([beads][framesPerBead][radialsteps],
 * host memory) to the GPU.  0 on success. */
QTRK_API int QTrkLUTMatchUploadProfiles(QTrkLUTMatch* m, const float* profiles, int framesPerBead);
/* Device buffer for [beads][framesPerBead][radialsteps] profiles, for callers
 * that fill it on the GPU; NULL on error. */
QTRK_API float* QTrkLUTMatchDeviceProfiles(QTrkLUTMatch* m, int framesPerBead);
/* Launches `repeats` matches of the current profiles (asynchronous, on the
 * object's own stream; HIP events bracket them).  0 on success. */
QTRK_API int QTrkLUTMatchRun(QTrkLUTMatch* m, int precision, int repeats);
/* Waits; returns the event-timed milliseconds of the last run, negative on error. */
QTRK_API float QTrkLUTMatchSync(QTrkLUTMatch* m);
/* Copies n = beads * framesPerBead results back (profile order): z in planes,
 * best plane, and for the BF16/FP16 paths the 8 candidate planes per profile
 * ([n][8], -1 where none; left untouched by the FP32 path).  Any output may be NULL.  0 on success. */
QTRK_API int QTrkLUTMatchDownload(QTrkLUTMatch* m, float* z, int* bestPlane, int* candidates, int n);

/*
 * Differentiable bead-image model rendered from radial Z-LUTs, its adjoint,
 * and a Poisson maximum-likelihood fit with per-ROI Fisher matrix and CRLB.
 *
 * mu(px) = bg + I * g(px; x, y, z), g = GenerateImageFromLUT (cputrack/utils.cpp:325-390:
 * cubic B-spline or linear interpolation between planes, linear in r, ovs x ovs
 * sub-samples, pixel centres at integer coordinates).  theta = (x, y, z, I, bg)
 * per ROI, x/y in pixels of the width x height ROI, z in LUT planes.
 * The reference evaluates this model on the CPU only, as a test image
 * generator and inside SampleFisherMatrix (cputrack/FisherMatrix.h:8-137),
 * whose 3x3 (x, y, z) Fisher matrix by central differences this model's
 * analytic 5x5 Fisher matrix extends.
 *
 * All array arguments are GPU memory; `stream` is a hipStream_t (NULL: the
 * default stream).  Calls are asynchronous on that stream.  0 on success.
 */
typedef struct QTrkLUTModel QTrkLUTModel;

/* luts: host memory, [count][planes][steps]; bins at minradius + i*(maxradius-minradius)/steps. */
QTRK_API QTrkLUTModel* QTrkLUTModelCreate(const float* luts, int count, int planes, int steps, float minradius,
                                          float maxradius, int width, int height, int ovs, bool splineInterp);
QTRK_API void QTrkLUTModelFree(QTrkLUTModel* m);
/* Forward: mu [n][height][width] and/or the Jacobian d mu / d theta [n][5][height][width]
 * (either may be NULL).  lutIndex [n] selects each ROI's LUT (NULL: LUT 0). */
QTRK_API int QTrkLUTModelForward(QTrkLUTModel* m, const float* theta, const int* lutIndex, int n, float* mu,
                                 float* jacobian, void* stream);
/* Backward (adjoint): gradTheta[i][k] = sum_px gradMu[i][px] * d mu[i][px] / d theta_k. */
QTRK_API int QTrkLUTModelBackward(QTrkLUTModel* m, const float* theta, const int* lutIndex, int n,
                                  const float* gradMu, float* gradTheta, void* stream);
/* Fits theta to photon counts data [n][height][width] from thetaInit with `iterations`
 * Fisher-scoring Newton steps (FP64 5x5 solves); writes theta [n][5] (may alias thetaInit)
 * and, at the fitted theta, the Fisher matrix [n][25] (FP64), the CRLB sqrt(diag(F^-1)) [n][5]
 * and the negative log-likelihood sum(mu - data*ln(mu)) [n]; those three may be NULL. */
QTRK_API int QTrkLUTModelFitMLE(QTrkLUTModel* m, const float* data, const float* thetaInit, const int* lutIndex,
                                int n, int iterations, float* theta, double* fisher, float* crlb, float* nll,
                                void* stream);
/* Waits for `stream` (NULL: the default stream). */
QTRK_API int QTrkLUTModelSync(QTrkLUTModel* m, void* stream);

#endif
