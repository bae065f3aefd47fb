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
 * Keys: "batch_size" (ROIs per GPU launch), "num_streams" (batches in flight),
 * "frame_copy_threads" (threads, the caller's included, that copy a
 * QTrkScheduleFrame frame into pinned memory; default 5), "queue_spin_us" (how long a
 * schedule call blocked on a full queue polls the oldest batch before it sleeps in a
 * blocking wait; default 2000).
 * GetConfigValues writes "key=value" lines. */
QTRK_API void QTrkSetConfigValue(QueuedTracker* qtrk, const char* name, const char* value);
QTRK_API void QTrkGetConfigValues(QueuedTracker* qtrk, char* dst, int maxStrLen);

/* QueuedTracker::SetImageZLUT / GetImageZLUT / GetImageZLUTSize
 * (cputrack/QueuedTracker.h:58-60), which the reference's C ABI does not export and
 * LabVIEW calls through the C++ interface (cputrack/lv_queuetrk_api.cpp:159-175).
 * dims = [beads, planes, height, width]; src null gives a zeroed image LUT; the radial
 * LUT is set alongside from radial_zlut ([beads][planes][zlut_radialsteps], null:
 * zeroed).  BuildLUT with BUILDLUT_IMAGELUT adds into plane `plane` of it.  Set
 * returns 1 (0 on an error). */
QTRK_API int QTrkSetImageZLUT(QueuedTracker* qtrk, float* src, float* radial_zlut, int* dims);
QTRK_API void QTrkGetImageZLUT(QueuedTracker* qtrk, float* dst);
QTRK_API void QTrkGetImageZLUTSize(QueuedTracker* qtrk, int* dims);
/* The image LUT's first and second differences across planes that QTrkFinalizeLUT computes, as the
 * reference's FinalizeLUT does into its image_lut_dz / image_lut_dz2 (cputrack/QueuedCPUTracker.cpp:
 * 661-696), which it exposes through no interface: dz = (U - L) / 2 and dz2 = U - 2 C + L of the
 * planes below, at and above, the first and last planes copied from their neighbours; each
 * [beads][planes][height][width] (either may be NULL).  1 when present, 0 when not (no image LUT
 * of at least 3 planes at QTrkFinalizeLUT, or set again since), -1 on error. */
QTRK_API int QTrkGetImageZLUTDerivatives(QueuedTracker* qtrk, float* dz, float* dz2);

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
/* Launches `repeats` localizations of the first n ROIs (asynchronous), as a
 * stream of consecutive batches: the repeats alternate between the batch
 * stream and a second one, so a launch's first workgroups fill the CUs the
 * previous launch's last ones leave idle; all write the same (identical)
 * results.  HIP events on the batch stream bracket all of them. */
QTRK_API int QTrkBatchRun(QTrkBatch* b, int n, int repeats);
/* How a run matches the radial profiles against the Z-LUT (modes with LT_LocalizeZ):
 * QTRK_MATCH_FP32 (the default) scores each ROI's profile inside its localization
 * workgroup, in the reference arithmetic.  QTRK_MATCH_BF16 / QTRK_MATCH_FP16 run the
 * localization in profile-only form, group the ROIs by bead (job.zlutIndex) on the GPU and
 * match each bead's profiles together as QTrkLUTMatch does (matrix-core scores, the 8 best
 * planes and the fit window rescored in the reference arithmetic, FP64 fit): results equal the
 * FP32 form's whenever the exact best plane is among those 8.  QTRK_MATCH_FP32_GROUPED (3)
 * groups the same way and scores every plane in the reference arithmetic (QTrkLUTMatch's
 * QTRK_MATCH_FP32): results identical to QTRK_MATCH_FP32's, each bead's LUT rows read once
 * per 16 of its ROIs instead of once per ROI.  Not with LT_ZLUTAlign,
 * LT_LocalizeZWeighted or LT_FourierLUT, and the compare profile
 * (QTrkEnableRadialZLUTCompareProfile) is not recorded.  The first call converts the
 * tracker's LUT (again after the LUT or its weights change).  0 on success. */
QTRK_API int QTrkBatchSetMatchPrecision(QTrkBatch* b, int precision);
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
 * No counterpart in the reference.  For callers that cannot allocate GPU memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * QTrkBatchRun's localization (one run, the match precision of QTrkBatchSetMatchPrecision
 * given here) of n ROIs on the caller's GPU buffers -- rois [n][height][width] of pdt,
 * contiguous; jobs [n]; results [n] -- asynchronously on `stream` (a hipStream_t; NULL: the
 * default stream).  It allocates nothing: QTRK_MATCH_FP32 needs no scratch, and the
 * BF16/FP16 grouped match keeps its scratch (radial profiles, grouping by bead) in a
 * fixed device array of QTRK_NOALLOC_SCRATCH_MB MiB that each GPU the library runs on
 * holds until the process ends.  A call needing more -- 4*n*zlut_radialsteps + 4*n +
 * 8*beads + 4 bytes, each part rounded up to 256 -- fails with an error, as does a
 * BF16/FP16 call before QTrkBatchSetMatchPrecision has converted the tracker's current
 * LUT for that precision.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams or threads) race.  Results are as QTrkBatchRun leaves them in
 * GPU memory, without the Z-bias correction QTrkBatchDownload applies.  0 on success.
 */
#define QTRK_NOALLOC_SCRATCH_MB 32
QTRK_API int QTrkLocalizeDeviceNoAlloc(QueuedTracker* qtrk, const void* rois, QTRK_PixelDataType pdt,
                                       const LocalizationJob* jobs, LocalizationResult* results, int n,
                                       int precision, void* stream);

/*
 * Z-bias table: per bead, the z error of the tracker's own localization of
 * its LUT rendered at biasPlanes depths, inverted on every later result
 * (QueuedTracker::ZLUTBiasCorrection, cputrack/QueuedTracker.cpp:115-140).
 *
 * QTrkComputeZLUTBiasTable replaces QueuedTracker::ComputeZBiasCorrection
 * (cputrack/QueuedTracker.cpp:143-189), bound in the reference by the LabVIEW
 * export qtrk_compute_zlut_bias_table (cputrack/lv_queuetrk_api.cpp:457-466):
 * the renders (GenerateImageFromLUT, smpPerPixel^2 samples per pixel, B-spline
 * or linear in z, at the ROI centre and z = plane/biasPlanes*planes) and their
 * localization (COM, 2 QI iterations, the LUT compare weighted by the Z-LUT bin
 * window) run on the GPU.  The table becomes active; result, when not NULL,
 * receives it as [numbeads][biasPlanes].  0 on success.
 *
 * QTrkSetZLUTBiasTable replaces SetZLUTBiasCorrection (:192-197) /
 * qtrk_set_zlut_bias_table (cputrack/lv_queuetrk_api.cpp:468-480): table is
 * [count][biasPlanes] with count equal to the Z-LUT's bead count; NULL clears
 * it.  QTrkSetRadialZLUT clears it too (cputrack/QueuedCPUTracker.cpp:368-369).
 *
 * QTrkGetZLUTBiasTable replaces GetZLUTBiasCorrection (:199-203): writes the
 * sizes (and the table when dst is not NULL); 1 when a table is set, 0 when
 * not, -1 on error.
 */
QTRK_API int QTrkComputeZLUTBiasTable(QueuedTracker* qtrk, int biasPlanes, float* result, int smpPerPixel,
                                      bool useSplineInterp);
QTRK_API int QTrkSetZLUTBiasTable(QueuedTracker* qtrk, const float* table, int biasPlanes, int count);
QTRK_API int QTrkGetZLUTBiasTable(QueuedTracker* qtrk, float* dst, int* biasPlanes, int* count);

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
/* device batches only (QTrkBatchSetMatchPrecision, QTrkLocalizeDeviceNoAlloc) */
#define QTRK_MATCH_FP32_GROUPED 3

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
 * Fisher-scoring Newton steps (FP64 5x5 solves); a start with I <= 0 first gets I and bg by
 * linear least squares at its (x, y, z).  Writes theta [n][5] (may alias thetaInit)
 * and, at the fitted theta, the Fisher matrix [n][25] (FP64), the CRLB sqrt(diag(F^-1)) [n][5]
 * and the negative log-likelihood sum(mu - data*ln(mu)) [n]; those three may be NULL.
 * The fit takes models of ovs <= 8 (at most 8x8 sub-samples per pixel); larger ones
 * fail with an error, as Forward and Backward still take them. */
QTRK_API int QTrkLUTModelFitMLE(QTrkLUTModel* m, const float* data, const float* thetaInit, const int* lutIndex,
                                int n, int iterations, float* theta, double* fisher, float* crlb, float* nll,
                                void* stream);
/* Lanes per ROI of QTrkLUTModelFitMLE: 64, 128, 256 (the default), 512 for models of at most 6x6
 * sub-samples per pixel, 1024 for at most 4x4.  More lanes shorten the fit of few large ROIs (one workgroup
 * per ROI; e.g. 32 ROIs of 128x128 px fill 32 of the GPU's 256 CUs); the per-lane partial sums, and
 * so the last bits of the result, depend on the lane count. */
QTRK_API int QTrkLUTModelSetFitLanes(QTrkLUTModel* m, int lanes);
/* Waits for `stream` (NULL: the default stream). */
QTRK_API int QTrkLUTModelSync(QTrkLUTModel* m, void* stream);

/*
 * Multi-tracker pipeline: the full chain of several trackers (each with its
 * own ROI size, settings, LUT and pixel type) served by one hipGraph launch.
 * Per tracker: COM -> QI -> radial profile (the tracker's kernel and settings)
 * -> grouped Z-LUT match of the batch's profiles against the tracker's LUT
 * (matchPrecision as in QTrkLUTMatch) -> optional LUT-model MLE on the ROI
 * pixels (mleIterations Fisher-scoring steps from the QI x/y and matched z,
 * image model from modelLUTs as in QTrkLUTModel, 4x4 sub-samples, B-spline).
 * The chains run as parallel branches of the graph.
 *
 * Order of use: create; add trackers (each with its LUT already set); set
 * every tracker's jobs (the fixed ROI layout of a batch: frame/bead per ROI;
 * the grouping by bead is derived from it); upload representative ROIs;
 * finalize (autotune = true times the launch-shape candidates of each stage on
 * those ROIs and keeps the fastest; then the graph is captured); then per
 * batch: fill the ROI buffers (upload or QTrkPipelineDeviceROIs), launch,
 * sync, download.  Trackers must not change their settings after finalize; a
 * tracker's Z-LUT or radial weights changed after it (SetRadialZLUT, SetRadialWeights,
 * BuildLUT, FinalizeLUT) makes QTrkPipelineLaunch fail until the pipeline is finalized
 * again, which converts the new LUT.  No counterpart in the reference (one tracker, one queue).
 */
typedef struct QTrkPipeline QTrkPipeline;

QTRK_API QTrkPipeline* QTrkPipelineCreate(void);
QTRK_API void QTrkPipelineFree(QTrkPipeline* p);
/* Returns the tracker's slot (>= 0), or -1.  modelLUTs: host [modelCount][modelPlanes][modelSteps]
 * (one LUT per bead, or one for all), or NULL for no MLE stage. */
QTRK_API int QTrkPipelineAddTracker(QTrkPipeline* p, QueuedTracker* qtrk, int roisPerBatch, QTRK_PixelDataType pdt,
                                    int matchPrecision, const float* modelLUTs, int modelCount, int modelPlanes,
                                    int modelSteps, float modelMinRadius, float modelMaxRadius, int mleIterations);
QTRK_API int QTrkPipelineSetJobs(QTrkPipeline* p, int slot, const LocalizationJob* jobs, int n);
QTRK_API void* QTrkPipelineDeviceROIs(QTrkPipeline* p, int slot);
QTRK_API int QTrkPipelineUploadROIs(QTrkPipeline* p, int slot, const void* rois, int n);
QTRK_API int QTrkPipelineFinalize(QTrkPipeline* p, bool autotune);
/* One graph launch over every tracker's batch (asynchronous). */
QTRK_API int QTrkPipelineLaunch(QTrkPipeline* p);
/* Waits; event-timed milliseconds of the last launch, negative on error. */
QTRK_API float QTrkPipelineSync(QTrkPipeline* p);
/* results: the localization records with z from the grouped match; theta [n][5] and crlb [n][5]
 * from the MLE stage.  Any output may be NULL. */
QTRK_API int QTrkPipelineDownload(QTrkPipeline* p, int slot, LocalizationResult* results, float* theta, float* crlb,
                                  int n);
/* Camera-paced run: `batches` graph launches, batch i submitted at i * periodMs after the first;
 * latencyMs[i] = host time from its submission until its completion is observed, gpuMs[i]
 * (may be NULL) = its event-timed GPU time. */
QTRK_API int QTrkPipelineStream(QTrkPipeline* p, int batches, float periodMs, float* latencyMs, float* gpuMs);
/* Per-tracker sizes and the launch shapes chosen (with the candidates' times). */
QTRK_API int QTrkPipelineGetReport(QTrkPipeline* p, char* dst, int maxLen);

#endif
