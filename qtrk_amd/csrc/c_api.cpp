// c_api.cpp -- extern "C" boundary of libqtrk.so (include/qtrk_c_api.h,
// include/qtrk_amd.h).  Every entry point forwards to the tracker inside a
// guard that turns exceptions into the thread's last error.
#include "ImageData.h"
#include "QueuedTracker.h"
#include "beadfinder.hpp"
#include "errors.hpp"
#include "hip_tracker.hpp"
#include "qtrk_amd.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

namespace qtrk {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// Trackers alive in this process, for the two reference entry points that
// take no tracker argument (cputrack/qtrk_c_api.h:153, 165-166).
static std::mutex g_reg_mtx;
static std::set<QueuedTracker*> g_trackers;
static QueuedTracker* g_last_tracker = nullptr;
}  // namespace qtrk

using namespace qtrk;

QueuedTracker* CreateQueuedTracker(const QTrkComputedConfig& cc)
{
	QueuedTracker* t = new QueuedHIPTracker(cc);
	std::lock_guard<std::mutex> g(g_reg_mtx);
	g_trackers.insert(t);
	g_last_tracker = t;
	return t;
}

static void unregister(QueuedTracker* t)
{
	std::lock_guard<std::mutex> g(g_reg_mtx);
	g_trackers.erase(t);
	if (g_last_tracker == t) g_last_tracker = g_trackers.empty() ? nullptr : *g_trackers.rbegin();
}

QTRK_API const char* QTrkGetLastError(void) { return g_last_error.c_str(); }
QTRK_API void QTrkClearLastError(void) { g_last_error.clear(); }

QTRK_API int QTrkGetDeviceInfo(char* dst, int maxStrLen)
{
	int ok = 0;
	GUARD({
		int d = 0;
		HIP_CHECK(hipGetDevice(&d));
		hipDeviceProp_t p;
		HIP_CHECK(hipGetDeviceProperties(&p, d));
		std::string s = std::string(p.name) + " (" + p.gcnArchName + "), " + std::to_string(p.multiProcessorCount) +
		                " CUs, " + std::to_string(p.clockRate / 1000) + " MHz, LDS/block " +
		                std::to_string(p.sharedMemPerBlock);
		if (dst && maxStrLen > 0) {
			strncpy(dst, s.c_str(), maxStrLen - 1);
			dst[maxStrLen - 1] = 0;
		}
		ok = 1;
	});
	return ok ? 0 : -1;
}

QTRK_API void TestDLLCallConv(int a) {}

QTRK_API void QTrkGetDefaultConfig(QTrkSettings* cfg)
{
	QTrkSettings s;
	memset(&s, 0, sizeof(s));
	// cputrack/qtrk_c_api.h:60-81
	s.width = s.height = 150;
	s.numThreads = -1;
	s.xc1_profileLength = 128;
	s.xc1_profileWidth = 32;
	s.xc1_iterations = 2;
	s.zlut_minradius = 1.0f;
	s.zlut_angular_coverage = 0.7f;
	s.zlut_radial_coverage = 3.0f;
	s.zlut_roi_coverage = 1.0f;
	s.qi_iterations = 4;
	s.qi_minradius = 1;
	s.qi_angular_coverage = 0.7f;
	s.qi_radial_coverage = 3.0f;
	s.qi_roi_coverage = 1.0f;
	s.qi_angstep_factor = 1.0f;
	s.cuda_device = -2;
	s.com_bgcorrection = 0.0f;
	s.gauss2D_iterations = 6;
	s.gauss2D_sigma = 4;
	s.downsample = 0;
	*cfg = s;
}

QTRK_API void QTrkGetComputedConfig(QTrkSettings* base, QTrkComputedConfig* cfg) { *cfg = QTrkComputeConfig(*base); }

QTRK_API QueuedTracker* QTrkCreateInstance(QTrkSettings* cfg)
{
	QueuedTracker* t = nullptr;
	GUARD(t = CreateQueuedTracker(QTrkComputeConfig(*cfg)));
	return t;
}

QTRK_API void QTrkFreeInstance(QueuedTracker* qtrk)
{
	if (!qtrk) return;
	unregister(qtrk);
	GUARD(delete qtrk);
}

QTRK_API void QTrkSetLocalizationMode(QueuedTracker* q, LocMode_t t) { GUARD(q->SetLocalizationMode(t)); }
QTRK_API void QTrkSetPixelCalibrationImages(QueuedTracker* q, float* offset, float* gain)
{
	GUARD(q->SetPixelCalibrationImages(offset, gain));
}
QTRK_API void QTrkSetPixelCalibrationFactors(float offsetFactor, float gainFactor)
{
	GUARD({
		std::lock_guard<std::mutex> g(g_reg_mtx);
		for (QueuedTracker* t : g_trackers) t->SetPixelCalibrationFactors(offsetFactor, gainFactor);
	});
}
QTRK_API void QTrkScheduleLocalization(QueuedTracker* q, void* data, int pitch, QTRK_PixelDataType pdt,
                                       const LocalizationJob* job)
{
	GUARD(q->ScheduleLocalization(data, pitch, pdt, job));
}
QTRK_API void QTrkClearResults(QueuedTracker* q) { GUARD(q->ClearResults()); }
QTRK_API void QTrkFlush(QueuedTracker* q) { GUARD(q->Flush()); }
QTRK_API int QTrkScheduleFrame(QueuedTracker* q, void* img, int pitch, int w, int h, ROIPosition* pos, int n,
                               QTRK_PixelDataType pdt, const LocalizationJob* job)
{
	int r = 0;
	GUARD(r = q->ScheduleFrame(img, pitch, w, h, pos, n, pdt, job));
	return r;
}
QTRK_API void QTrkEnableRadialZLUTCompareProfile(bool enabled)
{
	GUARD({
		std::lock_guard<std::mutex> g(g_reg_mtx);
		for (QueuedTracker* t : g_trackers) t->EnableRadialZLUTCompareProfile(enabled);
	});
}
QTRK_API void QTrkGetRadialZLUTCompareProfile(float* dst)
{
	GUARD({
		std::lock_guard<std::mutex> g(g_reg_mtx);
		if (!g_last_tracker) fail("no tracker instance");
		g_last_tracker->GetRadialZLUTCompareProfile(dst);
	});
}
QTRK_API void QTrkSetRadialZLUT(QueuedTracker* q, float* data, int count, int planes)
{
	GUARD(q->SetRadialZLUT(data, count, planes));
}
QTRK_API void QTrkGetRadialZLUT(QueuedTracker* q, float* dst) { GUARD(q->GetRadialZLUT(dst)); }
QTRK_API void QTrkGetRadialZLUTSize(QueuedTracker* q, int* count, int* planes, int* rsteps)
{
	GUARD(q->GetRadialZLUTSize(*count, *planes, *rsteps));
}
QTRK_API void QTrkSetRadialWeights(QueuedTracker* q, float* zcmp) { GUARD(q->SetRadialWeights(zcmp)); }
QTRK_API void QTrkBeginLUT(QueuedTracker* q, uint flags) { GUARD(q->BeginLUT(flags)); }
QTRK_API void QTrkBuildLUT(QueuedTracker* q, void* data, int pitch, QTRK_PixelDataType pdt, int plane,
                           vector2f* known_pos)
{
	GUARD(q->BuildLUT(data, pitch, pdt, plane, known_pos));
}
// Builds plane `plane` from ROIs cut out of one float frame: ROI i feeds bead i.
QTRK_API void QTrkBuildLUTFromFrame(QueuedTracker* q, ImageData* frame, int plane, ROIPosition* roipos, int numroi)
{
	GUARD({
		int count = 0, planes = 0, rs = 0;
		q->GetRadialZLUTSize(count, planes, rs);
		if (numroi != count) fail("QTrkBuildLUTFromFrame: numroi must equal the LUT bead count");
		const int w = q->cfg.width, h = q->cfg.height;
		std::vector<float> stack((size_t)w * h * numroi);
		for (int i = 0; i < numroi; i++) {
			const ROIPosition& p = roipos[i];
			if (p.x < 0 || p.y < 0 || p.x + w > frame->w || p.y + h > frame->h) fail("ROI outside frame");
			for (int y = 0; y < h; y++)
				memcpy(&stack[((size_t)i * h + y) * w], &frame->data[(size_t)(p.y + y) * frame->w + p.x], 4 * (size_t)w);
		}
		q->BuildLUT(stack.data(), 4 * w, QTrkFloat, plane, nullptr);
	});
}
QTRK_API void QTrkFinalizeLUT(QueuedTracker* q) { GUARD(q->FinalizeLUT()); }
QTRK_API int QTrkGetResultCount(QueuedTracker* q)
{
	int r = 0;
	GUARD(r = q->GetResultCount());
	return r;
}
QTRK_API int QTrkFetchResults(QueuedTracker* q, LocalizationResult* results, int maxResults)
{
	int r = 0;
	GUARD(r = q->FetchResults(results, maxResults));
	return r;
}
QTRK_API int QTrkGetQueueLength(QueuedTracker* q, int* maxQueueLen)
{
	int r = 0;
	GUARD(r = q->GetQueueLength(maxQueueLen));
	return r;
}
// QTrkIsIdle itself is defined in c_api_idle.cpp, with an int result
extern "C" __attribute__((visibility("hidden"))) int qtrk_is_idle(void* qtrk)
{
	bool r = false;
	GUARD(r = ((QueuedTracker*)qtrk)->IsIdle());
	return r ? 1 : 0;
}
static void copy_str(const std::string& s, char* dst, int n)
{
	if (!dst || n <= 0) return;
	strncpy(dst, s.c_str(), n);
}
QTRK_API void QTrkGetProfileReport(QueuedTracker* q, char* dst, int maxStrLen)
{
	GUARD(copy_str(q->GetProfileReport(), dst, maxStrLen));
}
QTRK_API void QTrkGetWarnings(QueuedTracker* q, char* dst, int maxStrLen) { GUARD(copy_str(q->GetWarnings(), dst, maxStrLen)); }

QTRK_API int QTrkSetImageZLUT(QueuedTracker* q, float* src, float* radial_zlut, int* dims)
{
	int ok = 0;
	GUARD(ok = q->SetImageZLUT(src, radial_zlut, dims) ? 1 : 0);
	return ok;
}
QTRK_API void QTrkGetImageZLUT(QueuedTracker* q, float* dst) { GUARD(q->GetImageZLUT(dst)); }
QTRK_API void QTrkGetImageZLUTSize(QueuedTracker* q, int* dims) { GUARD(q->GetImageZLUTSize(dims)); }
QTRK_API int QTrkGetImageZLUTDerivatives(QueuedTracker* q, float* dz, float* dz2)
{
	int r = -1;
	GUARD({
		auto* t = dynamic_cast<QueuedHIPTracker*>(q);
		if (!t) fail("not a MI355X tracker");
		r = t->GetImageZLUTDerivatives(dz, dz2) ? 1 : 0;
	});
	return r;
}

QTRK_API void QTrkSetConfigValue(QueuedTracker* q, const char* name, const char* value)
{
	GUARD(q->SetConfigValue(name ? name : "", value ? value : ""));
}
QTRK_API void QTrkGetConfigValues(QueuedTracker* q, char* dst, int maxStrLen)
{
	GUARD({
		std::string s;
		for (auto& kv : q->GetConfigValues()) s += kv.first + "=" + kv.second + "\n";
		copy_str(s, dst, maxStrLen);
	});
}

QTRK_API void QTrkFreeROIPositions(ROIPosition* data) { delete[] data; }
// QTrkFindBeads (cputrack/qtrk_c_api.cpp:33-58): the sample is the roi x roi block of the frame at
// (smpCornerPosX, smpCornerPosY), copied into `sample` when its size is roi x roi; the result is
// a new[] array for QTrkFreeROIPositions.  Returns NULL with *beadcount = 0 on an error.
QTRK_API ROIPosition* QTrkFindBeads(ImageData* imgData, int smpCornerPosX, int smpCornerPosY, int roi,
                                    float imgRelDist, float acceptance, int* beadcount, ImageData* sample)
{
	if (beadcount) *beadcount = 0;
	ROIPosition* out = nullptr;
	GUARD({
		if (!imgData || !beadcount) fail("null argument");
		float* smp = (sample && sample->data && sample->w == roi && sample->h == roi) ? sample->data : nullptr;
		std::vector<ROIPosition> found = find_beads(imgData->data, imgData->w, imgData->h, smpCornerPosX,
		                                            smpCornerPosY, roi, imgRelDist, acceptance, smp);
		out = new ROIPosition[found.size() ? found.size() : 1];
		std::copy(found.begin(), found.end(), out);
		*beadcount = (int)found.size();
	});
	return out;
}

// ------------------------------------------------------------ device batch

struct QTrkBatch {
	QueuedHIPTracker* trk = nullptr;
	int maxROIs = 0, pdt = 0;
	int prec = kMatchFP32;  // QTrkBatchSetMatchPrecision
	// a host copy of the uploaded jobs, whose largest bead count sizes the grouped match's grid
	// (cleared once QTrkBatchDeviceJobs hands out the device buffer: the jobs may then change there)
	std::vector<LocalizationJob> h_jobs;
	DeviceBuffer rois, jobs, results, best, prof, split, split2;
	bool diag = false;
	hipStream_t stream = nullptr, stream2 = nullptr;  // repeats alternate between the two
	hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
};

QTRK_API QTrkBatch* QTrkBatchCreate(QueuedTracker* q, int maxROIs, QTRK_PixelDataType pdt)
{
	QTrkBatch* b = nullptr;
	GUARD({
		auto* t = dynamic_cast<QueuedHIPTracker*>(q);
		if (!t) fail("not a MI355X tracker");
		if (maxROIs <= 0) fail("maxROIs must be > 0");
		HIP_CHECK(hipSetDevice(t->Device()));
		b = new QTrkBatch();
		b->trk = t;
		b->maxROIs = maxROIs;
		b->pdt = (int)pdt;
		const size_t rb = (size_t)t->cfg.width * t->cfg.height * PDT_BytesPerPixel(pdt);
		b->rois.ensure(rb * maxROIs);
		b->jobs.ensure(sizeof(DevJob) * maxROIs);
		b->results.ensure(sizeof(DevResult) * maxROIs);
		HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
		HIP_CHECK(hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking));
		HIP_CHECK(hipEventCreate(&b->e0));
		HIP_CHECK(hipEventCreate(&b->e1));
		HIP_CHECK(hipEventCreateWithFlags(&b->e2, hipEventDisableTiming));
	});
	return b;
}

QTRK_API void QTrkBatchFree(QTrkBatch* b)
{
	if (!b) return;
	hipStreamSynchronize(b->stream);
	hipStreamSynchronize(b->stream2);
	hipEventDestroy(b->e0);
	hipEventDestroy(b->e1);
	hipEventDestroy(b->e2);
	hipStreamDestroy(b->stream);
	hipStreamDestroy(b->stream2);
	delete b;
}

QTRK_API int QTrkBatchUpload(QTrkBatch* b, const void* rois, const LocalizationJob* jobs, int n)
{
	bool ok = GUARD({
		if (n < 0 || n > b->maxROIs) fail("n out of range");
		const size_t rb = (size_t)b->trk->cfg.width * b->trk->cfg.height * PDT_BytesPerPixel((QTRK_PixelDataType)b->pdt);
		HIP_CHECK(hipMemcpy(b->rois.p, rois, rb * n, hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(b->jobs.p, jobs, sizeof(DevJob) * n, hipMemcpyHostToDevice));
		b->h_jobs.assign(jobs, jobs + n);
	});
	return ok ? 0 : -1;
}

QTRK_API void* QTrkBatchDeviceROIs(QTrkBatch* b) { return b->rois.p; }
QTRK_API void* QTrkBatchDeviceJobs(QTrkBatch* b)
{
	b->h_jobs.clear();
	return b->jobs.p;
}
QTRK_API void* QTrkBatchDeviceResults(QTrkBatch* b) { return b->results.p; }

QTRK_API int QTrkBatchRun(QTrkBatch* b, int n, int repeats)
{
	bool ok = GUARD({
		if (n < 0 || n > b->maxROIs) fail("n out of range");
		HIP_CHECK(hipSetDevice(b->trk->Device()));
		b->trk->WaitLaunched();
		LocalizeParams P;
		const int pitch = b->trk->cfg.width * PDT_BytesPerPixel((QTRK_PixelDataType)b->pdt);
		b->trk->FillParams(P, b->pdt, n, b->rois.p, nullptr, pitch, b->jobs.as<DevJob>(), b->results.as<DevResult>());
		if (b->diag) {
			P.best_out = b->best.as<int>();
			P.prof_out = b->prof.as<float>();
		}
		if (b->prec != kMatchFP32) {
			// profile kernel + grouped match; repeats alternate between the two streams as below,
			// each stream with its own scratch (diagnostics: one stream, their buffers)
			const std::string why = b->trk->SplitMatchUsable();
			if (!why.empty()) fail("the grouped Z-LUT match cannot run: " + why);
			const int mp = match_kernel_precision(b->prec);
			const LutMatchTables* tab = b->trk->PrepareMatchTables(mp);
			const int nb = b->trk->cfg.zlut_radialsteps;
			const size_t bytes = QueuedHIPTracker::SplitScratch::Bytes(b->maxROIs, nb, tab->beads);
			const bool two = repeats > 1 && !b->diag;
			b->split.ensure(bytes);
			if (two) b->split2.ensure(bytes);
			QueuedHIPTracker::SplitScratch sc[2] = {
				QueuedHIPTracker::SplitScratch::Carve(b->split.p, b->maxROIs, nb, tab->beads),
				QueuedHIPTracker::SplitScratch::Carve(two ? b->split2.p : b->split.p, b->maxROIs, nb, tab->beads) };
			if (b->diag) sc[0].prof = b->prof.as<float>();
			// the most ROIs of one bead, from the uploaded jobs (0 when no ROI has a bead in the
			// LUT); unknown for jobs written on the GPU
			const int mpb = n <= (int)b->h_jobs.size()
			                    ? QueuedHIPTracker::MaxPerBead((const DevJob*)b->h_jobs.data(), n, tab->beads)
			                    : -1;
			HIP_CHECK(hipEventRecord(b->e0, b->stream));
			if (two) HIP_CHECK(hipStreamWaitEvent(b->stream2, b->e0, 0));
			for (int r = 0; r < repeats; r++) {
				const int k = two ? (r & 1) : 0;
				b->trk->LocalizeSplit(b->rois.p, b->pdt, b->jobs.as<DevJob>(), b->results.as<DevResult>(), n, mp,
				                      tab, sc[k], b->diag ? b->best.as<int>() : nullptr, k ? b->stream2 : b->stream,
				                      nullptr, 0, mpb == 0 ? 1 : mpb);
			}
			if (two) {
				HIP_CHECK(hipEventRecord(b->e2, b->stream2));
				HIP_CHECK(hipStreamWaitEvent(b->stream, b->e2, 0));
			}
			HIP_CHECK(hipEventRecord(b->e1, b->stream));
			return;
		}
		// Repeats of the batch alternate between two streams, so that a launch's first
		// workgroups fill the CUs the previous launch's last ones leave idle, as consecutive
		// camera batches do; e0 .. e1 spans both.  They write the same (identical) results.
		HIP_CHECK(hipEventRecord(b->e0, b->stream));
		if (repeats > 1) HIP_CHECK(hipStreamWaitEvent(b->stream2, b->e0, 0));
		for (int r = 0; r < repeats; r++) HIP_CHECK(launch_localize(P, (r & 1) ? b->stream2 : b->stream));
		if (repeats > 1) {
			HIP_CHECK(hipEventRecord(b->e2, b->stream2));
			HIP_CHECK(hipStreamWaitEvent(b->stream, b->e2, 0));
		}
		HIP_CHECK(hipEventRecord(b->e1, b->stream));
	});
	return ok ? 0 : -1;
}

QTRK_API int QTrkBatchSetMatchPrecision(QTrkBatch* b, int precision)
{
	bool ok = GUARD({
		if (!b) fail("null batch");
		if (precision < kMatchFP32 || precision > kMatchFP32Grouped) fail("unknown match precision");
		HIP_CHECK(hipSetDevice(b->trk->Device()));
		if (precision != kMatchFP32) {
			const std::string why = b->trk->SplitMatchUsable();
			if (!why.empty()) fail("the grouped Z-LUT match cannot run: " + why);
			const LutMatchTables* tab = b->trk->PrepareMatchTables(match_kernel_precision(precision));
			b->split.ensure(QueuedHIPTracker::SplitScratch::Bytes(b->maxROIs, b->trk->cfg.zlut_radialsteps, tab->beads));
		}
		b->prec = precision;
	});
	return ok ? 0 : -1;
}

QTRK_API float QTrkBatchSync(QTrkBatch* b)
{
	float ms = -1.0f;
	GUARD({
		HIP_CHECK(hipStreamSynchronize(b->stream));
		HIP_CHECK(hipEventElapsedTime(&ms, b->e0, b->e1));
	});
	return ms;
}

QTRK_API int QTrkBatchDownload(QTrkBatch* b, LocalizationResult* out, int n)
{
	bool ok = GUARD({
		if (n < 0 || n > b->maxROIs) fail("n out of range");
		HIP_CHECK(hipStreamSynchronize(b->stream));
		HIP_CHECK(hipMemcpy(out, b->results.p, sizeof(DevResult) * n, hipMemcpyDeviceToHost));
		if (b->trk->LocalizeMode() & LT_LocalizeZ) b->trk->ApplyZBias(out, n);
	});
	return ok ? 0 : -1;
}

QTRK_API int QTrkBatchEnableDiagnostics(QTrkBatch* b, bool enable)
{
	bool ok = GUARD({
		if (enable) {
			b->best.ensure(sizeof(int) * b->maxROIs);
			b->prof.ensure(sizeof(float) * b->maxROIs * (size_t)std::max(1, b->trk->cfg.zlut_radialsteps));
		}
		b->diag = enable;
	});
	return ok ? 0 : -1;
}

QTRK_API int QTrkBatchDownloadDiagnostics(QTrkBatch* b, int* bestPlane, float* profiles, int n)
{
	bool ok = GUARD({
		if (!b->diag) fail("diagnostics are not enabled");
		if (n < 0 || n > b->maxROIs) fail("n out of range");
		HIP_CHECK(hipStreamSynchronize(b->stream));
		if (bestPlane) HIP_CHECK(hipMemcpy(bestPlane, b->best.p, sizeof(int) * n, hipMemcpyDeviceToHost));
		if (profiles)
			HIP_CHECK(hipMemcpy(profiles, b->prof.p, sizeof(float) * n * (size_t)b->trk->cfg.zlut_radialsteps,
			                    hipMemcpyDeviceToHost));
	});
	return ok ? 0 : -1;
}

// ---------------------------------------------------------------- Z-bias table

QTRK_API int QTrkComputeZLUTBiasTable(QueuedTracker* qtrk, int biasPlanes, float* result, int smpPerPixel,
                                      bool useSplineInterp)
{
	bool ok = GUARD({
		QueuedHIPTracker* t = dynamic_cast<QueuedHIPTracker*>(qtrk);
		if (!t) fail("not a MI355X tracker");
		t->ComputeZBiasCorrection(biasPlanes, result, smpPerPixel, useSplineInterp);
	});
	return ok ? 0 : -1;
}

QTRK_API int QTrkSetZLUTBiasTable(QueuedTracker* qtrk, const float* table, int biasPlanes, int count)
{
	bool ok = GUARD({
		if (!qtrk) fail("null tracker");
		if (table) {
			int beads = 0, planes = 0, steps = 0;
			qtrk->GetRadialZLUTSize(beads, planes, steps);
			if (count != beads)
				fail("bias table must be [numbeads][biasPlanes]; the Z-LUT has " + std::to_string(beads) + " beads");
			if (biasPlanes < 1) fail("biasPlanes must be >= 1");
		}
		qtrk->SetZLUTBiasCorrection(table, biasPlanes, count);
	});
	return ok ? 0 : -1;
}

QTRK_API int QTrkGetZLUTBiasTable(QueuedTracker* qtrk, float* dst, int* biasPlanes, int* count)
{
	int r = -1;
	GUARD({
		if (!qtrk) fail("null tracker");
		r = qtrk->GetZLUTBiasCorrection(dst, biasPlanes, count) ? 1 : 0;
	});
	return r;
}
