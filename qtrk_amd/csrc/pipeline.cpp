// pipeline.cpp -- Pipeline: per-tracker chains, launch-shape tuning, graph
// capture and camera-paced streaming (the design is stated in pipeline.hpp).
#include "pipeline.hpp"

#include "errors.hpp"

#include <algorithm>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <sstream>

namespace qtrk {

static int pdt_bytes(int pdt) { return pdt == 0 ? 1 : pdt == 1 ? 2 : 4; }

Pipeline::Pipeline()
{
	HIP_CHECK(hipGetDevice(&dev));
	HIP_CHECK(hipStreamCreateWithFlags(&s0, hipStreamNonBlocking));
	HIP_CHECK(hipEventCreateWithFlags(&e_fork, hipEventDisableTiming));
	HIP_CHECK(hipEventCreate(&e0));
	HIP_CHECK(hipEventCreate(&e1));
}

Pipeline::~Pipeline()
{
	hipSetDevice(dev);
	hipStreamSynchronize(s0);
	release_graph();
	for (auto& t : ts) {
		if (t->stream) {
			hipStreamSynchronize(t->stream);
			hipStreamDestroy(t->stream);
		}
		if (t->done) hipEventDestroy(t->done);
	}
	hipEventDestroy(e_fork);
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	hipStreamDestroy(s0);
}

void Pipeline::release_graph()
{
	if (exec) hipGraphExecDestroy(exec);
	if (graph) hipGraphDestroy(graph);
	exec = nullptr;
	graph = nullptr;
	finalized = false;
}

PipelineTracker& Pipeline::at(int slot)
{
	if (slot < 0 || slot >= (int)ts.size()) fail("no such tracker slot");
	return *ts[slot];
}

int Pipeline::Add(QueuedHIPTracker* trk, int n, int pdt, int precision, const float* model_luts, int model_count,
                  int model_planes, int model_steps, float model_minr, float model_maxr, int iters)
{
	if (!trk) fail("not a MI355X tracker");
	if (trk->Device() != dev) fail("the tracker runs on another device");
	if (n <= 0) fail("roisPerBatch must be > 0");
	if (pdt < 0 || pdt > 2) fail("unknown pixel type");
	if (precision < kMatchFP32 || precision > kMatchFP16) fail("unknown match precision");
	if (!trk->DeviceZLUT() || trk->ZPlanes() <= 0) fail("the tracker has no radial Z-LUT");
	if (trk->LocalizeMode() & (LT_ZLUTAlign | LT_LocalizeZWeighted | LT_FourierLUT))
		fail("LT_ZLUTAlign, LT_LocalizeZWeighted and LT_FourierLUT need the per-ROI compare: the pipeline's "
		     "grouped match scores one normalised radial profile per ROI");
	if (iters < 0) fail("mleIterations must be >= 0");
	release_graph();
	std::unique_ptr<PipelineTracker> t(new PipelineTracker());
	t->trk = trk;
	t->n = n;
	t->pdt = pdt;
	t->width = trk->cfg.width;
	t->height = trk->cfg.height;
	t->beads = trk->ZCount();
	t->precision = precision;
	t->iters = iters;
	const int nb = trk->cfg.zlut_radialsteps;
	HIP_CHECK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
	HIP_CHECK(hipEventCreateWithFlags(&t->done, hipEventDisableTiming));
	t->rois.ensure((size_t)n * t->width * t->height * pdt_bytes(pdt));
	t->jobs.ensure((size_t)n * sizeof(DevJob));
	t->results.ensure((size_t)n * sizeof(DevResult));
	t->prof.ensure((size_t)n * nb * sizeof(float));
	t->z.ensure((size_t)n * sizeof(float));
	t->best.ensure((size_t)n * sizeof(int));
	t->cand.ensure((size_t)n * kMatchTopK * sizeof(int));
	t->start.ensure((size_t)(t->beads + 1) * sizeof(int));
	t->rowidx.ensure((size_t)n * sizeof(int));
	t->match.reset(new LutMatchTables(t->beads, trk->ZPlanes(), nb));
	t->match->set(trk->DeviceZLUT(), trk->DeviceZCmp(), true, s0);
	t->match->prepare(precision, s0);
	t->lut_version = trk->LUTVersion();
	if (model_luts) {
		if (model_count <= 0 || model_planes < 4 || model_steps < 2 || !(model_maxr > model_minr))
			fail("invalid model LUT");
		t->mle = true;
		const size_t mb = (size_t)model_count * model_planes * model_steps * sizeof(float);
		t->model_lut.ensure(mb);
		HIP_CHECK(hipMemcpyAsync(t->model_lut.p, model_luts, mb, hipMemcpyHostToDevice, s0));
		LutModelParams& m = t->model;
		m.luts = t->model_lut.as<float>();
		m.count = model_count;
		m.planes = model_planes;
		m.steps = model_steps;
		m.minr = model_minr;
		m.maxr = model_maxr;
		m.width = t->width;
		m.height = t->height;
		m.ovs = 4;
		m.spline = 1;
		if (lutmodel_lds_bytes(m, true) > 160 * 1024) fail("ROI too large for the MLE (LDS)");
		t->theta.ensure((size_t)n * kModelParams * sizeof(float));
		t->crlb.ensure((size_t)n * kModelParams * sizeof(float));
	}
	HIP_CHECK(hipStreamSynchronize(s0));
	ts.push_back(std::move(t));
	return (int)ts.size() - 1;
}

void Pipeline::SetJobs(int slot, const LocalizationJob* jobs, int n)
{
	PipelineTracker& t = at(slot);
	if (!jobs || n != t.n) fail("SetJobs needs exactly roisPerBatch jobs");
	release_graph();
	HIP_CHECK(hipMemcpyAsync(t.jobs.p, jobs, (size_t)n * sizeof(DevJob), hipMemcpyHostToDevice, s0));
	// static grouping of the batch's ROIs by bead for the grouped LUT match
	std::vector<int> count(t.beads + 1, 0), rows;
	for (int i = 0; i < n; i++) {
		const int b = jobs[i].zlutIndex;
		if (b >= 0 && b < t.beads) count[b + 1]++;
	}
	t.max_per_bead = 0;
	for (int b = 0; b < t.beads; b++) {
		t.max_per_bead = std::max(t.max_per_bead, count[b + 1]);
		count[b + 1] += count[b];
	}
	rows.resize(std::max(count[t.beads], 1));
	std::vector<int> fill(count.begin(), count.end() - 1);
	for (int i = 0; i < n; i++) {
		const int b = jobs[i].zlutIndex;
		if (b >= 0 && b < t.beads) rows[fill[b]++] = i;
	}
	HIP_CHECK(hipMemcpyAsync(t.start.p, count.data(), count.size() * sizeof(int), hipMemcpyHostToDevice, s0));
	HIP_CHECK(hipMemcpyAsync(t.rowidx.p, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice, s0));
	// ROIs without a LUT keep z = NaN
	std::vector<float> nanz(n, NAN);
	HIP_CHECK(hipMemcpyAsync(t.z.p, nanz.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice, s0));
	HIP_CHECK(hipStreamSynchronize(s0));
	t.jobs_set = true;
}

void Pipeline::Upload(int slot, const void* rois, int n)
{
	PipelineTracker& t = at(slot);
	if (!rois || n < 0 || n > t.n) fail("n must be in [0, roisPerBatch]");
	HIP_CHECK(hipMemcpyAsync(t.rois.p, rois, (size_t)n * t.width * t.height * pdt_bytes(t.pdt),
	                         hipMemcpyHostToDevice, s0));
	HIP_CHECK(hipStreamSynchronize(s0));
}

void* Pipeline::DeviceROIs(int slot) { return at(slot).rois.p; }

void Pipeline::enqueue_localize(PipelineTracker& t, hipStream_t s)
{
	LocalizeParams P;
	t.trk->FillParams(P, t.pdt, t.n, t.rois.p, nullptr, t.width * pdt_bytes(t.pdt), t.jobs.as<DevJob>(),
	                  t.results.as<DevResult>());
	P.mode |= kModeLocalizeZ | kModeNormalizeProfile;
	P.mode &= ~(kModeZLUTAlign | kModeLocalizeZWeighted | kModeFourierLUT);
	P.zluts = nullptr;  // profiles only: the grouped match scores them
	P.prof_out = t.prof.as<float>();
	P.best_out = nullptr;
	P.cmp_out = nullptr;
	P.lds_pix = t.lds_pix;
	HIP_CHECK(launch_localize(P, s));
}

void Pipeline::enqueue_match(PipelineTracker& t, hipStream_t s)
{
	if (t.max_per_bead <= 0) return;
	LutMatchParams p = t.match->params(t.precision);
	p.prof = t.prof.as<float>();
	p.start = t.start.as<int>();
	p.rowidx = t.rowidx.as<int>();
	p.z = t.z.as<float>();
	p.best = t.best.as<int>();
	p.cand = t.precision != kMatchFP32 ? t.cand.as<int>() : nullptr;
	HIP_CHECK(launch_lutmatch(p, t.precision, t.max_per_bead, s, t.precision != kMatchFP32 ? t.frame_tile : 0));
}

void Pipeline::enqueue_mle(PipelineTracker& t, hipStream_t s)
{
	if (!t.mle) return;
	LutModelParams m = t.model;
	m.n = t.n;
	// the start: the localization's x/y and the matched z, read by the fit kernel itself
	m.theta = nullptr;
	m.start_xy = (const float*)((const char*)t.results.p + offsetof(DevResult, x));
	m.start_stride = (int)(sizeof(DevResult) / sizeof(float));
	m.start_z = t.z.as<float>();
	m.lut_index = m.count > 1 ? (const int*)((const char*)t.jobs.p + offsetof(DevJob, zlutIndex)) : nullptr;
	m.lut_index_stride = (int)(sizeof(DevJob) / sizeof(int));
	m.data = t.rois.p;
	m.data_pdt = t.pdt;
	m.theta_out = t.theta.as<float>();
	m.iters = t.iters;
	m.crlb = t.crlb.as<float>();
	HIP_CHECK(launch_lutmodel_fit(m, s, t.mle_block));
}

void Pipeline::enqueue_chain(PipelineTracker& t, hipStream_t s)
{
	enqueue_localize(t, s);
	enqueue_match(t, s);
	enqueue_mle(t, s);
}

// Mean event time of `reps` runs of f on s0, after one untimed run.
template <typename F>
static float time_on(hipStream_t s, hipEvent_t a, hipEvent_t b, int reps, F f)
{
	f();
	HIP_CHECK(hipEventRecord(a, s));
	for (int i = 0; i < reps; i++) f();
	HIP_CHECK(hipEventRecord(b, s));
	HIP_CHECK(hipEventSynchronize(b));
	float ms = 0;
	HIP_CHECK(hipEventElapsedTime(&ms, a, b));
	return ms / reps;
}

void Pipeline::tune(PipelineTracker& t)
{
	std::ostringstream os;
	// 1. ROI staging in LDS: as float, or in the source format (u8/u16) when no calibration rewrites pixels
	std::vector<int> lp = { 2 };
	if (t.pdt != 2 && !t.trk->HasCalibration()) lp.push_back(t.pdt);
	float best = 1e30f;
	int pick = 2;
	for (int v : lp) {
		t.lds_pix = v;
		LocalizeParams P;
		t.trk->FillParams(P, t.pdt, t.n, t.rois.p, nullptr, 0, nullptr, nullptr);
		P.mode |= kModeLocalizeZ;
		P.lds_pix = v;
		if (localize_lds_bytes(P) > 160 * 1024) continue;
		const float ms = time_on(s0, e0, e1, 3, [&] { enqueue_localize(t, s0); });
		os << " localize[lds " << (v == 2 ? "f32" : v == 1 ? "u16" : "u8") << "] " << ms << " ms;";
		if (ms < best) { best = ms; pick = v; }
	}
	t.lds_pix = pick;
	enqueue_localize(t, s0);
	// 2. profiles per matrix-core LUT-match workgroup
	if (t.precision != kMatchFP32 && t.max_per_bead > 0) {
		std::vector<int> ft = { 32 };
		if (t.max_per_bead > 32) ft.push_back(64);
		best = 1e30f;
		int pf = ft.back();
		for (int v : ft) {
			t.frame_tile = v;
			const float ms = time_on(s0, e0, e1, 3, [&] { enqueue_match(t, s0); });
			os << " match[tile " << v << "] " << ms << " ms;";
			if (ms < best) { best = ms; pf = v; }
		}
		t.frame_tile = pf;
	}
	enqueue_match(t, s0);
	// 3. lanes per MLE workgroup
	if (t.mle) {
		best = 1e30f;
		int pb = 256;
		for (int v : { 64, 128, 256, 512, 1024 }) {
			if (!lutmodel_fit_block_ok(v, t.model.ovs) || lutmodel_lds_bytes(t.model, true, v) > 160 * 1024) continue;
			t.mle_block = v;
			const float ms = time_on(s0, e0, e1, 1, [&] { enqueue_mle(t, s0); });
			os << " mle[block " << v << "] " << ms << " ms;";
			if (ms < best) { best = ms; pb = v; }
		}
		t.mle_block = pb;
	}
	HIP_CHECK(hipStreamSynchronize(s0));
	t.chain_ms = time_on(s0, e0, e1, 1, [&] { enqueue_chain(t, s0); });
	std::ostringstream ch;
	ch << "chain " << t.chain_ms << " ms; chosen: lds " << (t.lds_pix == 2 ? "f32" : t.lds_pix == 1 ? "u16" : "u8") << ", match tile "
	   << t.frame_tile << ", mle block " << t.mle_block << ";" << os.str();
	t.tuning = ch.str();
}

void Pipeline::Finalize(bool autotune)
{
	if (ts.empty()) fail("no trackers added");
	for (auto& t : ts)
		if (!t->jobs_set) fail("SetJobs was not called for every tracker");
	release_graph();
	HIP_CHECK(hipSetDevice(dev));
	for (auto& t : ts) {
		if (t->trk->ZPlanes() != t->match->planes || t->trk->ZCount() != t->match->beads)
			fail("a tracker's LUT changed size after it was added");
		if (t->trk->LUTVersion() != t->lut_version) {  // LUT or weights set, built or normalised since
			t->match->set(t->trk->DeviceZLUT(), t->trk->DeviceZCmp(), true, s0);
			t->match->prepare(t->precision, s0);
			HIP_CHECK(hipStreamSynchronize(s0));
			t->lut_version = t->trk->LUTVersion();
		}
		if (autotune) tune(*t);
		else {
			t->lds_pix = 2;
			t->frame_tile = 0;
			t->mle_block = 256;
			t->chain_ms = 0;
			t->tuning = "defaults";
		}
	}
	// The graph's kernels are submitted one node after another in capture order, some microseconds
	// each, so a branch captured last starts that much later.  Capture stage by stage (every
	// tracker's localization, then every match, then every MLE), the trackers with the longest
	// chain first: each branch starts within the first few submissions and the longest (the
	// critical path of a batch) first of all.
	std::vector<PipelineTracker*> order;
	for (auto& t : ts) order.push_back(t.get());
	std::stable_sort(order.begin(), order.end(),
	                 [](const PipelineTracker* a, const PipelineTracker* b) { return a->chain_ms > b->chain_ms; });
	// one untimed pass sets every kernel's LDS attribute before capture
	for (auto& t : ts) enqueue_chain(*t, s0);
	HIP_CHECK(hipStreamSynchronize(s0));
	// capture: s0 forks to one stream per tracker and joins them
	HIP_CHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
	try {
		HIP_CHECK(hipEventRecord(e_fork, s0));
		for (PipelineTracker* t : order) HIP_CHECK(hipStreamWaitEvent(t->stream, e_fork, 0));
		for (PipelineTracker* t : order) enqueue_localize(*t, t->stream);
		for (PipelineTracker* t : order) enqueue_match(*t, t->stream);
		for (PipelineTracker* t : order) enqueue_mle(*t, t->stream);
		for (PipelineTracker* t : order) {
			HIP_CHECK(hipEventRecord(t->done, t->stream));
			HIP_CHECK(hipStreamWaitEvent(s0, t->done, 0));
		}
	} catch (...) {
		hipGraph_t g = nullptr;
		hipStreamEndCapture(s0, &g);
		if (g) hipGraphDestroy(g);
		throw;
	}
	HIP_CHECK(hipStreamEndCapture(s0, &graph));
	HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
	finalized = true;
}

void Pipeline::Launch()
{
	if (!finalized) fail("call QTrkPipelineFinalize first");
	for (auto& t : ts)
		if (t->trk->LUTVersion() != t->lut_version)
			fail("a tracker's Z-LUT or radial weights changed after QTrkPipelineFinalize: finalize again");
	HIP_CHECK(hipSetDevice(dev));
	HIP_CHECK(hipEventRecord(e0, s0));
	HIP_CHECK(hipGraphLaunch(exec, s0));
	HIP_CHECK(hipEventRecord(e1, s0));
}

float Pipeline::Sync()
{
	HIP_CHECK(hipSetDevice(dev));
	HIP_CHECK(hipStreamSynchronize(s0));
	float ms = 0;
	HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
	return ms;
}

void Pipeline::Download(int slot, LocalizationResult* out, float* theta, float* crlb, int n)
{
	PipelineTracker& t = at(slot);
	if (n < 0 || n > t.n) fail("n must be in [0, roisPerBatch]");
	HIP_CHECK(hipSetDevice(dev));
	HIP_CHECK(hipStreamSynchronize(s0));
	if (out) {
		static_assert(sizeof(DevResult) == sizeof(LocalizationResult), "result layout");
		HIP_CHECK(hipMemcpy(out, t.results.p, (size_t)n * sizeof(DevResult), hipMemcpyDeviceToHost));
		std::vector<float> z(n);
		HIP_CHECK(hipMemcpy(z.data(), t.z.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
		for (int i = 0; i < n; i++) out[i].pos.z = z[i];
		t.trk->ApplyZBias(out, n);  // the chain always localizes in z
	}
	if ((theta || crlb) && !t.mle) fail("this tracker's chain has no MLE stage");
	if (theta) HIP_CHECK(hipMemcpy(theta, t.theta.p, (size_t)n * kModelParams * sizeof(float), hipMemcpyDeviceToHost));
	if (crlb) HIP_CHECK(hipMemcpy(crlb, t.crlb.p, (size_t)n * kModelParams * sizeof(float), hipMemcpyDeviceToHost));
}

// Camera-paced submission: batch i is launched at t0 + i * period (or as soon
// after as the host gets there); its latency is the host time from that
// submission until its completion event is observed.
void Pipeline::Stream(int batches, double period_ms, float* latency_ms, float* gpu_ms)
{
	if (!finalized) fail("call QTrkPipelineFinalize first");
	if (batches <= 0 || !latency_ms) fail("invalid arguments");
	HIP_CHECK(hipSetDevice(dev));
	using clk = std::chrono::steady_clock;
	std::vector<hipEvent_t> ea(batches), eb(batches);
	for (int i = 0; i < batches; i++) {
		HIP_CHECK(hipEventCreate(&ea[i]));
		HIP_CHECK(hipEventCreate(&eb[i]));
	}
	std::vector<clk::time_point> sub(batches);
	std::vector<double> lat(batches, -1.0);
	int next_done = 0;
	auto poll = [&] {
		while (next_done < batches && lat[next_done] < 0 && next_done < batches) {
			if (sub[next_done].time_since_epoch().count() == 0) break;
			const hipError_t q = hipEventQuery(eb[next_done]);
			if (q == hipErrorNotReady) break;
			HIP_CHECK(q);
			lat[next_done] = std::chrono::duration<double, std::milli>(clk::now() - sub[next_done]).count();
			next_done++;
		}
	};
	const clk::time_point t0 = clk::now();
	for (int i = 0; i < batches; i++) {
		const clk::time_point due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double, std::milli>(period_ms * i));
		while (clk::now() < due) poll();
		sub[i] = clk::now();
		HIP_CHECK(hipEventRecord(ea[i], s0));
		HIP_CHECK(hipGraphLaunch(exec, s0));
		HIP_CHECK(hipEventRecord(eb[i], s0));
		poll();
	}
	while (next_done < batches) poll();
	for (int i = 0; i < batches; i++) {
		latency_ms[i] = (float)lat[i];
		if (gpu_ms) HIP_CHECK(hipEventElapsedTime(&gpu_ms[i], ea[i], eb[i]));
	}
	for (int i = 0; i < batches; i++) {
		hipEventDestroy(ea[i]);
		hipEventDestroy(eb[i]);
	}
}

std::string Pipeline::Report() const
{
	std::ostringstream os;
	for (size_t i = 0; i < ts.size(); i++) {
		const PipelineTracker& t = *ts[i];
		os << "tracker " << i << " (" << t.width << "x" << t.height << ", " << t.n << " ROIs, " << t.beads
		   << " beads, " << t.match->planes << " planes x " << t.match->res << " steps, match "
		   << (t.precision == kMatchFP32 ? "fp32" : t.precision == kMatchBF16 ? "bf16" : "fp16") << ", mle "
		   << (t.mle ? std::to_string(t.iters) + " iterations" : std::string("off")) << "): " << t.tuning << "\n";
	}
	return os.str();
}

}  // namespace qtrk

// ------------------------------------------------------------ C ABI

#include "qtrk_amd.h"

using qtrk::Pipeline;

struct QTrkPipeline {
	Pipeline p;
};

QTRK_API QTrkPipeline* QTrkPipelineCreate(void)
{
	QTrkPipeline* p = nullptr;
	GUARD({ p = new QTrkPipeline(); });
	return p;
}

QTRK_API void QTrkPipelineFree(QTrkPipeline* p) { delete p; }

QTRK_API int QTrkPipelineAddTracker(QTrkPipeline* p, QueuedTracker* qtrk, int roisPerBatch, QTRK_PixelDataType pdt,
                                    int matchPrecision, const float* modelLUTs, int modelCount, int modelPlanes,
                                    int modelSteps, float modelMinRadius, float modelMaxRadius, int mleIterations)
{
	int slot = -1;
	GUARD({
		if (!p) qtrk::fail("null pipeline");
		HIP_CHECK(hipSetDevice(p->p.Device()));
		slot = p->p.Add(dynamic_cast<qtrk::QueuedHIPTracker*>(qtrk), roisPerBatch, (int)pdt, matchPrecision, modelLUTs,
		                modelCount, modelPlanes, modelSteps, modelMinRadius, modelMaxRadius, mleIterations);
	});
	return slot;
}

QTRK_API int QTrkPipelineSetJobs(QTrkPipeline* p, int slot, const LocalizationJob* jobs, int n)
{
	return GUARD({
		if (!p) qtrk::fail("null pipeline");
		HIP_CHECK(hipSetDevice(p->p.Device()));
		p->p.SetJobs(slot, jobs, n);
	}) ? 0 : -1;
}

QTRK_API void* QTrkPipelineDeviceROIs(QTrkPipeline* p, int slot)
{
	void* d = nullptr;
	GUARD({
		if (!p) qtrk::fail("null pipeline");
		d = p->p.DeviceROIs(slot);
	});
	return d;
}

QTRK_API int QTrkPipelineUploadROIs(QTrkPipeline* p, int slot, const void* rois, int n)
{
	return GUARD({
		if (!p) qtrk::fail("null pipeline");
		HIP_CHECK(hipSetDevice(p->p.Device()));
		p->p.Upload(slot, rois, n);
	}) ? 0 : -1;
}

QTRK_API int QTrkPipelineFinalize(QTrkPipeline* p, bool autotune)
{
	return GUARD({
		if (!p) qtrk::fail("null pipeline");
		p->p.Finalize(autotune);
	}) ? 0 : -1;
}

QTRK_API int QTrkPipelineLaunch(QTrkPipeline* p)
{
	return GUARD({
		if (!p) qtrk::fail("null pipeline");
		p->p.Launch();
	}) ? 0 : -1;
}

QTRK_API float QTrkPipelineSync(QTrkPipeline* p)
{
	float ms = -1;
	GUARD({
		if (!p) qtrk::fail("null pipeline");
		ms = p->p.Sync();
	});
	return ms;
}

QTRK_API int QTrkPipelineDownload(QTrkPipeline* p, int slot, LocalizationResult* results, float* theta, float* crlb,
                                  int n)
{
	return GUARD({
		if (!p) qtrk::fail("null pipeline");
		p->p.Download(slot, results, theta, crlb, n);
	}) ? 0 : -1;
}

QTRK_API int QTrkPipelineStream(QTrkPipeline* p, int batches, float periodMs, float* latencyMs, float* gpuMs)
{
	return GUARD({
		if (!p) qtrk::fail("null pipeline");
		p->p.Stream(batches, periodMs, latencyMs, gpuMs);
	}) ? 0 : -1;
}

QTRK_API int QTrkPipelineGetReport(QTrkPipeline* p, char* dst, int maxLen)
{
	return GUARD({
		if (!p || !dst || maxLen <= 0) qtrk::fail("invalid arguments");
		const std::string r = p->p.Report();
		snprintf(dst, (size_t)maxLen, "%s", r.c_str());
	}) ? 0 : -1;
}
