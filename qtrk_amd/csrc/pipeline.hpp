// pipeline.hpp -- several trackers' full localization chains in one hipGraph.
//
// Per tracker (its own ROI size, settings, LUT and pixel type) the chain is
//   COM -> QI -> radial profile     (localize_kernel, the tracker's settings, no per-ROI LUT compare)
//   -> grouped Z-LUT match           (lutmatch, a converted copy of the tracker's LUT, BF16/FP16/FP32,
//                                     profiles grouped by bead; remade at finalize when the LUT changed)
//   -> start assembly -> LUT-model MLE (lutmodel_fit_kernel on the ROI pixels, the tracker's image LUT).
// The chains of all trackers are captured as parallel branches of one graph,
// so one hipGraphLaunch serves every tracker's batch.  Launch shapes (LDS
// staging format of the ROI, profiles per LUT-match workgroup, lanes per MLE
// workgroup) are chosen per tracker by timing each candidate on the data
// present when the pipeline is finalized.
#pragma once

#include "devbuf.hpp"
#include "hip_tracker.hpp"
#include "localize.hpp"
#include "lutmatch.hpp"
#include "lutmodel.hpp"

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

namespace qtrk {


struct PipelineTracker {
	QueuedHIPTracker* trk = nullptr;
	int n = 0, pdt = 0, width = 0, height = 0, beads = 0;
	int precision = kMatchBF16, iters = 5;
	uint64_t lut_version = 0;  // the tracker's LUTVersion() the match tables were made from
	bool mle = false;
	DeviceBuffer rois, jobs, results, prof, z, best, cand, theta, crlb;
	DeviceBuffer start, rowidx;
	int max_per_bead = 0;
	bool jobs_set = false;
	std::unique_ptr<LutMatchTables> match;
	DeviceBuffer model_lut;
	LutModelParams model = {};
	// launch shapes (autotuned)
	int lds_pix = 2, frame_tile = 0, mle_block = 256;
	float chain_ms = 0;  // the tuned chain's time alone on the GPU (0 when not autotuned)
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	std::string tuning;
};

class Pipeline {
public:
	Pipeline();
	~Pipeline();
	int Add(QueuedHIPTracker* trk, int rois_per_batch, int pdt, int precision, const float* model_luts, int model_count,
	        int model_planes, int model_steps, float model_minr, float model_maxr, int iters);
	void SetJobs(int slot, const LocalizationJob* jobs, int n);
	void Upload(int slot, const void* rois, int n);
	void* DeviceROIs(int slot);
	void Finalize(bool autotune);
	void Launch();
	float Sync();
	void Download(int slot, LocalizationResult* results, float* theta, float* crlb, int n);
	void Stream(int batches, double period_ms, float* latency_ms, float* gpu_ms);
	std::string Report() const;
	int Device() const { return dev; }

private:
	PipelineTracker& at(int slot);
	void enqueue_localize(PipelineTracker& t, hipStream_t s);
	void enqueue_match(PipelineTracker& t, hipStream_t s);
	void enqueue_mle(PipelineTracker& t, hipStream_t s);
	void enqueue_chain(PipelineTracker& t, hipStream_t s);
	void tune(PipelineTracker& t);
	void release_graph();

	int dev = 0;
	std::vector<std::unique_ptr<PipelineTracker>> ts;
	hipStream_t s0 = nullptr;
	hipEvent_t e_fork = nullptr, e0 = nullptr, e1 = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	bool finalized = false;
};

}  // namespace qtrk
