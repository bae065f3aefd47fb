// hip_tracker.hpp -- QueuedHIPTracker: the MI355X implementation of the
// QueuedTracker interface (cputrack/QueuedTracker.h:22-96).
//
// Host model: scheduled ROIs are copied into pinned batch slots; a slot is
// launched (H2D, fused localization kernel, D2H, event) when it holds
// batch_size ROIs or on Flush; completed slots are harvested into the result
// deque whenever a caller touches the queue.  There is no scheduler thread:
// every call that needs progress polls the slot events.
#pragma once

#include "QueuedTracker.h"
#include "copy_pool.hpp"
#include "devbuf.hpp"
#include "localize.hpp"
#include "lutmatch.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace qtrk {

// A kissfft plan of one length with its leaf order and twiddles in GPU memory.
struct HostFFT {
	int n = 0, nst = 0;
	int radix[kMaxFFTStages] = {}, m[kMaxFFTStages] = {}, fs[kMaxFFTStages] = {};
	DeviceBuffer perm, tw_fwd, tw_inv;
	void build(int len);
	FFTPlan plan() const;
};

class QueuedHIPTracker : public QueuedTracker {
public:
	explicit QueuedHIPTracker(const QTrkComputedConfig& cc);
	~QueuedHIPTracker() override;

	void SetLocalizationMode(LocMode_t locType) override;
	void SetPixelCalibrationImages(float* offset, float* gain) override;
	void SetPixelCalibrationFactors(float offsetFactor, float gainFactor) override;
	void ScheduleLocalization(void* data, int pitch, QTRK_PixelDataType pdt, const LocalizationJob* jobInfo) override;
	int ScheduleFrame(void* imgptr, int pitch, int width, int height, ROIPosition* positions, int numROI,
	                  QTRK_PixelDataType pdt, const LocalizationJob* jobInfo) override;
	void ClearResults() override;
	void Flush() override;
	void SetRadialZLUT(float* data, int count, int planes) override;
	void GetRadialZLUT(float* dst) override;
	void GetRadialZLUTSize(int& count, int& planes, int& radialsteps) override;
	void SetRadialWeights(float* zcmp) override;
	void EnableRadialZLUTCompareProfile(bool enabled) override;
	void GetRadialZLUTCompareProfile(float* dst) override;
	void BeginLUT(uint flags) override;
	void BuildLUT(void* data, int pitch, QTRK_PixelDataType pdt, int plane, vector2f* known_pos = 0) override;
	void FinalizeLUT() override;
	void GetImageZLUTSize(int* dims) override;
	void GetImageZLUT(float* dst) override;
	bool SetImageZLUT(float* src, float* radial_zlut, int* dims) override;
	// The image LUT's first and second differences across planes that FinalizeLUT computed
	// ([beads][planes][height][width] each; either pointer may be null); false when there are
	// none (no image LUT of at least 3 planes when FinalizeLUT ran, or it was set again since).
	bool GetImageZLUTDerivatives(float* dz, float* dz2);
	int GetResultCount() override;
	int FetchResults(LocalizationResult* results, int maxResults) override;
	int GetQueueLength(int* maxQueueLen = 0) override;
	bool IsIdle() override;
	void SetConfigValue(std::string name, std::string value) override;
	ConfigValueMap GetConfigValues() override;
	std::string GetProfileReport() override;
	std::string GetWarnings() override;

	// Fills the kernel parameters from the tracker's current state for n ROIs
	// at `src` (contiguous when roi_offset is null).
	void FillParams(LocalizeParams& P, int pdt, int n, const void* src, const int64_t* roi_offset, int pitch,
	                const DevJob* jobs, DevResult* results);
	int Device() const { return dev; }
	int LocalizeMode() const { return localizeMode; }
	void WaitLaunched();
	// the radial Z-LUT [ZCount()][ZPlanes()][zlut_radialsteps] and weights in GPU memory (null when unset)
	// QueuedTracker::ComputeZBiasCorrection (cputrack/QueuedTracker.cpp:143-189) on the GPU:
	// every bead's LUT rendered at biasPlanes depths and localized; the [ZCount()][biasPlanes]
	// table becomes the active bias correction and is copied to `result` when not null.
	void ComputeZBiasCorrection(int biasPlanes, float* result, int smpPerPixel, bool useSplineInterp);
	// The active Z-bias correction applied to n results' z (as ProcessJob does,
	// cputrack/QueuedCPUTracker.cpp:346-347).
	void ApplyZBias(LocalizationResult* r, int n);
	const float* DeviceZLUT() const { return zcount > 0 ? d_zlut.as<float>() : nullptr; }
	const float* DeviceZCmp() const { return has_zcmp ? d_zcmp.as<float>() : nullptr; }
	int ZCount() const { return zcount; }
	int ZPlanes() const { return zplanes; }
	// Counts changes of the radial Z-LUT or its weights (set, built, normalised), so that copies
	// of them (the grouped-match tables) can tell they are stale.
	uint64_t LUTVersion() const { return lutVersion; }

	// The grouped-match tables of the tracker's LUT and weights in precision `prec`
	// (kMatchBF16/kMatchFP16 tile means and residuals; kMatchFP32 the FP32 copy): made, or remade
	// after the LUT changed, by PrepareMatchTables (which allocates); CurrentMatchTables returns
	// them only when current and prepared (null otherwise) and allocates nothing.
	const LutMatchTables* PrepareMatchTables(int prec);
	const LutMatchTables* CurrentMatchTables(int prec) const;

	// Scratch of LocalizeSplit, carved from one block: each part 256-byte aligned.
	struct SplitScratch {
		float* prof = nullptr;  // [n][zlut_radialsteps] radial profiles
		int* counts = nullptr;  // [beads] grouping counts, then fill cursors
		int* start = nullptr;   // [beads + 1] first slot of each bead
		int* rowidx = nullptr;  // [n] rows grouped by bead
		static size_t Bytes(int n, int nb, int beads);
		static SplitScratch Carve(void* base, int n, int nb, int beads);
	};
	// Localizes n ROIs in GPU memory (contiguous [n][height][width] of pdt) into results, with the
	// Z-LUT match of all ROIs together instead of inside each ROI's workgroup: localize_kernel in
	// profile-only form (COM -> QI -> the radial profile; z left 0), the rows grouped by bead, then
	// the grouped match of precision `prec` with the tables `tab`, which writes each matched ROI's z
	// into its result.  best_out ([n], or null) receives the best planes.  Asynchronous on
	// `stream`; allocates nothing.  Needs LT_LocalizeZ without LT_ZLUTAlign, LT_LocalizeZWeighted
	// and LT_FourierLUT (SplitMatchUsable).
	// roi_offset/pitch: ROI i at rois + roi_offset[i] with rows pitch bytes apart (roi_offset null:
	// contiguous ROIs, pitch 0: width * bytes per pixel).  max_per_bead: the most ROIs of one bead
	// among the jobs when the caller knows it (> 0), which sizes the match's grid; otherwise (<= 0)
	// the match's workgroups loop over their bead's tiles.
	void LocalizeSplit(const void* rois, int pdt, const DevJob* jobs, DevResult* results, int n, int prec,
	                   const LutMatchTables* tab, const SplitScratch& s, int* best_out, hipStream_t stream,
	                   const int64_t* roi_offset = nullptr, int pitch = 0, int max_per_bead = 0);
	// The most jobs of one bead index in [0, beads) among n host jobs (0 when none).
	static int MaxPerBead(const DevJob* jobs, int n, int beads);
	// Empty when the current mode and LUT can take the split localization, else the reason.
	std::string SplitMatchUsable() const;
	bool HasCalibration() const { return has_off || has_gain; }

private:
	struct Slot {
		enum State { Free, Filling, Launched } state = Free;
		int n = 0, pdt = 0;
		uint64_t seq = 0;
		PinnedBuffer h_rois, h_jobs, h_results;
		DeviceBuffer d_rois, d_jobs, d_results;
		// frame mode (ScheduleFrame): h_rois holds row blocks of camera frames, rows
		// `frame_pitch` bytes apart, frame_bytes in all, and ROI i starts at byte h_off[i]
		bool frame = false;
		int frame_pitch = 0;
		size_t frame_bytes = 0;
		size_t frame_sent = 0;  // bytes of the row blocks already sent to the GPU
		PinnedBuffer h_off;
		DeviceBuffer d_off;
		hipStream_t stream = nullptr;
		hipEvent_t ev_start = nullptr, ev_k0 = nullptr, ev_k1 = nullptr, ev_done = nullptr;
		hipEvent_t ev_h2d = nullptr;  // after the slot's last frame block on the copy stream
		DeviceBuffer d_split;  // LocalizeSplit's scratch, when match_precision is not fp32
	};
	// the grouped-match tables for a launch on `stream` (made or remade there when stale); the
	// LUT setters wait for every batch first, so no batch in flight uses tables being remade
	const LutMatchTables* MatchTablesLocked(int prec, hipStream_t stream);

	void BuildTables();
	void AllocSlots();
	void SendFrameBlocks(Slot* s);
	void FreeSlots();
	Slot* TakeFreeSlot();
	void WaitEventLocked(std::unique_lock<std::mutex>& lk, hipEvent_t ev);
	void WaitOldestLocked(std::unique_lock<std::mutex>& lk);
	void QuiesceLocked(std::unique_lock<std::mutex>& lk);
	void EndQuiesceLocked();
	void Launch(Slot* s);
	void Harvest();
	void WaitAllLocked(std::unique_lock<std::mutex>& lk);
	int QueueLengthLocked() const;
	void Warn(const std::string& w);

	int dev = 0;
	// The synchronous set-up work (LUT build and normalisation, bias table) runs after the batches
	// in flight have finished, on the first slot's stream: every stream takes one of the process's
	// few hardware queues (GPU_MAX_HW_QUEUES), and streams that share one run one after another.
	hipStream_t aux_stream() const { return slots[0]->stream; }

	// per-config device tables
	DeviceBuffer d_qi_steps, d_qi_stride, d_qi_dirs, d_qi_offs, d_qi_window, d_z_dirs;
	HostFFT qi_fft;   // length 2*qi_radialsteps
	HostFFT xc_fft;   // length xc1_profileLength (LT_XCor1D), when 2..kMaxFFTLen
	HostFFT qa_fft;   // length 2*zlut_radialsteps (LT_ZLUTAlign), when <= kMaxFFTLen
	HostFFT fx_fft, fy_fft;  // ROI width and height (LT_FourierLUT, BUILDLUT_FOURIER), when <= kMaxFFTLen
	int qa_steps = 0, qa_stride = 0;

	// radial Z-LUT [count][planes][zlut_radialsteps], weights, compare profiles
	std::vector<float> h_zlut;
	bool h_zlut_stale = false;
	DeviceBuffer d_zlut, d_zcmp, d_cmp;
	int zcount = 0, zplanes = 0;
	bool has_zcmp = false, cmp_enabled = false;

	// per-bead calibration images
	DeviceBuffer d_calib_off, d_calib_gain;
	bool has_off = false, has_gain = false;
	float gc_offsetFactor = 1.0f, gc_gainFactor = 1.0f;

	LocMode_t localizeMode = 0;
	uint buildFlags = 0;
	DeviceBuffer d_build_rois, d_build_pos;
	// the image LUT, [beads][planes][height][width] as ilut_dims (QueuedTracker.h's order)
	DeviceBuffer d_image_lut;
	// its first and second differences across planes, made by FinalizeLUT
	DeviceBuffer d_image_lut_dz, d_image_lut_dz2;
	bool ilut_diff = false;
	int ilut_dims[4] = { 0, 0, 0, 0 };
	PinnedBuffer h_build_rois;

	int batchSize = 1024, numSlots = 4;
	std::vector<Slot*> slots;
	Slot* filling = nullptr;
	uint64_t seqCounter = 0;
	// threads, the caller's included, that copy a frame's rows into pinned memory
	int frameCopyThreads = 5;
	std::unique_ptr<CopyPool> copier;  // made on the first frame

	int matchPrec = kMatchFP32;  // "match_precision": how queued batches match the Z-LUT
	// frame blocks (ScheduleFrame) go to the GPU one after another on this stream, so a slot's
	// kernel starts as soon as its own blocks are over rather than when every slot's copies are
	hipStream_t copyStream = nullptr;
	uint64_t lutVersion = 0;
	std::unique_ptr<LutMatchTables> matchTab;
	uint64_t matchTabVersion = 0;

	mutable std::mutex mtx;
	// threads blocked on a slot's event with mtx released (WaitEventLocked); while `quiescing`
	// (the slots are being remade) scheduling threads wait; both are signalled on slots_cv
	int waiters = 0;
	bool quiescing = false;
	std::condition_variable slots_cv;
	// how long a wait on a batch polls its event before a blocking synchronisation (queue_spin_us)
	std::chrono::microseconds spinWait{ 2000 };
	std::mutex res_mtx;
	std::deque<LocalizationResult> results;

	std::string warnings;
	double t_kernel_ms = 0, t_copy_ms = 0;
	uint64_t batchesDone = 0, roisDone = 0;
	size_t lds_limit = 0;
};

}  // namespace qtrk
