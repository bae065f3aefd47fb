// noalloc.hip -- QTrkLocalizeDeviceNoAlloc: the device batch's localization (QTrkBatchRun, with
// the match precision of QTrkBatchSetMatchPrecision) on the caller's device buffers, for callers
// that cannot allocate GPU memory.  Its scratch (radial profiles, grouping by bead) lives in a
// file-scope device array of kNoAllocScratchBytes that every GPU the library runs on holds until
// the process ends; all calls on one GPU share it.
#include "errors.hpp"
#include "hip_tracker.hpp"
#include "qtrk_amd.h"

#include <mutex>
#include <string>

using namespace qtrk;

static constexpr size_t kNoAllocScratchBytes = (size_t)QTRK_NOALLOC_SCRATCH_MB << 20;
static constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The array's address on `dev` (each GPU holds its own copy); looked up once per device, with the
// device current.  `scratch` itself is only a host-side stand-in and is never passed to a kernel.
static void* scratch_on(int dev)
{
	static std::mutex mu;
	static void* addr[kMaxDevices] = {};
	if (dev < 0 || dev >= kMaxDevices) fail("device index out of range");
	std::lock_guard<std::mutex> g(mu);
	if (!addr[dev]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) fail(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
		addr[dev] = p;
	}
	return addr[dev];
}

QTRK_API int QTrkLocalizeDeviceNoAlloc(QueuedTracker* qtrk, const void* rois, QTRK_PixelDataType pdt,
                                       const LocalizationJob* jobs, LocalizationResult* results, int n, int precision,
                                       void* stream)
{
	bool ok = GUARD({
		auto* t = dynamic_cast<QueuedHIPTracker*>(qtrk);
		if (!t) fail("not a MI355X tracker");
		if (n < 0) fail("n must be >= 0");
		if ((int)pdt < 0 || (int)pdt > 2) fail("invalid pixel data type");
		if (precision < kMatchFP32 || precision > kMatchFP32Grouped) fail("unknown match precision");
		if (n > 0 && (!rois || !jobs || !results)) fail("null buffer");
		HIP_CHECK(hipSetDevice(t->Device()));
		const hipStream_t s = (hipStream_t)stream;
		void* base = scratch_on(t->Device());
		if (n == 0) return;
		const DevJob* dj = (const DevJob*)jobs;
		DevResult* dr = (DevResult*)results;
		if (precision == kMatchFP32) {
			// the per-ROI compare inside the localization kernel: no scratch
			LocalizeParams P;
			t->FillParams(P, (int)pdt, n, rois, nullptr, t->cfg.width * PDT_BytesPerPixel(pdt), dj, dr);
			HIP_CHECK(launch_localize(P, s));
			return;
		}
		const std::string why = t->SplitMatchUsable();
		if (!why.empty()) fail("the grouped Z-LUT match cannot run: " + why);
		const int mp = match_kernel_precision(precision);
		const LutMatchTables* tab = t->CurrentMatchTables(mp);
		if (!tab)
			fail("the tracker's LUT is not converted for this match precision (or changed since): call "
			     "QTrkBatchSetMatchPrecision on a batch of this tracker first");
		const size_t need = QueuedHIPTracker::SplitScratch::Bytes(n, t->cfg.zlut_radialsteps, tab->beads);
		if (need > kNoAllocScratchBytes)
			fail(std::to_string(n) + " ROIs need " + std::to_string(need) + " bytes of scratch; the limit is " +
			     std::to_string(kNoAllocScratchBytes) + " (" + std::to_string(QTRK_NOALLOC_SCRATCH_MB) + " MiB)");
		const QueuedHIPTracker::SplitScratch sc =
		    QueuedHIPTracker::SplitScratch::Carve(base, n, t->cfg.zlut_radialsteps, tab->beads);
		t->LocalizeSplit(rois, (int)pdt, dj, dr, n, mp, tab, sc, nullptr, s);
	});
	return ok ? 0 : -1;
}
