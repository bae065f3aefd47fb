// hip_tracker.cpp -- host runtime of the MI355X tracker.
#include "hip_tracker.hpp"
#include "errors.hpp"
#include "lutmodel.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>
#include <thread>

namespace qtrk {

// a batch slot holds the row blocks of several frames, each starting at a
// multiple of kFrameBlockAlign bytes, up to kFrameSlotMaxBytes unless one block
// alone is larger
static constexpr size_t kFrameBlockAlign = 256;
static constexpr size_t kFrameSlotMaxBytes = (size_t)64 << 20;
// a slot's row blocks go to the GPU once this many bytes are waiting (or at launch): a camera
// frame's block goes at once, while one-ROI frames (the reference's speed test schedules every ROI
// as one) share a copy instead of paying a copy call and an event each
static constexpr size_t kFrameSendBytes = (size_t)1 << 20;

void DeviceBuffer::ensure(size_t n)
{
	if (n <= bytes && p) return;
	release();
	if (n == 0) return;
	HIP_CHECK(hipMalloc(&p, n));
	bytes = n;
}
void DeviceBuffer::release()
{
	if (p) hipFree(p);
	p = nullptr;
	bytes = 0;
}
void PinnedBuffer::ensure(size_t n)
{
	if (n <= bytes && p) return;
	release();
	if (n == 0) return;
	HIP_CHECK(hipHostMalloc(&p, n, hipHostMallocDefault));
	bytes = n;
}
void PinnedBuffer::release()
{
	if (p) hipHostFree(p);
	p = nullptr;
	bytes = 0;
}

static void upload(DeviceBuffer& b, const void* src, size_t bytes)
{
	b.ensure(bytes);
	if (bytes) HIP_CHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
}

// ComputeRadialBinWindow, cputrack/utils.cpp:665-677
static std::vector<float> radial_bin_window(int n)
{
	std::vector<float> w(n);
	for (int i = 0; i < n; i++) {
		float x = i / (float)n;
		float omx = 1 - x;
		float fall = 1.0f - expf(-(omx * omx) / 0.05f);
		float rise = 1.0f - expf(-(x * x) / 0.01f);
		w[i] = sqrtf(fall * rise * x * 2);
	}
	return w;
}

// kissfft factorisation of n (cputrack/kissfft.h:29-46): 4s first, then 2s, ...
static int fft_factor(int n, int* radix, int* rem)
{
	int m = n, f = 4, ns = 0;
	do {
		while (m % f) {
			switch (f) { case 4: f = 2; break; case 2: f = 3; break; default: f += 2; break; }
			if (f * f > m) f = m;
		}
		m /= f;
		if (ns >= kMaxFFTStages) fail("FFT length has too many factors");
		radix[ns] = f;
		rem[ns] = m;
		ns++;
	} while (m > 1);
	return ns;
}

// Input index that lands in each output slot at kissfft's leaves.
static void fft_leaves(const int* radix, const int* rem, int stage, int out, int in, int fstride, std::vector<int>& perm)
{
	for (int q = 0; q < radix[stage]; q++) {
		if (rem[stage] == 1) perm[out + q] = in + q * fstride;
		else fft_leaves(radix, rem, stage + 1, out + q * rem[stage], in + q * fstride, fstride * radix[stage], perm);
	}
}

void HostFFT::build(int len)
{
	n = len;
	int rem[kMaxFFTStages];
	nst = fft_factor(n, radix, rem);
	int f = 1;
	for (int s = 0; s < nst; s++) {
		m[s] = rem[s];
		fs[s] = f;
		f *= radix[s];
	}
	std::vector<int> leaves(n);
	fft_leaves(radix, rem, 0, 0, 0, 1, leaves);
	upload(perm, leaves.data(), leaves.size() * 4);
	// twiddles exp(+-2 pi i k / n) with the float phase step of kissfft's traits
	// (cputrack/kissfft.h:11-16); exp(i a) is (cosf a, sinf a)
	std::vector<float> tf(2 * (size_t)n), ti(2 * (size_t)n);
	const float phf = -2 * acosf(-1.0f) / n, phi = 2 * acosf(-1.0f) / n;
	for (int i = 0; i < n; i++) {
		tf[2 * i] = cosf(i * phf);
		tf[2 * i + 1] = sinf(i * phf);
		ti[2 * i] = cosf(i * phi);
		ti[2 * i + 1] = sinf(i * phi);
	}
	upload(tw_fwd, tf.data(), tf.size() * 4);
	upload(tw_inv, ti.data(), ti.size() * 4);
}

FFTPlan HostFFT::plan() const
{
	FFTPlan p{};
	p.n = n;
	p.nst = nst;
	for (int s = 0; s < kMaxFFTStages; s++) {
		p.radix[s] = radix[s];
		p.m[s] = m[s];
		p.fs[s] = fs[s];
	}
	p.perm = perm.as<int>();
	p.tw_fwd = tw_fwd.as<float2>();
	p.tw_inv = tw_inv.as<float2>();
	return p;
}

QueuedHIPTracker::QueuedHIPTracker(const QTrkComputedConfig& cc)
{
	cfg = cc;
	int ndev = 0;
	HIP_CHECK(hipGetDeviceCount(&ndev));
	if (ndev <= 0) fail("no HIP device");
	dev = cfg.cuda_device >= 0 ? cfg.cuda_device : 0;
	if (dev >= ndev) fail("cuda_device out of range");
	HIP_CHECK(hipSetDevice(dev));
	hipDeviceProp_t prop;
	HIP_CHECK(hipGetDeviceProperties(&prop, dev));
	lds_limit = prop.sharedMemPerBlock;
	if (cfg.width < 2 || cfg.height < 2) fail("ROI width and height must be >= 2");
	if (cfg.downsample != 0) fail("downsample != 0 is not supported");
	BuildTables();
	AllocSlots();
}

QueuedHIPTracker::~QueuedHIPTracker()
{
	hipSetDevice(dev);
	{
		std::unique_lock<std::mutex> lk(mtx);
		for (Slot* s : slots)
			if (s->state == Slot::Launched) hipEventSynchronize(s->ev_done);
		slots_cv.wait(lk, [&] { return waiters == 0; });
		// a filling frame slot's blocks may still be on their way to the GPU
		if (copyStream) hipStreamSynchronize(copyStream);
	}
	FreeSlots();
}

void QueuedHIPTracker::BuildTables()
{
	const int A = cfg.qi_angstepspq, nr = cfg.qi_radialsteps, iters = cfg.qi_iterations;
	if (A < 1) fail("qi_angstepspq must be >= 1");
	if (nr < 1) fail("qi_radialsteps must be >= 1");
	if (cfg.zlut_radialsteps < 0 || cfg.zlut_angularsteps < 0) fail("invalid Z-LUT sampling sizes");
	if (iters < 0) fail("qi_iterations must be >= 0");

	// QI directions (cputrack/cpu_tracker.cpp:294-300): 0.5*pi*(j+0.5)/A with a
	// double 0.5, rounded to float once.
	std::vector<float> qd(2 * (size_t)A);
	for (int j = 0; j < A; j++) {
		float ang = (float)(0.5 * (double)3.141593f * (double)(j + 0.5f) / (double)(float)A);
		qd[2 * j] = cosf(ang);
		qd[2 * j + 1] = sinf(ang);
	}
	upload(d_qi_dirs, qd.data(), qd.size() * 4);
	// each quadrant's sample offsets from the centre, (mx*dx)*r and (my*dy)*r in float as the
	// reference forms them (cputrack/cpu_tracker.cpp:649-659): the sign flips are exact, so the
	// products are the kernel's own; one table for every ROI, read instead of formed per sample
	{
		const float minr = cfg.qi_minradius, maxr = cfg.qi_maxradius;
		const float rstep = (maxr - minr) / (float)nr;
		std::vector<float> offs(8 * (size_t)A * nr);
		for (int q = 0; q < 4; q++) {
			const float mx = (q == 0 || q == 3) ? 1.0f : -1.0f, my = q < 2 ? 1.0f : -1.0f;
			for (int j = 0; j < A; j++)
				for (int i = 0; i < nr; i++) {
					const float r = minr + rstep * (float)i;
					float* o = &offs[2 * (((size_t)q * A + j) * nr + i)];
					o[0] = (mx * qd[2 * j]) * r;
					o[1] = (my * qd[2 * j + 1]) * r;
				}
		}
		upload(d_qi_offs, offs.data(), offs.size() * 4);
	}

	// Angular samples and table stride per iteration (cputrack/cpu_tracker.cpp:314,
	// 356, 642-657): the float count is truncated, clamped to >= 4, and the
	// stride is trunc((float)A / count).
	std::vector<int> steps(std::max(iters, 1)), stride(std::max(iters, 1));
	float angsteps = A / powf(cfg.qi_angstep_factor, (float)iters);
	for (int k = 0; k < iters; k++) {
		int s = (int)angsteps;
		if (s < MIN_RADPROFILE_SMP_COUNT) s = MIN_RADPROFILE_SMP_COUNT;
		float angstepf = (float)A / s;
		steps[k] = s;
		stride[k] = (int)angstepf;
		angsteps *= cfg.qi_angstep_factor;
	}
	upload(d_qi_steps, steps.data(), steps.size() * 4);
	upload(d_qi_stride, stride.data(), stride.size() * 4);

	const std::vector<float> wnd = radial_bin_window(nr);
	upload(d_qi_window, wnd.data(), wnd.size() * 4);

	// FFT of length 2*nr (cputrack/kissfft.h:11-47)
	if (2 * nr > kMaxFFTLen) fail("qi_radialsteps > 256 is not supported");
	qi_fft.build(2 * nr);
	for (int s = 0; s < qi_fft.nst; s++)
		if (qi_fft.radix[s] != 2 && qi_fft.radix[s] != 4) fail("QI FFT length must be a power of two");
	// LT_XCor1D's transform of the profile length and LT_ZLUTAlign's of twice the Z-LUT
	// radial steps, built when the sizes fit; SetLocalizationMode rejects the mode otherwise
	if (cfg.xc1_profileLength >= 2 && cfg.xc1_profileLength <= kMaxFFTLen) xc_fft.build(cfg.xc1_profileLength);
	if (cfg.zlut_radialsteps >= 1 && 2 * cfg.zlut_radialsteps <= kMaxFFTLen) qa_fft.build(2 * cfg.zlut_radialsteps);
	if (cfg.width <= kMaxFFTLen && cfg.height <= kMaxFFTLen) {
		fx_fft.build(cfg.width);
		fy_fft.build(cfg.height);
	}
	// QuadrantAlign's quadrant profiles take all qi_angstepspq directions
	// (cputrack/cpu_tracker.cpp:476, 642-657): at least 4 samples, stride trunc(A / samples)
	qa_steps = A < MIN_RADPROFILE_SMP_COUNT ? MIN_RADPROFILE_SMP_COUNT : A;
	qa_stride = (int)((float)A / qa_steps);

	// Radial-profile directions, cputrack/utils.cpp:278-282
	const int ZA = cfg.zlut_angularsteps;
	std::vector<float> zd(2 * (size_t)std::max(ZA, 1));
	for (int j = 0; j < ZA; j++) {
		float ang = 2 * 3.141593f * j / (float)ZA;
		zd[2 * j] = cosf(ang);
		zd[2 * j + 1] = sinf(ang);
	}
	upload(d_z_dirs, zd.data(), zd.size() * 4);

	LocalizeParams P{};
	FillParams(P, 2, 1, nullptr, nullptr, 0, nullptr, nullptr);
	P.zplanes = 0;
	size_t need = localize_lds_bytes(P);
	if (need + 8 * 1024 > lds_limit)
		fail("ROI too large for the per-workgroup LDS (" + std::to_string(need) + " bytes needed)");
}

void QueuedHIPTracker::AllocSlots()
{
	if (!copyStream) HIP_CHECK(hipStreamCreateWithFlags(&copyStream, hipStreamNonBlocking));
	for (int i = 0; i < numSlots; i++) {
		Slot* s = new Slot();
		HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
		HIP_CHECK(hipEventCreate(&s->ev_start));
		HIP_CHECK(hipEventCreateWithFlags(&s->ev_h2d, hipEventDisableTiming));
		HIP_CHECK(hipEventCreate(&s->ev_k0));
		HIP_CHECK(hipEventCreate(&s->ev_k1));
		// blocking synchronisation: a long wait on a full queue sleeps instead of spinning
		HIP_CHECK(hipEventCreateWithFlags(&s->ev_done, hipEventBlockingSync));
		size_t rb = (size_t)cfg.width * cfg.height * 2;
		s->h_rois.ensure(rb * batchSize);
		s->h_jobs.ensure(sizeof(DevJob) * batchSize);
		s->h_results.ensure(sizeof(DevResult) * batchSize);
		s->d_rois.ensure(rb * batchSize);
		s->d_jobs.ensure(sizeof(DevJob) * batchSize);
		s->d_results.ensure(sizeof(DevResult) * batchSize);
		slots.push_back(s);
	}
}

void QueuedHIPTracker::FreeSlots()
{
	for (Slot* s : slots) {
		if (s->stream) hipStreamDestroy(s->stream);
		hipEventDestroy(s->ev_start);
		hipEventDestroy(s->ev_h2d);
		hipEventDestroy(s->ev_k0);
		hipEventDestroy(s->ev_k1);
		hipEventDestroy(s->ev_done);
		delete s;
	}
	slots.clear();
	filling = nullptr;
	if (copyStream) hipStreamDestroy(copyStream);
	copyStream = nullptr;
}

void QueuedHIPTracker::Warn(const std::string& w)
{
	if (warnings.size() < 8192) warnings += w + "\n";
}

void QueuedHIPTracker::FillParams(LocalizeParams& P, int pdt, int n, const void* src, const int64_t* roi_offset,
                                  int pitch, const DevJob* jobs, DevResult* res)
{
	P = LocalizeParams{};
	P.src = (const uint8_t*)src;
	P.roi_offset = roi_offset;
	P.pitch = pitch;
	P.pdt = pdt;
	P.n = n;
	P.jobs = jobs;
	P.results = res;
	P.width = cfg.width;
	P.height = cfg.height;
	P.mode = localizeMode;
	P.com_bgcorrection = cfg.com_bgcorrection;
	P.gauss_iters = cfg.gauss2D_iterations;
	P.gauss_sigma = cfg.gauss2D_sigma;
	P.qi_iters = cfg.qi_iterations;
	P.qi_nr = cfg.qi_radialsteps;
	P.qi_A = cfg.qi_angstepspq;
	P.qi_minr = cfg.qi_minradius;
	P.qi_maxr = cfg.qi_maxradius;
	P.qi_iter_steps = d_qi_steps.as<int>();
	P.qi_iter_stride = d_qi_stride.as<int>();
	P.qi_dirs = d_qi_dirs.as<float2>();
	P.qi_offs = d_qi_offs.as<float2>();
	P.qi_window = d_qi_window.as<float>();
	P.qi_fft = qi_fft.plan();
	P.xc_len = cfg.xc1_profileLength;
	P.xc_width = cfg.xc1_profileWidth;
	P.xc_iters = cfg.xc1_iterations;
	P.xc_fft = xc_fft.plan();
	P.qa_steps = qa_steps;
	P.qa_stride = qa_stride;
	P.qa_fft = qa_fft.plan();
	P.fx_fft = fx_fft.plan();
	P.fy_fft = fy_fft.plan();
	P.fourier = 0;
	P.z_nb = cfg.zlut_radialsteps;
	P.z_A = cfg.zlut_angularsteps;
	P.z_minr = cfg.zlut_minradius;
	P.z_maxr = cfg.zlut_maxradius;
	P.z_dirs = d_z_dirs.as<float2>();
	P.zluts = zcount > 0 ? d_zlut.as<float>() : nullptr;
	P.zcount = zcount;
	P.zplanes = zplanes;
	P.zcmp = has_zcmp ? d_zcmp.as<float>() : nullptr;
	P.cmp_out = cmp_enabled && zcount > 0 ? d_cmp.as<float>() : nullptr;
	P.calib_offset = has_off ? d_calib_off.as<float>() : nullptr;
	P.calib_gain = has_gain ? d_calib_gain.as<float>() : nullptr;
	// The ROI is staged in LDS as float: measured faster than u16 staging, whose
	// unaligned pixel-pair reads cost more than the occupancy they buy.
	P.lds_pix = 2;
}

// ------------------------------------------------------------- queueing

// A free slot, marked Filling, after harvesting finished batches; nullptr when every slot is busy.
QueuedHIPTracker::Slot* QueuedHIPTracker::TakeFreeSlot()
{
	Harvest();
	for (Slot* s : slots)
		if (s->state == Slot::Free) {
			s->state = Slot::Filling;
			s->n = 0;
			s->frame = false;
			return s;
		}
	return nullptr;
}

// Waits for `ev` with the lock released.  The event is polled for up to spinWait (a batch that
// is about to finish resumes the caller at once) and then waited on with a blocking
// synchronisation, so a long wait does not keep a core busy.  `waiters` keeps the slots (and
// their events) alive meanwhile: code that frees them waits until it is back to zero.  Other
// threads may change `filling` and the slot states while this waits.
void QueuedHIPTracker::WaitEventLocked(std::unique_lock<std::mutex>& lk, hipEvent_t ev)
{
	waiters++;
	const auto spin = spinWait;
	lk.unlock();
	const auto t0 = std::chrono::steady_clock::now();
	hipError_t e;
	while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
		if (std::chrono::steady_clock::now() - t0 > spin) {
			e = hipEventSynchronize(ev);
			break;
		}
		std::this_thread::yield();
	}
	lk.lock();
	if (--waiters == 0) slots_cv.notify_all();
	HIP_CHECK(e);
}

// Queue full: block like the reference (cputrack/QueuedCPUTracker.cpp:437-440) until the
// oldest launched batch completes.  Callers re-examine `filling` afterwards.
void QueuedHIPTracker::WaitOldestLocked(std::unique_lock<std::mutex>& lk)
{
	Slot* oldest = nullptr;
	for (Slot* s : slots)
		if (s->state == Slot::Launched && (!oldest || s->seq < oldest->seq)) oldest = s;
	if (!oldest) fail("no batch slot available");
	WaitEventLocked(lk, oldest->ev_done);
}

// Flushes, waits for every batch, and waits until no thread is blocked on a slot's event, so the
// slots can be freed or remade.  Returns with the lock held, no slot in use and `quiescing` set:
// scheduling threads wait until the caller clears it (EndQuiesceLocked).
void QueuedHIPTracker::QuiesceLocked(std::unique_lock<std::mutex>& lk)
{
	quiescing = true;
	while (true) {
		if (filling) {
			if (filling->n > 0) Launch(filling);
			else filling->state = Slot::Free;
			filling = nullptr;
		}
		WaitAllLocked(lk);
		if (waiters == 0 && !filling) return;
		slots_cv.wait(lk, [&] { return waiters == 0; });
	}
}

void QueuedHIPTracker::EndQuiesceLocked()
{
	quiescing = false;
	slots_cv.notify_all();
}

void QueuedHIPTracker::ScheduleLocalization(void* data, int pitch, QTRK_PixelDataType pdt, const LocalizationJob* jobInfo)
{
	if ((int)pdt < 0 || (int)pdt > 2) fail("invalid pixel data type");
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	// several threads may schedule: a thread that waited for a slot finds `filling` as the
	// others left it
	while (true) {
		if (quiescing) {  // the slots are being remade (SetConfigValue)
			slots_cv.wait(lk, [&] { return !quiescing; });
			continue;
		}
		if (filling && filling->n > 0 && (filling->pdt != (int)pdt || filling->frame)) {
			Launch(filling);
			filling = nullptr;
		}
		if (filling || (filling = TakeFreeSlot())) break;
		WaitOldestLocked(lk);
	}
	Slot* s = filling;
	s->frame = false;
	s->pdt = (int)pdt;
	const size_t row = (size_t)cfg.width * PDT_BytesPerPixel(pdt);
	const size_t rb = row * cfg.height;
	if (s->h_rois.bytes < rb * batchSize) {
		s->h_rois.ensure(rb * batchSize);
		s->d_rois.ensure(rb * batchSize);
	}
	uint8_t* dst = s->h_rois.as<uint8_t>() + rb * s->n;
	const uint8_t* srcp = (const uint8_t*)data;
	// non-temporal stores: the GPU's copy engine, not this CPU, reads the slot next (1.16 against
	// 1.05 M calls/s for plain memcpy, 100x100 u16 ROIs)
	if ((size_t)pitch == row)
		copy_streaming(dst, srcp, rb);
	else
		for (int y = 0; y < cfg.height; y++) copy_streaming(dst + row * y, srcp + (size_t)pitch * y, row);
	memcpy(s->h_jobs.as<DevJob>() + s->n, jobInfo, sizeof(DevJob));
	s->n++;
	if (s->n >= batchSize) {
		Launch(s);
		filling = nullptr;
	}
}

void QueuedHIPTracker::Launch(Slot* s)
{
	const int bpp = PDT_BytesPerPixel((QTRK_PixelDataType)s->pdt);
	const size_t rb = (size_t)cfg.width * cfg.height * bpp;
	LocalizeParams P;
	if (s->frame) {
		// the row blocks are on their way already on the copy stream (ScheduleFrame), ev_start
		// recorded before the first; the rest go now, and the kernel waits for the last
		SendFrameBlocks(s);
		HIP_CHECK(hipStreamWaitEvent(s->stream, s->ev_h2d, 0));
		HIP_CHECK(hipMemcpyAsync(s->d_off.p, s->h_off.p, sizeof(int64_t) * s->n, hipMemcpyHostToDevice, s->stream));
		HIP_CHECK(hipMemcpyAsync(s->d_jobs.p, s->h_jobs.p, sizeof(DevJob) * s->n, hipMemcpyHostToDevice, s->stream));
		FillParams(P, s->pdt, s->n, s->d_rois.p, s->d_off.as<int64_t>(), s->frame_pitch, s->d_jobs.as<DevJob>(),
		           s->d_results.as<DevResult>());
	} else {
		HIP_CHECK(hipEventRecord(s->ev_start, s->stream));
		HIP_CHECK(hipMemcpyAsync(s->d_rois.p, s->h_rois.p, rb * s->n, hipMemcpyHostToDevice, s->stream));
		HIP_CHECK(hipMemcpyAsync(s->d_jobs.p, s->h_jobs.p, sizeof(DevJob) * s->n, hipMemcpyHostToDevice, s->stream));
		FillParams(P, s->pdt, s->n, s->d_rois.p, nullptr, cfg.width * bpp, s->d_jobs.as<DevJob>(),
		           s->d_results.as<DevResult>());
	}
	HIP_CHECK(hipEventRecord(s->ev_k0, s->stream));
	if (matchPrec != kMatchFP32 && SplitMatchUsable().empty()) {
		// match_precision bf16/fp16/fp32_grouped: profile kernel + grouped match (modes it cannot
		// take run fused)
		const int mp = match_kernel_precision(matchPrec);
		const LutMatchTables* tab = MatchTablesLocked(mp, s->stream);
		const int nb = cfg.zlut_radialsteps;
		s->d_split.ensure(SplitScratch::Bytes(batchSize, nb, tab->beads));
		const SplitScratch sc = SplitScratch::Carve(s->d_split.p, batchSize, nb, tab->beads);
		const int mpb = MaxPerBead(s->h_jobs.as<DevJob>(), s->n, tab->beads);
		if (mpb > 0)
			LocalizeSplit(P.src, P.pdt, P.jobs, P.results, s->n, mp, tab, sc, nullptr, s->stream, P.roi_offset,
			              P.pitch, mpb);
		else
			HIP_CHECK(launch_localize(P, s->stream));  // no ROI has a bead in the LUT: nothing to match
	} else {
		HIP_CHECK(launch_localize(P, s->stream));
	}
	HIP_CHECK(hipEventRecord(s->ev_k1, s->stream));
	HIP_CHECK(hipMemcpyAsync(s->h_results.p, s->d_results.p, sizeof(DevResult) * s->n, hipMemcpyDeviceToHost, s->stream));
	HIP_CHECK(hipEventRecord(s->ev_done, s->stream));
	s->seq = ++seqCounter;
	s->state = Slot::Launched;
}

// The slot's row blocks not yet sent go to the GPU in one copy, in schedule order on the one copy
// stream; ev_h2d follows it, and the slot's kernel waits for ev_h2d when it launches.  The slot is
// taken again only after that kernel, so the pinned and device blocks are not overwritten while a
// copy or the kernel still reads them.
void QueuedHIPTracker::SendFrameBlocks(Slot* s)
{
	if (s->frame_sent >= s->frame_bytes) return;
	HIP_CHECK(hipMemcpyAsync(s->d_rois.as<uint8_t>() + s->frame_sent, s->h_rois.as<uint8_t>() + s->frame_sent,
	                         s->frame_bytes - s->frame_sent, hipMemcpyHostToDevice, copyStream));
	HIP_CHECK(hipEventRecord(s->ev_h2d, copyStream));
	s->frame_sent = s->frame_bytes;
}

// QueuedTracker::ScheduleFrame (cputrack/QueuedTracker.cpp:91-111) with the
// extraction on the GPU: ROIs outside the frame are skipped by the reference's
// rule and job.zlutIndex = i + jobInfo->zlutIndex as there; then the frame
// rows the ROIs cover are copied once into the filling batch slot, after the
// row blocks of earlier frames of the same pixel type and pitch, and on to the
// GPU at once; the kernel reads each ROI at its byte offset in the slot.  A slot launches when it holds
// batch_size ROIs, when its row blocks would outgrow it, or on Flush, as
// scheduled ROIs do.  Returns the number of ROIs queued.
int QueuedHIPTracker::ScheduleFrame(void* imgptr, int pitch, int width, int height, ROIPosition* positions,
                                    int numROI, QTRK_PixelDataType pdt, const LocalizationJob* jobInfo)
{
	if ((int)pdt < 0 || (int)pdt > 2) fail("invalid pixel data type");
	if (numROI <= 0) return 0;
	if (!imgptr || !positions || !jobInfo) fail("null argument");
	const int bpp = PDT_BytesPerPixel(pdt);
	if (pitch < width * bpp) fail("pitch is smaller than width * bytes per pixel");
	std::vector<int> sel;
	for (int i = 0; i < numROI; i++) {
		const ROIPosition& p = positions[i];
		if (p.x < 0 || p.y < 0 || p.x + cfg.width > width || p.y + cfg.height > height) continue;
		sel.push_back(i);
	}
	if (sel.empty()) return 0;
	const uint8_t* img = (const uint8_t*)imgptr;
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	if (!copier) copier.reset(new CopyPool(frameCopyThreads - 1, true));
	size_t c0 = 0;
	while (c0 < sel.size()) {
		if (quiescing) {  // the slots are being remade (SetConfigValue)
			slots_cv.wait(lk, [&] { return !quiescing; });
			continue;
		}
		// ROIs queued before the frame (or another thread's) go first, in their own launch
		if (filling && filling->n > 0 &&
		    !(filling->frame && filling->pdt == (int)pdt && filling->frame_pitch == pitch)) {
			Launch(filling);
			filling = nullptr;
		}
		if (!filling) {
			Slot* f = TakeFreeSlot();
			if (!f) {
				WaitOldestLocked(lk);  // `filling` may be another thread's frame slot afterwards
				continue;
			}
			filling = f;
		}
		if (filling->n == 0) {
			filling->frame = true;
			filling->pdt = (int)pdt;
			filling->frame_pitch = pitch;
			filling->frame_bytes = 0;
			filling->frame_sent = 0;
		}
		Slot* s = filling;
		const size_t c1 = std::min(sel.size(), c0 + (size_t)(batchSize - s->n));
		int y0 = height, y1 = 0;
		for (size_t k = c0; k < c1; k++) {
			y0 = std::min(y0, positions[sel[k]].y);
			y1 = std::max(y1, positions[sel[k]].y + cfg.height);
		}
		// rows y0 .. y1-1; the last one only as far as the frame's width
		const size_t bytes = (size_t)pitch * (y1 - y0 - 1) + (size_t)width * bpp;
		const size_t at = (s->frame_bytes + kFrameBlockAlign - 1) & ~(kFrameBlockAlign - 1);
		if (s->n > 0 && at + bytes > std::min(s->h_rois.bytes, s->d_rois.bytes)) {
			Launch(s);
			filling = nullptr;
			continue;
		}
		if (s->n == 0) {
			// room for as many blocks like this one as make a batch, within kFrameSlotMaxBytes
			const size_t per = c1 - c0, blocks = ((size_t)batchSize + per - 1) / per;
			const size_t block = (bytes + kFrameBlockAlign - 1) & ~(kFrameBlockAlign - 1);
			const size_t cap = std::max(bytes, std::min(blocks * block, kFrameSlotMaxBytes));
			s->h_rois.ensure(cap);
			s->d_rois.ensure(cap);
			s->h_off.ensure(sizeof(int64_t) * (size_t)batchSize);
			s->d_off.ensure(sizeof(int64_t) * (size_t)batchSize);
		}
		if (s->n == 0) HIP_CHECK(hipEventRecord(s->ev_start, copyStream));
		copier->Copy(s->h_rois.as<uint8_t>() + at, img + (size_t)pitch * y0, bytes);
		s->frame_bytes = at + bytes;
		if (s->frame_bytes - s->frame_sent >= kFrameSendBytes) SendFrameBlocks(s);
		int64_t* off = s->h_off.as<int64_t>() + s->n;
		DevJob* jobs = s->h_jobs.as<DevJob>() + s->n;
		for (size_t k = c0; k < c1; k++) {
			const int i = sel[k];
			off[k - c0] = (int64_t)(at + (size_t)pitch * (positions[i].y - y0) + (size_t)positions[i].x * bpp);
			LocalizationJob job = *jobInfo;
			job.zlutIndex = i + jobInfo->zlutIndex;
			memcpy(&jobs[k - c0], &job, sizeof(DevJob));
		}
		s->n += (int)(c1 - c0);
		if (s->n >= batchSize) {
			Launch(s);
			filling = nullptr;
		}
		c0 = c1;
	}
	return (int)sel.size();
}

void QueuedHIPTracker::Harvest()
{
	std::vector<Slot*> done;
	for (Slot* s : slots)
		if (s->state == Slot::Launched && hipEventQuery(s->ev_done) == hipSuccess) done.push_back(s);
	std::sort(done.begin(), done.end(), [](Slot* a, Slot* b) { return a->seq < b->seq; });
	for (Slot* s : done) {
		float kms = 0, all = 0;
		hipEventElapsedTime(&kms, s->ev_k0, s->ev_k1);
		hipEventElapsedTime(&all, s->ev_start, s->ev_done);
		t_kernel_ms += kms;
		t_copy_ms += all - kms;
		batchesDone++;
		roisDone += s->n;
		const DevResult* r = s->h_results.as<DevResult>();
		const bool bias = localizeMode & LT_LocalizeZ;
		{
			std::lock_guard<std::mutex> g(res_mtx);
			for (int i = 0; i < s->n; i++) {
				LocalizationResult lr;
				memcpy(&lr, &r[i], sizeof(lr));
				if (bias) ApplyZBias(&lr, 1);
				results.push_back(lr);
			}
		}
		s->n = 0;
		s->state = Slot::Free;
	}
}

void QueuedHIPTracker::ApplyZBias(LocalizationResult* r, int n)
{
	if (!HasZLUTBiasCorrection()) return;
	for (int i = 0; i < n; i++) r[i].pos.z = ZLUTBiasCorrection(r[i].pos.z, zplanes, r[i].job.zlutIndex);
}

void QueuedHIPTracker::ComputeZBiasCorrection(int biasPlanes, float* result, int smpPerPixel, bool spline)
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	WaitAllLocked(lk);
	if (zcount <= 0 || zplanes <= 0) fail("no radial Z-LUT");
	if (biasPlanes < 1) fail("bias_planes must be >= 1");
	if (smpPerPixel < 1) fail("smpPerPixel must be >= 1");
	if (zplanes < (spline ? 4 : 2)) fail("too few Z-LUT planes for the interpolation");
	if (cfg.zlut_radialsteps < 2 || !(cfg.zlut_maxradius > cfg.zlut_minradius)) fail("invalid Z-LUT radii");
	const int W = cfg.width, H = cfg.height, nb = cfg.zlut_radialsteps;
	const int total = zcount * biasPlanes;
	// renders staged in chunks of at most 512 MB
	const int chunk = (int)std::max<size_t>(1, std::min<size_t>(total, (512u << 20) / ((size_t)W * H * 4)));

	// The reference localizes each render with a plain CPUTracker: 2 QI iterations with the QI
	// bin window, COM without background correction, the normalised profile compared with the
	// Z-LUT bin window as radial weights (cputrack/QueuedTracker.cpp:155-181).
	const int A = cfg.qi_angstepspq, iters = 2;
	std::vector<int> steps(iters), stride(iters);
	float angsteps = A / powf(cfg.qi_angstep_factor, (float)iters);
	for (int k = 0; k < iters; k++) {
		int st = (int)angsteps;
		if (st < MIN_RADPROFILE_SMP_COUNT) st = MIN_RADPROFILE_SMP_COUNT;
		steps[k] = st;
		stride[k] = (int)((float)A / st);
		angsteps *= cfg.qi_angstep_factor;
	}
	const std::vector<float> zw = radial_bin_window(nb);
	DeviceBuffer d_steps, d_stride, d_zw, d_theta, d_idx, d_img, d_jobs, d_res;
	upload(d_steps, steps.data(), steps.size() * 4);
	upload(d_stride, stride.data(), stride.size() * 4);
	upload(d_zw, zw.data(), zw.size() * 4);
	d_theta.ensure((size_t)chunk * kModelParams * 4);
	d_idx.ensure((size_t)chunk * 4);
	d_img.ensure((size_t)chunk * W * H * 4);
	d_jobs.ensure((size_t)chunk * sizeof(DevJob));
	d_res.ensure((size_t)chunk * sizeof(DevResult));
	std::vector<float> theta((size_t)chunk * kModelParams), zs(chunk), table(total);
	std::vector<int> idx(chunk);
	std::vector<DevJob> jobs(chunk);
	std::vector<DevResult> res(chunk);
	for (int c0 = 0; c0 < total; c0 += chunk) {
		const int n = std::min(chunk, total - c0);
		for (int i = 0; i < n; i++) {
			const int bead = (c0 + i) / biasPlanes, plane = (c0 + i) % biasPlanes;
			zs[i] = (float)plane / (float)biasPlanes * (float)zplanes;
			float* th = &theta[(size_t)i * kModelParams];
			th[0] = (float)(W / 2);
			th[1] = (float)(H / 2);
			th[2] = zs[i];
			th[3] = 1.0f;  // mu = 0 + 1 * g: the rendered LUT image itself
			th[4] = 0.0f;
			idx[i] = bead;
			jobs[i] = DevJob{ 0u, 0u, bead, 0.0f, 0.0f, 0.0f };
		}
		HIP_CHECK(hipMemcpyAsync(d_theta.p, theta.data(), (size_t)n * kModelParams * 4, hipMemcpyHostToDevice,
		                         aux_stream()));
		HIP_CHECK(hipMemcpyAsync(d_idx.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, aux_stream()));
		HIP_CHECK(hipMemcpyAsync(d_jobs.p, jobs.data(), (size_t)n * sizeof(DevJob), hipMemcpyHostToDevice,
		                         aux_stream()));
		LutModelParams m{};
		m.luts = d_zlut.as<float>();
		m.count = zcount;
		m.planes = zplanes;
		m.steps = nb;
		m.minr = cfg.zlut_minradius;
		m.maxr = cfg.zlut_maxradius;
		m.width = W;
		m.height = H;
		m.ovs = smpPerPixel;
		m.spline = spline ? 1 : 0;
		m.n = n;
		m.theta = d_theta.as<float>();
		m.lut_index = d_idx.as<int>();
		m.lut_index_stride = 1;
		m.mu = d_img.as<float>();
		HIP_CHECK(launch_lutmodel_forward(m, aux_stream()));
		LocalizeParams P;
		FillParams(P, 2, n, d_img.p, nullptr, W * 4, d_jobs.as<DevJob>(), d_res.as<DevResult>());
		P.mode = kModeQI | kModeLocalizeZ | kModeNormalizeProfile;
		P.com_bgcorrection = 0.0f;
		P.qi_iters = iters;
		P.qi_iter_steps = d_steps.as<int>();
		P.qi_iter_stride = d_stride.as<int>();
		P.zcmp = d_zw.as<float>();
		P.cmp_out = nullptr;
		P.calib_offset = nullptr;
		P.calib_gain = nullptr;
		HIP_CHECK(launch_localize(P, aux_stream()));
		HIP_CHECK(hipMemcpyAsync(res.data(), d_res.p, (size_t)n * sizeof(DevResult), hipMemcpyDeviceToHost,
		                         aux_stream()));
		HIP_CHECK(hipStreamSynchronize(aux_stream()));
		for (int i = 0; i < n; i++) table[c0 + i] = res[i].z - zs[i];
	}
	SetZLUTBiasCorrection(table.data(), biasPlanes, zcount);
	if (result) std::copy(table.begin(), table.end(), result);
}

void QueuedHIPTracker::WaitAllLocked(std::unique_lock<std::mutex>& lk)
{
	while (true) {
		Harvest();
		hipEvent_t ev = nullptr;
		for (Slot* s : slots)
			if (s->state == Slot::Launched) { ev = s->ev_done; break; }
		if (!ev) return;
		WaitEventLocked(lk, ev);
	}
}

void QueuedHIPTracker::WaitLaunched()
{
	std::unique_lock<std::mutex> lk(mtx);
	WaitAllLocked(lk);
}

int QueuedHIPTracker::QueueLengthLocked() const
{
	int q = 0;
	for (const Slot* s : slots)
		if (s->state != Slot::Free) q += s->n;
	return q;
}

void QueuedHIPTracker::Flush()
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	if (filling && filling->n > 0) {
		Launch(filling);
		filling = nullptr;
	}
	Harvest();
}

int QueuedHIPTracker::GetQueueLength(int* maxQueueLen)
{
	std::unique_lock<std::mutex> lk(mtx);
	Harvest();
	if (maxQueueLen) *maxQueueLen = numSlots * batchSize;
	return QueueLengthLocked();
}

bool QueuedHIPTracker::IsIdle() { return GetQueueLength(nullptr) == 0; }

int QueuedHIPTracker::GetResultCount()
{
	{
		std::unique_lock<std::mutex> lk(mtx);
		Harvest();
	}
	std::lock_guard<std::mutex> g(res_mtx);
	return (int)results.size();
}

// LIFO like the CPU tracker (cputrack/QueuedCPUTracker.cpp:458-469).
int QueuedHIPTracker::FetchResults(LocalizationResult* out, int maxResults)
{
	{
		std::unique_lock<std::mutex> lk(mtx);
		Harvest();
	}
	std::lock_guard<std::mutex> g(res_mtx);
	int n = 0;
	while (n < maxResults && !results.empty()) {
		out[n++] = results.back();
		results.pop_back();
	}
	return n;
}

void QueuedHIPTracker::ClearResults()
{
	std::lock_guard<std::mutex> g(res_mtx);
	results.clear();
}

// Flush, wait for idle, then switch (cudatrack/QueuedCUDATracker.cu:420-428).
void QueuedHIPTracker::SetLocalizationMode(LocMode_t mode)
{
	const int supported = LT_QI | LT_Gaussian2D | LT_LocalizeZ | LT_NormalizeProfile | LT_ClearFirstFourPixels |
	                      LT_XCor1D | LT_ZLUTAlign | LT_LocalizeZWeighted | LT_FourierLUT;
	if (mode & ~supported) {
		std::ostringstream os;
		os << "localization mode bits 0x" << std::hex << (mode & ~supported) << " are not implemented";
		fail(os.str());
	}
	if ((mode & LT_XCor1D) && xc_fft.n == 0)
		fail("LT_XCor1D needs xc1_profileLength in 2.." + std::to_string(kMaxFFTLen));
	if ((mode & LT_XCor1D) && (cfg.xc1_iterations < 0 || cfg.xc1_profileWidth < 0))
		fail("LT_XCor1D needs xc1_iterations >= 0 and xc1_profileWidth >= 0");
	if ((mode & LT_ZLUTAlign) && (mode & LT_LocalizeZ) && qa_fft.n == 0)
		fail("LT_ZLUTAlign needs zlut_radialsteps in 1.." + std::to_string(kMaxFFTLen / 2));
	if ((mode & LT_FourierLUT) && (mode & LT_LocalizeZ) && fx_fft.n == 0)
		fail("LT_FourierLUT needs a ROI of at most " + std::to_string(kMaxFFTLen) + " pixels a side");
	{
		LocalizeParams P{};
		FillParams(P, 2, 1, nullptr, nullptr, 0, nullptr, nullptr);
		P.mode = mode;
		P.zplanes = zplanes;
		const size_t need = localize_lds_bytes(P);
		if (need > lds_limit)
			fail("mode needs " + std::to_string(need) + " bytes of LDS per ROI, more than the device's " +
			     std::to_string(lds_limit));
	}
	Flush();
	std::unique_lock<std::mutex> lk(mtx);
	WaitAllLocked(lk);
	localizeMode = mode;
}

// ------------------------------------------------------------- Z-LUT

void QueuedHIPTracker::SetRadialZLUT(float* data, int count, int planes)
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	WaitAllLocked(lk);
	const int res = cfg.zlut_radialsteps;
	SetZLUTBiasCorrection(nullptr, 0, 0);  // the reference drops its bias table here (:368-369)
	if (count < 0 || planes < 0) fail("negative LUT size");
	{
		LocalizeParams P;
		FillParams(P, 2, 1, nullptr, nullptr, 0, nullptr, nullptr);
		P.zplanes = planes;
		if (localize_lds_bytes(P) > lds_limit) fail("too many LUT planes for the per-workgroup LDS");
	}
	const size_t total = (size_t)count * planes * res;
	if (total > 0) {
		h_zlut.assign(total, 0.0f);
		if (data) memcpy(h_zlut.data(), data, total * 4);
		zcount = count;
		zplanes = planes;
		upload(d_zlut, h_zlut.data(), total * 4);
		d_cmp.ensure((size_t)count * planes * 4);
		HIP_CHECK(hipMemset(d_cmp.p, 0, (size_t)count * planes * 4));
	} else {
		h_zlut.clear();
		zcount = zplanes = 0;
	}
	h_zlut_stale = false;
	lutVersion++;
}

void QueuedHIPTracker::GetRadialZLUT(float* dst)
{
	std::unique_lock<std::mutex> lk(mtx);
	if (h_zlut_stale && !h_zlut.empty()) {
		HIP_CHECK(hipMemcpy(h_zlut.data(), d_zlut.p, h_zlut.size() * 4, hipMemcpyDeviceToHost));
		h_zlut_stale = false;
	}
	if (!h_zlut.empty()) memcpy(dst, h_zlut.data(), h_zlut.size() * 4);
}

void QueuedHIPTracker::GetRadialZLUTSize(int& count, int& planes, int& radialsteps)
{
	count = zcount;
	planes = zplanes;
	radialsteps = cfg.zlut_radialsteps;
}

void QueuedHIPTracker::SetRadialWeights(float* w)
{
	std::unique_lock<std::mutex> lk(mtx);
	WaitAllLocked(lk);
	if (w) {
		upload(d_zcmp, w, (size_t)cfg.zlut_radialsteps * 4);
		has_zcmp = true;
	} else
		has_zcmp = false;
	lutVersion++;
}

void QueuedHIPTracker::EnableRadialZLUTCompareProfile(bool enabled)
{
	std::unique_lock<std::mutex> lk(mtx);
	cmp_enabled = enabled;
}

void QueuedHIPTracker::GetRadialZLUTCompareProfile(float* dst)
{
	std::unique_lock<std::mutex> lk(mtx);
	WaitAllLocked(lk);
	if (zcount > 0 && d_cmp.p)
		HIP_CHECK(hipMemcpy(dst, d_cmp.p, (size_t)zcount * zplanes * 4, hipMemcpyDeviceToHost));
}

// Per-bead calibration, cputrack/QueuedCPUTracker.cpp:165-196: only with a
// LUT loaded; an image is taken only while none is set, and NULL removes it.
void QueuedHIPTracker::SetPixelCalibrationImages(float* offset, float* gain)
{
	std::unique_lock<std::mutex> lk(mtx);
	WaitAllLocked(lk);
	if (zcount <= 0) return;
	const size_t bytes = (size_t)cfg.width * cfg.height * zcount * 4;
	if (!has_gain && gain) { upload(d_calib_gain, gain, bytes); has_gain = true; }
	else if (has_gain && !gain) { d_calib_gain.release(); has_gain = false; }
	if (!has_off && offset) { upload(d_calib_off, offset, bytes); has_off = true; }
	else if (has_off && !offset) { d_calib_off.release(); has_off = false; }
}

// Stored but, as in the CPU tracker, not applied (cputrack/QueuedCPUTracker.cpp:266).
void QueuedHIPTracker::SetPixelCalibrationFactors(float offsetFactor, float gainFactor)
{
	std::unique_lock<std::mutex> lk(mtx);
	gc_offsetFactor = offsetFactor;
	gc_gainFactor = gainFactor;
}

void QueuedHIPTracker::BeginLUT(uint flags)
{
	std::unique_lock<std::mutex> lk(mtx);
	buildFlags = flags;
}

// QueuedCPUTracker::BuildLUT (cputrack/QueuedCPUTracker.cpp:580-646) on the
// GPU: image i of the stack belongs to bead i; COM -> QI (no window) or the
// known position, radial profile, added into plane `plane` of bead i's LUT.
void QueuedHIPTracker::BuildLUT(void* data, int pitch, QTRK_PixelDataType pdt, int plane, vector2f* known_pos)
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	WaitAllLocked(lk);
	const bool ilut = (buildFlags & BUILDLUT_IMAGELUT) && ilut_dims[2] * ilut_dims[3] > 0;
	if (ilut && (ilut_dims[0] < zcount || plane >= ilut_dims[1]))
		fail("BUILDLUT_IMAGELUT: the image LUT has fewer beads or planes than the build needs");
	if ((buildFlags & BUILDLUT_FOURIER) && fx_fft.n == 0)
		fail("BUILDLUT_FOURIER needs a ROI of at most " + std::to_string(kMaxFFTLen) + " pixels a side");
	if (zcount <= 0) fail("BuildLUT needs a LUT allocated with SetRadialZLUT");
	if (plane < 0 || plane >= zplanes) fail("BuildLUT plane out of range");
	const int bpp = PDT_BytesPerPixel(pdt);
	const size_t row = (size_t)cfg.width * bpp, rb = row * cfg.height;
	h_build_rois.ensure(rb * zcount);
	for (int i = 0; i < zcount; i++)
		for (int y = 0; y < cfg.height; y++)
			memcpy(h_build_rois.as<uint8_t>() + rb * i + row * y,
			       (const uint8_t*)data + (size_t)pitch * cfg.height * i + (size_t)pitch * y, row);
	d_build_rois.ensure(rb * zcount);
	HIP_CHECK(hipMemcpyAsync(d_build_rois.p, h_build_rois.p, rb * zcount, hipMemcpyHostToDevice, aux_stream()));
	LocalizeParams P;
	FillParams(P, (int)pdt, zcount, d_build_rois.p, nullptr, (int)row, nullptr, nullptr);
	P.mode = (buildFlags & BUILDLUT_NORMALIZE) ? LT_NormalizeProfile : 0;
	P.build = 1;
	P.build_plane = plane;
	P.build_lut = d_zlut.as<float>();
	P.zcmp = nullptr;
	P.cmp_out = nullptr;
	// BUILDLUT_FOURIER: the profile of each image's power spectrum (cputrack/QueuedCPUTracker.cpp:630-637)
	P.fourier = (buildFlags & BUILDLUT_FOURIER) ? 1 : 0;
	if (localize_lds_bytes(P) > lds_limit) fail("BUILDLUT_FOURIER: the ROI's spectrum does not fit the LDS");
	if (known_pos) {
		upload(d_build_pos, known_pos, sizeof(vector2f) * zcount);
		P.known_pos = d_build_pos.as<float>();
	}
	if (ilut) {
		P.image_lut = d_image_lut.as<float>();
		P.ilut_planes = ilut_dims[1];
		P.ilut_h = ilut_dims[2];
		P.ilut_w = ilut_dims[3];
	}
	HIP_CHECK(launch_localize(P, aux_stream()));
	HIP_CHECK(hipStreamSynchronize(aux_stream()));
	h_zlut_stale = true;
	lutVersion++;
}

// QueuedCPUTracker::GetImageZLUTSize / GetImageZLUT / SetImageZLUT
// (cputrack/QueuedCPUTracker.cpp:519-531, 548-570): dims = [beads, planes, height,
// width]; a null src gives a zeroed image LUT; the radial LUT is set alongside with
// [beads][planes] of radial_zlut (null: zeroed).
void QueuedHIPTracker::GetImageZLUTSize(int* dims)
{
	std::unique_lock<std::mutex> lk(mtx);
	for (int i = 0; i < 4; i++) dims[i] = ilut_dims[i];
}

void QueuedHIPTracker::GetImageZLUT(float* dst)
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	WaitAllLocked(lk);
	const size_t n = (size_t)ilut_dims[0] * ilut_dims[1] * ilut_dims[2] * ilut_dims[3];
	if (n && d_image_lut.p) HIP_CHECK(hipMemcpy(dst, d_image_lut.p, n * 4, hipMemcpyDeviceToHost));
}

bool QueuedHIPTracker::SetImageZLUT(float* src, float* radial_zlut, int* dims)
{
	if (!dims) fail("null dims");
	for (int i = 0; i < 4; i++)
		if (dims[i] < 0) fail("image LUT dimensions must be >= 0");
	{
		std::unique_lock<std::mutex> lk(mtx);
		HIP_CHECK(hipSetDevice(dev));
		WaitAllLocked(lk);
		const size_t n = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
		d_image_lut.release();
		ilut_diff = false;
		for (int i = 0; i < 4; i++) ilut_dims[i] = dims[i];
		if (n) {
			d_image_lut.ensure(n * 4);
			if (src) HIP_CHECK(hipMemcpy(d_image_lut.p, src, n * 4, hipMemcpyHostToDevice));
			else HIP_CHECK(hipMemset(d_image_lut.p, 0, n * 4));
		}
	}
	SetRadialZLUT(radial_zlut, dims[0], dims[1]);
	return true;
}

bool QueuedHIPTracker::GetImageZLUTDerivatives(float* dz, float* dz2)
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	WaitAllLocked(lk);
	if (!ilut_diff) return false;
	const size_t n = (size_t)ilut_dims[0] * ilut_dims[1] * ilut_dims[2] * ilut_dims[3];
	if (dz) HIP_CHECK(hipMemcpy(dz, d_image_lut_dz.p, n * 4, hipMemcpyDeviceToHost));
	if (dz2) HIP_CHECK(hipMemcpy(dz2, d_image_lut_dz2.p, n * 4, hipMemcpyDeviceToHost));
	return true;
}

// Normalizes every plane profile (cputrack/QueuedCPUTracker.cpp:648-659) and computes the image
// LUT's first and second differences across planes (:661-696; read back with
// GetImageZLUTDerivatives).
void QueuedHIPTracker::FinalizeLUT()
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	WaitAllLocked(lk);
	const size_t npx = (size_t)ilut_dims[2] * ilut_dims[3];
	ilut_diff = false;
	if (npx > 0 && ilut_dims[0] > 0 && ilut_dims[1] >= 3 && d_image_lut.p) {
		const size_t n = (size_t)ilut_dims[0] * ilut_dims[1] * npx;
		d_image_lut_dz.ensure(n * 4);
		d_image_lut_dz2.ensure(n * 4);
		HIP_CHECK(launch_image_lut_diff(d_image_lut.as<float>(), d_image_lut_dz.as<float>(),
		                                d_image_lut_dz2.as<float>(), ilut_dims[0], ilut_dims[1], (int)npx, aux_stream()));
		HIP_CHECK(hipStreamSynchronize(aux_stream()));
		ilut_diff = true;
	}
	if (zcount <= 0) return;
	HIP_CHECK(launch_normalize_rows(d_zlut.as<float>(), zcount * zplanes, cfg.zlut_radialsteps, aux_stream()));
	HIP_CHECK(hipStreamSynchronize(aux_stream()));
	HIP_CHECK(hipMemcpy(h_zlut.data(), d_zlut.p, h_zlut.size() * 4, hipMemcpyDeviceToHost));
	h_zlut_stale = false;
	lutVersion++;
}

// ------------------------------------------------- split localization

const LutMatchTables* QueuedHIPTracker::PrepareMatchTables(int prec)
{
	std::unique_lock<std::mutex> lk(mtx);
	HIP_CHECK(hipSetDevice(dev));
	if (prec < kMatchFP32 || prec > kMatchFP16) fail("unknown match precision");
	if (zcount <= 0 || zplanes <= 0) fail("no radial Z-LUT");
	WaitAllLocked(lk);
	const LutMatchTables* t = MatchTablesLocked(prec, aux_stream());
	HIP_CHECK(hipStreamSynchronize(aux_stream()));
	return t;
}

const LutMatchTables* QueuedHIPTracker::MatchTablesLocked(int prec, hipStream_t stream)
{
	if (!matchTab || matchTabVersion != lutVersion || matchTab->beads != zcount || matchTab->planes != zplanes) {
		if (!matchTab || matchTab->beads != zcount || matchTab->planes != zplanes)
			matchTab.reset(new LutMatchTables(zcount, zplanes, cfg.zlut_radialsteps));
		matchTab->set(d_zlut.as<float>(), has_zcmp ? d_zcmp.as<float>() : nullptr, true, stream);
		matchTabVersion = lutVersion;
	}
	matchTab->prepare(prec, stream);
	return matchTab.get();
}

const LutMatchTables* QueuedHIPTracker::CurrentMatchTables(int prec) const
{
	std::unique_lock<std::mutex> lk(mtx);
	if (!matchTab || matchTabVersion != lutVersion || matchTab->beads != zcount || matchTab->planes != zplanes)
		return nullptr;
	if (prec != kMatchFP32 && !(prec >= 0 && prec < 3 && matchTab->prepared[prec])) return nullptr;
	return matchTab.get();
}

std::string QueuedHIPTracker::SplitMatchUsable() const
{
	if (!(localizeMode & LT_LocalizeZ)) return "the localization mode has no LT_LocalizeZ";
	if (localizeMode & (LT_ZLUTAlign | LT_LocalizeZWeighted | LT_FourierLUT))
		return "LT_ZLUTAlign, LT_LocalizeZWeighted and LT_FourierLUT need the per-ROI compare";
	if (zcount <= 0 || zplanes <= 0) return "no radial Z-LUT";
	return "";
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t QueuedHIPTracker::SplitScratch::Bytes(int n, int nb, int beads)
{
	return align256((size_t)n * nb * sizeof(float)) + align256((size_t)beads * sizeof(int)) +
	       align256((size_t)(beads + 1) * sizeof(int)) + align256((size_t)n * sizeof(int));
}

QueuedHIPTracker::SplitScratch QueuedHIPTracker::SplitScratch::Carve(void* base, int n, int nb, int beads)
{
	SplitScratch s;
	char* p = (char*)base;
	s.prof = (float*)p;
	p += align256((size_t)n * nb * sizeof(float));
	s.counts = (int*)p;
	p += align256((size_t)beads * sizeof(int));
	s.start = (int*)p;
	p += align256((size_t)(beads + 1) * sizeof(int));
	s.rowidx = (int*)p;
	return s;
}

void QueuedHIPTracker::LocalizeSplit(const void* rois, int pdt, const DevJob* jobs, DevResult* results, int n,
                                     int prec, const LutMatchTables* tab, const SplitScratch& s, int* best_out,
                                     hipStream_t stream, const int64_t* roi_offset, int pitch, int max_per_bead)
{
	if (n <= 0) return;
	LocalizeParams P;
	FillParams(P, pdt, n, rois, roi_offset, pitch > 0 ? pitch : cfg.width * PDT_BytesPerPixel((QTRK_PixelDataType)pdt),
	           jobs, results);
	P.zluts = nullptr;  // profiles only: z stays 0 (the fused kernel's value for ROIs without a LUT)
	P.prof_out = s.prof;
	P.best_out = best_out;  // -1 here; the match writes the best plane of every grouped ROI
	P.cmp_out = nullptr;
	HIP_CHECK(launch_localize(P, stream));
	const int beads = tab->beads;
	HIP_CHECK(launch_group_by_bead((const int*)((const char*)jobs + offsetof(DevJob, zlutIndex)),
	                               (int)(sizeof(DevJob) / sizeof(int)), n, beads, s.counts, s.start, s.rowidx,
	                               stream));
	LutMatchParams p = tab->params(prec);
	p.prof = s.prof;
	p.start = s.start;
	p.rowidx = s.rowidx;
	p.z = (float*)((char*)results + offsetof(DevResult, z));
	p.zstride = (int)(sizeof(DevResult) / sizeof(float));
	p.best = best_out;
	p.cand = nullptr;
	if (max_per_bead > 0)
		HIP_CHECK(launch_lutmatch(p, prec, max_per_bead, stream, 0));
	else  // workgroups for the average bead; beads with more profiles loop over their tiles
		HIP_CHECK(launch_lutmatch(p, prec, std::max(1, (n + beads - 1) / beads), stream, 0, true));
}

int QueuedHIPTracker::MaxPerBead(const DevJob* jobs, int n, int beads)
{
	std::vector<int> c((size_t)std::max(beads, 0), 0);
	int m = 0;
	for (int i = 0; i < n; i++) {
		const int b = jobs[i].zlutIndex;
		if (b >= 0 && b < beads) m = std::max(m, ++c[b]);
	}
	return m;
}

// ------------------------------------------------------------- config

void QueuedHIPTracker::SetConfigValue(std::string name, std::string value)
{
	if (name == "batch_size" || name == "num_streams") {
		int v = atoi(value.c_str());
		if (v < 1) fail(name + " must be >= 1");
		std::unique_lock<std::mutex> lk(mtx);
		HIP_CHECK(hipSetDevice(dev));
		struct End {
			QueuedHIPTracker* t;
			~End() { t->EndQuiesceLocked(); }
		} end{ this };
		QuiesceLocked(lk);
		FreeSlots();
		if (name == "batch_size") batchSize = v;
		else numSlots = v;
		AllocSlots();
	} else if (name == "match_precision") {
		int v = value == "fp32"           ? kMatchFP32
		        : value == "bf16"         ? kMatchBF16
		        : value == "fp16"         ? kMatchFP16
		        : value == "fp32_grouped" ? kMatchFP32Grouped
		                                  : -1;
		if (v < 0) fail("match_precision must be fp32, bf16, fp16 or fp32_grouped");
		std::unique_lock<std::mutex> lk(mtx);
		HIP_CHECK(hipSetDevice(dev));
		if (filling && filling->n > 0) {  // ROIs queued before go with the old setting
			Launch(filling);
			filling = nullptr;
		}
		WaitAllLocked(lk);
		matchPrec = v;
	} else if (name == "queue_spin_us") {
		int v = atoi(value.c_str());
		if (v < 0 || v > 1000000) fail("queue_spin_us must be in 0..1000000");
		std::unique_lock<std::mutex> lk(mtx);
		spinWait = std::chrono::microseconds(v);
	} else if (name == "frame_copy_threads") {
		int v = atoi(value.c_str());
		if (v < 1 || v > 64) fail("frame_copy_threads must be in 1..64");
		std::unique_lock<std::mutex> lk(mtx);
		frameCopyThreads = v;
		copier.reset();  // remade with the new size on the next frame
	}
}

QueuedTracker::ConfigValueMap QueuedHIPTracker::GetConfigValues()
{
	ConfigValueMap m;
	m["batch_size"] = std::to_string(batchSize);
	m["num_streams"] = std::to_string(numSlots);
	m["frame_copy_threads"] = std::to_string(frameCopyThreads);
	m["queue_spin_us"] = std::to_string((long long)spinWait.count());
	m["match_precision"] = matchPrec == kMatchBF16          ? "bf16"
	                       : matchPrec == kMatchFP16        ? "fp16"
	                       : matchPrec == kMatchFP32Grouped ? "fp32_grouped"
	                                                        : "fp32";
	return m;
}

std::string QueuedHIPTracker::GetProfileReport()
{
	std::unique_lock<std::mutex> lk(mtx);
	std::ostringstream os;
	double f = batchesDone ? 1.0 / batchesDone : 0.0;
	os << "MI355X tracker: " << batchesDone << " batches (" << roisDone << " ROIs), batch size " << batchSize
	   << ", " << numSlots << " streams\n";
	os << "kernel ms/batch: " << t_kernel_ms * f << "\ncopies+queue ms/batch: " << t_copy_ms * f << "\n";
	return os.str();
}

std::string QueuedHIPTracker::GetWarnings()
{
	std::unique_lock<std::mutex> lk(mtx);
	return warnings;
}

}  // namespace qtrk
