// lutmatch_host.cpp -- the QTrkLUTMatch C-ABI object: device buffers, the
// one-time LUT conversion per precision, and launches on the object's stream.
#include "errors.hpp"
#include "hip_tracker.hpp"
#include "lutmatch.hpp"
#include "qtrk_amd.h"

#include <memory>
#include <vector>

using namespace qtrk;

LutMatchTables::LutMatchTables(int beads_, int planes_, int res_)
{
	if (beads_ <= 0 || planes_ <= 0 || res_ <= 0) fail("beads, planes and radialsteps must be > 0");
	if (planes_ > 8192) fail("at most 8192 planes are supported");
	beads = beads_;
	planes = planes_;
	res = res_;
	ptiles = (planes + 31) / 32;
	ksteps = (res + 15) / 16;
	mres = (res + 7) & ~7;
	mtiles = (ptiles + 15) / 16;
	idx_bits = 0;
	while ((1 << idx_bits) < ptiles * 32) idx_bits++;
	lut.ensure((size_t)beads * planes * res * sizeof(float));
	w.ensure((size_t)res * sizeof(float));
}

void LutMatchTables::set(const float* luts, const float* weights, bool device_src, hipStream_t stream)
{
	const hipMemcpyKind k = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
	HIP_CHECK(hipMemcpyAsync(lut.p, luts, (size_t)beads * planes * res * sizeof(float), k, stream));
	have_w = weights != nullptr;
	if (weights) HIP_CHECK(hipMemcpyAsync(w.p, weights, (size_t)res * sizeof(float), k, stream));
	HIP_CHECK(hipStreamSynchronize(stream));
	prepared[1] = prepared[2] = false;
}

void LutMatchTables::prepare(int prec, hipStream_t stream)
{
	if (prec == kMatchFP32 || prepared[prec]) return;
	if (res > kMatchMaxRes) fail("the BF16/FP16 match supports at most 256 radial steps");
	frag[prec].ensure((size_t)beads * ptiles * ksteps * 64 * 16);
	norm.ensure((size_t)beads * ptiles * 32 * sizeof(float));
	mfrag[prec].ensure((size_t)beads * mtiles * ksteps * 64 * 16);
	raw_mean.ensure((size_t)beads * ptiles * mres * sizeof(float));
	HIP_CHECK(launch_lutmatch_prepare(params(prec), prec, frag[prec].as<uint4>(), mfrag[prec].as<uint4>(),
	                                  norm.as<float>(), raw_mean.as<float>(), mres, stream));
	prepared[prec] = true;
}

LutMatchParams LutMatchTables::params(int prec) const
{
	LutMatchParams p = {};
	p.beads = beads;
	p.planes = planes;
	p.res = res;
	p.ptiles = ptiles;
	p.ksteps = ksteps;
	p.idx_bits = idx_bits;
	p.lut = lut.as<float>();
	p.frag = prec > 0 ? frag[prec].as<uint4>() : nullptr;
	p.norm = norm.as<float>();
	p.mfrag = prec > 0 ? mfrag[prec].as<uint4>() : nullptr;
	p.mtiles = mtiles;
	p.w = have_w ? w.as<float>() : nullptr;
	return p;
}

struct QTrkLUTMatch {
	int dev = 0;
	std::unique_ptr<LutMatchTables> tab;
	DeviceBuffer prof, start, z, best, cand;
	bool have_lut = false;
	int frames = 0;
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;

	LutMatchParams params(int prec) const
	{
		LutMatchParams p = tab->params(prec);
		p.prof = prof.as<float>();
		p.start = start.as<int>();
		p.rowidx = nullptr;
		p.z = z.as<float>();
		p.best = best.as<int>();
		p.cand = prec > 0 ? cand.as<int>() : nullptr;
		return p;
	}

	void set_frames(int f)
	{
		if (f <= 0) fail("framesPerBead must be > 0");
		const size_t n = (size_t)tab->beads * f;
		prof.ensure(n * tab->res * sizeof(float));
		z.ensure(n * sizeof(float));
		best.ensure(n * sizeof(int));
		cand.ensure(n * kMatchTopK * sizeof(int));
		if (f != frames) {
			std::vector<int> s(tab->beads + 1);
			for (int b = 0; b <= tab->beads; b++) s[b] = b * f;
			HIP_CHECK(hipMemcpyAsync(start.p, s.data(), s.size() * sizeof(int), hipMemcpyHostToDevice, stream));
			HIP_CHECK(hipStreamSynchronize(stream));
			frames = f;
		}
	}
};

QTRK_API QTrkLUTMatch* QTrkLUTMatchCreate(int beads, int planes, int radialsteps)
{
	QTrkLUTMatch* m = nullptr;
	GUARD({
		std::unique_ptr<QTrkLUTMatch> o(new QTrkLUTMatch());
		HIP_CHECK(hipGetDevice(&o->dev));
		o->tab.reset(new LutMatchTables(beads, planes, radialsteps));
		o->start.ensure((size_t)(beads + 1) * sizeof(int));
		HIP_CHECK(hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking));
		HIP_CHECK(hipEventCreate(&o->e0));
		HIP_CHECK(hipEventCreate(&o->e1));
		m = o.release();
	});
	return m;
}

QTRK_API void QTrkLUTMatchFree(QTrkLUTMatch* m)
{
	if (!m) return;
	hipSetDevice(m->dev);
	hipStreamSynchronize(m->stream);
	hipEventDestroy(m->e0);
	hipEventDestroy(m->e1);
	hipStreamDestroy(m->stream);
	delete m;
}

QTRK_API int QTrkLUTMatchSetLUT(QTrkLUTMatch* m, const float* luts, const float* radialweights)
{
	return GUARD({
		if (!m || !luts) fail("null argument");
		HIP_CHECK(hipSetDevice(m->dev));
		m->tab->set(luts, radialweights, false, m->stream);
		m->have_lut = true;
	}) ? 0 : -1;
}

QTRK_API float* QTrkLUTMatchDeviceProfiles(QTrkLUTMatch* m, int framesPerBead)
{
	float* p = nullptr;
	GUARD({
		if (!m) fail("null argument");
		HIP_CHECK(hipSetDevice(m->dev));
		m->set_frames(framesPerBead);
		p = m->prof.as<float>();
	});
	return p;
}

QTRK_API int QTrkLUTMatchUploadProfiles(QTrkLUTMatch* m, const float* profiles, int framesPerBead)
{
	return GUARD({
		if (!m || !profiles) fail("null argument");
		HIP_CHECK(hipSetDevice(m->dev));
		m->set_frames(framesPerBead);
		HIP_CHECK(hipMemcpyAsync(m->prof.p, profiles, (size_t)m->tab->beads * framesPerBead * m->tab->res * sizeof(float),
		                         hipMemcpyHostToDevice, m->stream));
		HIP_CHECK(hipStreamSynchronize(m->stream));
	}) ? 0 : -1;
}

QTRK_API int QTrkLUTMatchRun(QTrkLUTMatch* m, int precision, int repeats)
{
	return GUARD({
		if (!m) fail("null argument");
		if (precision < QTRK_MATCH_FP32 || precision > QTRK_MATCH_FP16) fail("unknown precision");
		if (!m->have_lut) fail("no LUT set");
		if (m->frames <= 0) fail("no profiles set");
		if (repeats < 1) repeats = 1;
		HIP_CHECK(hipSetDevice(m->dev));
		const LutMatchTables& t = *m->tab;
		if (precision == QTRK_MATCH_FP32) {
			if (lutmatch_lds_bytes(precision, t.res, t.planes, t.ksteps, m->frames, 0, t.beads) > 160 * 1024)
				fail("too many planes for the FP32 match (at most about 4,900)");
		} else {
			m->tab->prepare(precision, m->stream);
		}
		const LutMatchParams p = m->params(precision);
		HIP_CHECK(hipEventRecord(m->e0, m->stream));
		for (int i = 0; i < repeats; i++) HIP_CHECK(launch_lutmatch(p, precision, m->frames, m->stream));
		HIP_CHECK(hipEventRecord(m->e1, m->stream));
	}) ? 0 : -1;
}

QTRK_API float QTrkLUTMatchSync(QTrkLUTMatch* m)
{
	float ms = -1;
	GUARD({
		if (!m) fail("null argument");
		HIP_CHECK(hipSetDevice(m->dev));
		HIP_CHECK(hipStreamSynchronize(m->stream));
		HIP_CHECK(hipEventElapsedTime(&ms, m->e0, m->e1));
	});
	return ms;
}

QTRK_API int QTrkLUTMatchDownload(QTrkLUTMatch* m, float* z, int* bestPlane, int* candidates, int n)
{
	return GUARD({
		if (!m) fail("null argument");
		if (n < 0 || (size_t)n > (size_t)m->tab->beads * m->frames) fail("n exceeds beads * framesPerBead");
		HIP_CHECK(hipSetDevice(m->dev));
		HIP_CHECK(hipStreamSynchronize(m->stream));
		if (z) HIP_CHECK(hipMemcpy(z, m->z.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
		if (bestPlane) HIP_CHECK(hipMemcpy(bestPlane, m->best.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
		if (candidates)
			HIP_CHECK(hipMemcpy(candidates, m->cand.p, (size_t)n * kMatchTopK * sizeof(int), hipMemcpyDeviceToHost));
	}) ? 0 : -1;
}
