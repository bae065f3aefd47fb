// lutmodel_host.cpp -- the QTrkLUTModel C-ABI object: the LUTs in GPU memory
// and launches of the forward, adjoint and MLE kernels on a caller's stream.
#include "errors.hpp"
#include "hip_tracker.hpp"
#include "lutmodel.hpp"
#include "qtrk_amd.h"

#include <algorithm>
#include <memory>

using namespace qtrk;

struct QTrkLUTModel {
	int dev = 0;
	LutModelParams base = {};
	DeviceBuffer luts;
	hipStream_t own = nullptr;  // uploads only
	int fit_lanes = 256;        // lanes per ROI of FitMLE (QTrkLUTModelSetFitLanes)

	// NULL is the device's default (null) stream, as in every HIP call
	static hipStream_t stream(void* s) { return (hipStream_t)s; }
	LutModelParams params(const float* theta, const int* lutIndex, int n) const
	{
		LutModelParams p = base;
		p.luts = luts.as<float>();
		p.theta = theta;
		p.lut_index = lutIndex;
		p.lut_index_stride = 1;
		p.n = n;
		return p;
	}
	void check(int n) const
	{
		if (n < 0) fail("n must be >= 0");
		if (lutmodel_lds_bytes(base, true, fit_lanes) > 160 * 1024) fail("ROI too large for the fit (LDS)");
	}
};

QTRK_API QTrkLUTModel* QTrkLUTModelCreate(const float* luts, int count, int planes, int steps, float minradius,
                                          float maxradius, int width, int height, int ovs, bool splineInterp)
{
	QTrkLUTModel* m = nullptr;
	GUARD({
		if (!luts) fail("null LUT");
		if (count <= 0 || steps < 2 || width <= 0 || height <= 0 || ovs <= 0) fail("invalid sizes");
		if (planes < (splineInterp ? 4 : 2)) fail("too few planes for the interpolation");
		if (!(maxradius > minradius)) fail("maxradius must exceed minradius");
		std::unique_ptr<QTrkLUTModel> o(new QTrkLUTModel());
		HIP_CHECK(hipGetDevice(&o->dev));
		LutModelParams& b = o->base;
		b.count = count;
		b.planes = planes;
		b.steps = steps;
		b.minr = minradius;
		b.maxr = maxradius;
		b.width = width;
		b.height = height;
		b.ovs = ovs;
		b.spline = splineInterp ? 1 : 0;
		const size_t bytes = (size_t)count * planes * steps * sizeof(float);
		o->luts.ensure(bytes);
		HIP_CHECK(hipStreamCreateWithFlags(&o->own, hipStreamNonBlocking));
		HIP_CHECK(hipMemcpyAsync(o->luts.p, luts, bytes, hipMemcpyHostToDevice, o->own));
		HIP_CHECK(hipStreamSynchronize(o->own));
		m = o.release();
	});
	return m;
}

QTRK_API void QTrkLUTModelFree(QTrkLUTModel* m)
{
	if (!m) return;
	hipSetDevice(m->dev);
	hipStreamSynchronize(m->own);
	hipStreamDestroy(m->own);
	delete m;
}

QTRK_API int QTrkLUTModelForward(QTrkLUTModel* m, const float* theta, const int* lutIndex, int n, float* mu,
                                 float* jacobian, void* stream)
{
	return GUARD({
		if (!m || !theta || (!mu && !jacobian)) fail("null argument");
		m->check(n);
		HIP_CHECK(hipSetDevice(m->dev));
		LutModelParams p = m->params(theta, lutIndex, n);
		p.mu = mu;
		p.jac = jacobian;
		HIP_CHECK(launch_lutmodel_forward(p, m->stream(stream)));
	}) ? 0 : -1;
}

QTRK_API int QTrkLUTModelBackward(QTrkLUTModel* m, const float* theta, const int* lutIndex, int n,
                                  const float* gradMu, float* gradTheta, void* stream)
{
	return GUARD({
		if (!m || !theta || !gradMu || !gradTheta) fail("null argument");
		m->check(n);
		HIP_CHECK(hipSetDevice(m->dev));
		LutModelParams p = m->params(theta, lutIndex, n);
		p.gmu = gradMu;
		p.grad = gradTheta;
		HIP_CHECK(launch_lutmodel_backward(p, m->stream(stream)));
	}) ? 0 : -1;
}

QTRK_API int QTrkLUTModelFitMLE(QTrkLUTModel* m, const float* data, const float* thetaInit, const int* lutIndex,
                                int n, int iterations, float* theta, double* fisher, float* crlb, float* nll,
                                void* stream)
{
	return GUARD({
		if (!m || !data || !thetaInit || !theta) fail("null argument");
		if (iterations < 0) fail("iterations must be >= 0");
		if (m->base.ovs > kMaxFitOvs) fail("the fit supports at most " + std::to_string(kMaxFitOvs) + " sub-samples per pixel side");
		m->check(n);
		HIP_CHECK(hipSetDevice(m->dev));
		LutModelParams p = m->params(thetaInit, lutIndex, n);
		p.data = data;
		p.data_pdt = 2;
		p.theta_out = theta;
		p.iters = iterations;
		p.fisher = fisher;
		p.crlb = crlb;
		p.nll = nll;
		HIP_CHECK(launch_lutmodel_fit(p, m->stream(stream), m->fit_lanes));
	}) ? 0 : -1;
}

QTRK_API int QTrkLUTModelSetFitLanes(QTrkLUTModel* m, int lanes)
{
	return GUARD({
		if (!m) fail("null argument");
		if (!lutmodel_fit_block_ok(lanes, std::min(m->base.ovs, kMaxFitOvs)))
			fail("lanes per ROI must be 64, 128, 256, 512 (at most " + std::to_string(kMaxWideFitOvs) +
			     " sub-samples per pixel side) or 1024 (at most " + std::to_string(kMax1024FitOvs) + ")");
		if (lutmodel_lds_bytes(m->base, true, lanes) > 160 * 1024) fail("ROI too large for the fit (LDS)");
		m->fit_lanes = lanes;
	}) ? 0 : -1;
}

QTRK_API int QTrkLUTModelSync(QTrkLUTModel* m, void* stream)
{
	return GUARD({
		if (!m) fail("null argument");
		HIP_CHECK(hipSetDevice(m->dev));
		HIP_CHECK(hipStreamSynchronize(m->stream(stream)));
	}) ? 0 : -1;
}
