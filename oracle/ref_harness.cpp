// ref_harness.cpp -- exposes the reference's own standalone numerics headers
// (cputrack/LsqQuadraticFit.h, cputrack/CubicBSpline.h), compiled unmodified
// from where they lie, through a C interface, so tests can pin the oracle's
// restatement of them bit for bit.  Test infrastructure only.
//
// LsqQuadraticFit.h uses the `uint` typedef that cputrack/std_incl.h
// provides to every translation unit that includes it; this harness
// declares the same typedef before including it.
#include <cmath>
typedef unsigned int uint;
#include "LsqQuadraticFit.h"
#include "CubicBSpline.h"

extern "C" {

float ref_max_interp_f7(float* data, int len, const float* w7)
{
	return ComputeMaxInterp<float, 7>::Compute(data, len, w7);
}

double ref_max_interp_d7(double* data, int len, const double* w7)
{
	return ComputeMaxInterp<double, 7>::Compute(data, len, w7);
}

void ref_lsqfit_f(int n, const float* x, const float* y, const float* w, float* abcd)
{
	LsqSqQuadFit<float> f(n, x, y, w);
	abcd[0] = f.a; abcd[1] = f.b; abcd[2] = f.c; abcd[3] = f.d;
}

void ref_lsqfit_d(int n, const double* x, const double* y, const double* w, double* abcd)
{
	LsqSqQuadFit<double> f(n, x, y, w);
	abcd[0] = f.a; abcd[1] = f.b; abcd[2] = f.c; abcd[3] = f.d;
}

void ref_bspline_weights(float t, float* w) { ComputeBSplineWeights(w, t); }

void ref_bspline_derivatives(float t, const float* k, float* d1d2)
{
	float a, b, kk[4] = { k[0], k[1], k[2], k[3] };
	ComputeBSplineDerivatives(t, kk, a, b);
	d1d2[0] = a; d1d2[1] = b;
}

float ref_spline_fit_max(float* data, int len) { return ComputeSplineFitMaxPos(data, len); }
}
