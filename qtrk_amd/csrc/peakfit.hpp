// peakfit.hpp -- the reference's 7-point weighted quadratic peak fit, shared by
// the localization and LUT-match kernels.
//
// ComputeMaxInterp<T,7> (cputrack/LsqQuadraticFit.h:132-176): window
// [max(i-3,0), min(i+4,len)) around the argmax, x = index - iMax, weights
// indexed from 0 even when the window is truncated, Cramer solve of the
// weighted normal equations (:18-24, :64-102); the vertex -b/2a is used
// unless |a| < 1e-9 or fewer than 3 points exist.
#pragma once

#include <hip/hip_runtime.h>

namespace qtrk {

// QI_LSQFIT_WEIGHTS and ZLUT_LSQFIT_WEIGHTS (cputrack/QueuedTracker.h:106-107, 115-116);
// the Z weights are float literals widened to double, as the reference's
// double instantiation sees them.
static __constant__ float kQIW[7] = { 0.14f, 0.5f, 0.85f, 1.0f, 0.85f, 0.5f, 0.14f };
static __constant__ double kZW[7] = { (double)0.15f, (double)0.5f, (double)0.85f, (double)1.0f,
                                      (double)0.85f, (double)0.5f, (double)0.15f };

// Weighted quadratic least squares through n <= 7 points (x = i - iMax),
// returning a and b of a*x^2 + b*x + c.  Fixed 7-slot loops keep the arrays
// in registers; slots >= n are skipped so the sums run in the reference order.
template <typename T>
__device__ __forceinline__ void lsq_quad(int n, const T (&X)[7], const T (&Y)[7], const T (&Wt)[7], T& a, T& b)
{
	T Sx = 0, Sy = 0, Sx2 = 0, Sx3 = 0, Sxy = 0, Sx4 = 0, Sx2y = 0, Sw = 0;
#pragma unroll
	for (int i = 0; i < 7; i++) {
		if (i < n) {
			const T x = X[i], y = Y[i];
			Sx += x * Wt[i];
			Sy += y * Wt[i];
			const T sq = x * x;
			Sx2 += x * x * Wt[i];
			Sx3 += sq * x * Wt[i];
			Sx4 += sq * sq * Wt[i];
			Sxy += x * y * Wt[i];
			Sx2y += sq * y * Wt[i];
			Sw += Wt[i];
		}
	}
	const T s40 = Sx4, s30 = Sx3, s20 = Sx2, s10 = Sx, s21 = Sx2y, s11 = Sxy, s01 = Sy, s00 = Sw;
	const T d = s40 * (s20 * s00 - s10 * s10) - s30 * (s30 * s00 - s10 * s20) + s20 * (s30 * s10 - s20 * s20);
	a = (s21 * (s20 * s00 - s10 * s10) - s11 * (s30 * s00 - s10 * s20) + s01 * (s30 * s10 - s20 * s20)) / d;
	b = (s40 * (s11 * s00 - s01 * s10) - s30 * (s21 * s00 - s01 * s20) + s20 * (s21 * s10 - s11 * s20)) / d;
}

// Sub-sample peak position: the 7-point fit around the argmax iMax; get(k)
// returns the value at index k.
template <typename T, typename Get>
__device__ __forceinline__ T peak_fit(int iMax, int len, const T* wts, Get get)
{
	const int s = iMax - 3 > 0 ? iMax - 3 : 0;
	const int e = iMax + 4 < len ? iMax + 4 : len;
	const int np = e - s;
	if (np < 3) return (T)iMax;
	T xs[7], ys[7], w[7];
#pragma unroll
	for (int i = 0; i < 7; i++) {
		xs[i] = (T)(s + i - iMax);
		ys[i] = i < np ? get(s + i) : (T)0;
		w[i] = wts[i];
	}
	T a, b;
	lsq_quad<T>(np, xs, ys, w, a, b);
	const T lim = (T)1e-9f;
	if ((a < 0 ? -a : a) < lim) return (T)iMax;
	return (T)iMax + (-b / (2 * a));
}

}  // namespace qtrk
