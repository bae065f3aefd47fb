// lutmodel.hpp -- differentiable bead-image model rendered from a radial Z-LUT,
// its hand-written adjoint, and a Poisson maximum-likelihood fit of
// theta = (x, y, z, I, bg) with per-ROI 5x5 Fisher matrix and CRLB.
//
// Model: mu(px) = bg + I * g(px; x, y, z), where g is the reference's
// GenerateImageFromLUT (cputrack/utils.cpp:325-390): the mean over ovs x ovs
// sub-samples of the z-interpolated LUT row (cubic B-spline over planes
// iz-1..iz+2 with iz = clamp((int)z, 1, planes-3), cputrack/CubicBSpline.h:15-24,
// or linear), itself interpolated linearly in r = (|p - c| - minr) * steps /
// (maxr - minr) clamped to [0, steps-2], pixel centres at integer coordinates.
// The forward pass reproduces that arithmetic operation for operation.
//
// Derivatives (the adjoint): d g / d x = mean over sub-samples of
// (zi[i+1] - zi[i]) * steps/(maxr-minr) * (-(X)/|p - c|), zero where r is
// clamped; likewise for y; d g / d z = the same interpolation of dzi/dz, the
// derivative of the B-spline blend (ComputeBSplineDerivatives,
// cputrack/CubicBSpline.h:41-50); d mu / d I = g; d mu / d bg = 1.
#pragma once

#include <hip/hip_runtime.h>

namespace qtrk {

constexpr int kModelParams = 5;  // x, y, z, I, bg
constexpr int kMaxFitOvs = 8;    // sub-samples per pixel side the fit supports (forward/backward: any)
constexpr int kMaxWideFitOvs = 6;  // ... with 512 lanes per ROI
constexpr int kMax1024FitOvs = 4;  // ... with 1024 lanes per ROI

struct LutModelParams {
	const float* luts;       // [count][planes][steps]
	int count, planes, steps;
	float minr, maxr;
	int width, height, ovs, spline;
	int n;
	const float* theta;      // [n][5] model parameters (input of forward/backward, start of the fit)
	const int* lut_index;    // LUT of ROI i at lut_index[i * lut_index_stride], or null (LUT 0)
	int lut_index_stride;    // in ints (1 for a plain array, 6 for the zlutIndex of 24-byte job records)
	// forward
	float* mu;               // [n][height][width] or null
	float* jac;              // [n][5][height][width] or null
	// backward: grad[k] = sum_px gmu * d mu / d theta_k
	const float* gmu;        // [n][height][width]
	float* grad;             // [n][5]
	// fit
	const void* data;        // [n][height][width] photon counts
	int data_pdt;            // pixel type of data: 0 u8, 1 u16, 2 f32
	float* theta_out;        // [n][5] fitted parameters (may alias theta)
	int iters;               // Fisher-scoring Newton iterations; a start with I <= 0 gets I and bg
	                         // by linear least squares at its (x, y, z) first
	double* fisher;          // [n][25] Fisher matrix at the fitted parameters, or null
	float* crlb;             // [n][5] sqrt(diag(F^-1)), or null
	float* nll;              // [n] negative log-likelihood (without the data-only log-factorial term), or null
	// fit start from localizations instead of theta (when start_xy is set): x and y at
	// start_xy[i * start_stride] and [+1], z at start_z[i], each clamped into the ROI and the LUT
	// (NaN: the middle), I = bg = -1 (the least-squares start)
	const float* start_xy;
	int start_stride;        // in floats
	const float* start_z;
};

hipError_t launch_lutmodel_forward(const LutModelParams& p, hipStream_t stream);
hipError_t launch_lutmodel_backward(const LutModelParams& p, hipStream_t stream);
// block: lanes per ROI, 64, 128, 256, 512 (up to kMaxWideFitOvs sub-samples per side) or 1024 (up to
// kMax1024FitOvs)
hipError_t launch_lutmodel_fit(const LutModelParams& p, hipStream_t stream, int block = 256);
bool lutmodel_fit_block_ok(int block, int ovs);
size_t lutmodel_lds_bytes(const LutModelParams& p, bool fit, int block = 256);

}  // namespace qtrk
