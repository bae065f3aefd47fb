// lutmodel.hip -- the LUT image model (forward), its adjoint (backward) and
// the Poisson MLE with Fisher/CRLB, for gfx950.  One workgroup (256 lanes)
// per ROI: the z-blended LUT row zi[] and its z-derivative dzi[] are built once
// per ROI (per fit iteration) in LDS, then lanes take pixels.
//
// The forward pass follows GenerateImageFromLUT's float arithmetic operation
// for operation (this file is built with -ffp-contract=off and IEEE sqrt):
// B-spline weights as ComputeBSplineWeights (cputrack/CubicBSpline.h:15-24),
// zi[r] summed over the 4 planes in order, sub-samples summed ox-major, the
// sum divided by ovs*ovs (cputrack/utils.cpp:325-390).
#include "lutmodel.hpp"

#include <cfloat>
#include <cstdint>

namespace qtrk {

constexpr int kModelBlock = 256;  // lanes per ROI of the forward/backward passes; the fit takes 64-1024

struct ModelLds {
	int zi, dzi, row, data, red, th, total;
};
__host__ __device__ inline int m16(int x) { return (x + 15) & ~15; }
// `waves`: wavefronts per workgroup whose partial sums block_sum_d keeps (at least 4)
__host__ __device__ inline ModelLds model_lds(const LutModelParams& P, bool fit, int waves = 4)
{
	ModelLds L;
	int o = 0;
	L.zi = o;    o += m16(P.steps * 4);
	L.dzi = o;   o += m16(P.steps * 4);
	L.row = o;   o += fit ? m16(P.steps * 16) : 0;
	L.data = o;  o += fit ? m16(P.width * P.height * 4) : 0;
	L.red = o;   o += m16(((waves > 4 ? waves : 4) + 1) * 24 * 8);  // the totals + one row per wavefront
	L.th = o;    o += m16(8 * 4);
	L.total = o;
	return L;
}
size_t lutmodel_lds_bytes(const LutModelParams& p, bool fit, int block)
{
	return (size_t)model_lds(p, fit, block / 64).total;
}

// zi[r] and dzi[r] for depth z: the cubic B-spline blend of planes iz-1..iz+2
// and its derivative d/dz (or the linear blend of planes iz, iz+1 and its slope).
__device__ void build_z_row(const LutModelParams& P, const float* lut, float z, float* zi, float* dzi)
{
	const int S = P.steps;
	// (int)z as the reference computes it, defined for every float (NaN -> 0)
	const int izt = z != z ? 0 : (int)fminf(fmaxf(z, -2.0e9f), 2.0e9f);
	if (P.spline) {
		int iz = izt;
		if (iz > P.planes - 3) iz = P.planes - 3;
		if (iz < 1) iz = 1;
		const float t = z - iz;
		const float t2 = t * t, t3 = t2 * t, omt = 1 - t;
		const float w0 = 1.0f / 6 * omt * omt * omt;
		const float w1 = 1.0f / 6 * (3 * t3 - 6 * t2 + 4);
		const float w2 = 1.0f / 6 * (-3 * t3 + 3 * t2 + 3 * t + 1);
		const float w3 = 1.0f / 6 * t3;
		const float d0 = -omt * omt, d1 = 3 * t2 - 4 * t, d2 = -3 * t2 + 2 * t + 1, d3 = t2;
		const float* L0 = lut + (size_t)(iz - 1) * S;
		for (int r = threadIdx.x; r < S; r += blockDim.x) {
			const float a = L0[r], b = L0[S + r], c = L0[2 * S + r], d = L0[3 * S + r];
			float v = 0;
			v += w0 * a;
			v += w1 * b;
			v += w2 * c;
			v += w3 * d;
			zi[r] = v;
			dzi[r] = 0.5f * (d0 * a + d1 * b + d2 * c + d3 * d);
		}
	} else {
		const int iz = izt;
		for (int r = threadIdx.x; r < S; r += blockDim.x) {
			if (iz < 0) {
				zi[r] = lut[r];
				dzi[r] = 0;
			} else if (iz >= P.planes - 1) {
				zi[r] = lut[(size_t)(P.planes - 1) * S + r];
				dzi[r] = 0;
			} else {
				const float a = lut[(size_t)iz * S + r], b = lut[(size_t)(iz + 1) * S + r];
				zi[r] = a + (b - a) * (z - iz);
				dzi[r] = b - a;
			}
		}
	}
}

__device__ __forceinline__ int lut_of(const LutModelParams& P, int roi)
{
	if (!P.lut_index) return 0;
	const int li = P.lut_index[(size_t)roi * P.lut_index_stride];
	return li >= 0 && li < P.count ? li : 0;
}

// g and its derivatives with respect to the bead position at pixel (x, y).
struct PixelG {
	float g, gx, gy, gz;
};
// OVS > 0: the sub-sample count as a compile-time constant (a column's sub-samples
// unroll, their LUT reads in flight together); 0: P.ovs at run time.
template <bool DERIV, int OVS>
__device__ __forceinline__ PixelG pixel_g(const LutModelParams& P, int x, int y, float px, float py, float radialcov,
                                          float step, const float* zi, const float* dzi)
{
	float s = 0, sx = 0, sy = 0, sz = 0;
	const float rmax = (float)(P.steps - 2);
	const int ovs = OVS > 0 ? OVS : P.ovs;
#pragma unroll 1
	for (int ox = 0; ox < ovs; ox++) {  // one column's sub-samples unrolled
#pragma unroll
		for (int oy = 0; oy < ovs; oy++) {
			const float X = x + (ox + 0.5f) * step - px - 0.5f;
			const float Y = y + (oy + 0.5f) * step - py - 0.5f;
			const float pr = sqrtf(X * X + Y * Y);
			const float ru = (pr - P.minr) * radialcov;
			const bool inside = ru >= 0 && ru <= rmax;  // the slope in r is zero where r is clamped
			// clamp to [0, steps-2]; a NaN radius (NaN position) maps to bin 0
			const float r = fminf(fmaxf(ru, 0.0f), rmax);
			const int i = (int)r;
			const float fr = r - i;
			const float a = zi[i], b = zi[i + 1];
			s += a + (b - a) * fr;
			if (DERIV) {
				const float da = dzi[i], db = dzi[i + 1];
				sz += da + (db - da) * fr;
				if (inside && pr > 0) {
					const float k = (b - a) * radialcov / pr;
					sx -= k * X;
					sy -= k * Y;
				}
			}
		}
	}
	const float inv = 1.0f / (float)(ovs * ovs);
	PixelG o;
	o.g = s / (float)(ovs * ovs);
	o.gx = sx * inv;
	o.gy = sy * inv;
	o.gz = sz * inv;
	return o;
}

// The same model for the fit, where nothing has to reproduce the reference's
// rounding: the hardware reciprocal square root replaces the IEEE square root
// and division, the row segment (a, b - a, da, db - da) of each radial bin is
// one 16-byte LDS read from a table built per iteration, the bin clamp is one
// med3 into that table, whose end entries have zero slope (so the r-derivative
// of clamped samples is zero without a test), the sub-sample offsets are
// hoisted per pixel, and products are fused.  Agrees with pixel_g to float
// rounding (at r exactly on the last bin's edge the derivative is taken as the
// clamped side's).
// OVS, the sub-samples per pixel side, is a compile-time constant so that the loops
// unroll without exits and the table reads of several sub-samples are in flight at once.
template <int OVS>
__device__ __forceinline__ PixelG pixel_g_fast(const LutModelParams& P, int x, int y, float px, float py,
                                               float radialcov, float step, const float4* row)
{
	constexpr int kMaxOvs = OVS;
	const float tmax = (float)(P.steps - 1);
	const float mr1 = P.minr * radialcov - 1.0f;  // t = (r - minr) * radialcov + 1
	// offsets in bin units (X' = X radialcov): r' = |(X', Y')| = r radialcov, and the slope
	// term q.y radialcov / r times X is q.y / r' times X', scaled back once at the end
	const float sstep = step * radialcov;
	const float X0 = ((float)x + 0.5f * step - px - 0.5f) * radialcov;
	const float Y0 = ((float)y + 0.5f * step - py - 0.5f) * radialcov;
	constexpr int ovs = OVS;
	float Ys[kMaxOvs], Y2[kMaxOvs];
#pragma unroll
	for (int o = 0; o < kMaxOvs; o++) {
		Ys[o] = __builtin_fmaf((float)o, sstep, Y0);
		// + 1e-30 keeps r'^2 > 0 at the centre itself (its slope term is then 0 * finite)
		Y2[o] = __builtin_fmaf(Ys[o], Ys[o], 1e-30f);
	}
	// the slope sums sum_{ox,oy} k X_ox and sum k Y_oy are taken as sum_ox X_ox (sum_oy k)
	// and sum_oy Y_oy (sum_ox k): per sub-sample two fused adds into the marginals
	float sa = 0, sb = 0, za = 0, zb = 0, sx = 0, kr[ovs];
#pragma unroll
	for (int oy = 0; oy < ovs; oy++) kr[oy] = 0.0f;
#pragma unroll
	for (int ox = 0; ox < ovs; ox++) {
		const float X = __builtin_fmaf((float)ox, sstep, X0);
		const float X2 = X * X;
		// a column's table reads are all issued before the first is used
		float rs[ovs], fr[ovs];
		float4 q[ovs];
#pragma unroll
		for (int oy = 0; oy < ovs; oy++) {
			const float r2 = X2 + Y2[oy];
			rs[oy] = __builtin_amdgcn_rsqf(r2);
			const float t = __builtin_amdgcn_fmed3f(__builtin_fmaf(r2, rs[oy], -mr1), 0.0f, tmax);
			fr[oy] = __builtin_amdgcn_fractf(t);
			q[oy] = row[(int)t];
		}
		float kc = 0.0f;
#pragma unroll
		for (int oy = 0; oy < ovs; oy++) {
			sa += q[oy].x;
			sb = __builtin_fmaf(q[oy].y, fr[oy], sb);
			za += q[oy].z;
			zb = __builtin_fmaf(q[oy].w, fr[oy], zb);
			// k = q.y / r', zero in the clamped entries
			kc = __builtin_fmaf(q[oy].y, rs[oy], kc);
			kr[oy] = __builtin_fmaf(q[oy].y, rs[oy], kr[oy]);
		}
		sx = __builtin_fmaf(-kc, X, sx);
	}
	float sy = 0.0f;
#pragma unroll
	for (int oy = 0; oy < ovs; oy++) sy = __builtin_fmaf(-kr[oy], Ys[oy], sy);
	const float inv = 1.0f / (float)(ovs * ovs);
	const float dinv = inv * radialcov;
	PixelG o;
	o.g = (sa + sb) * inv;
	o.gx = sx * dinv;
	o.gy = sy * dinv;
	o.gz = (za + zb) * inv;
	return o;
}

// The fit's row table from zi and dzi, indexed by t = r + 1 clamped to [0, steps-1]:
// entry 0 holds the radii below the LUT (zi[0], zero slope), entries 1 .. steps-2 the
// bins 0 .. steps-3, entry steps-1 the radii at and beyond the last bin the clamp keeps
// (zi[steps-2], zero slope).  Zero slopes make the clamped samples' r-derivative zero
// without a test.
__device__ __forceinline__ void build_row_table(const LutModelParams& P, const float* zi, const float* dzi, float4* row)
{
	const int S = P.steps;
	for (int r = threadIdx.x; r < S; r += blockDim.x) {
		if (r == 0) row[0] = make_float4(zi[0], 0.0f, dzi[0], 0.0f);
		else if (r == S - 1) row[r] = make_float4(zi[S - 2], 0.0f, dzi[S - 2], 0.0f);
		else row[r] = make_float4(zi[r - 1], zi[r] - zi[r - 1], dzi[r - 1], dzi[r] - dzi[r - 1]);
	}
}

// ------------------------------------------------------------ forward / backward

template <bool JAC, int OVS>
__global__ __launch_bounds__(kModelBlock) void lutmodel_forward_kernel(LutModelParams P)
{
	extern __shared__ __align__(16) unsigned char smem[];
	const int roi = blockIdx.x;
	const ModelLds L = model_lds(P, false);
	float* zi = (float*)(smem + L.zi);
	float* dzi = (float*)(smem + L.dzi);
	const float* th = P.theta + (size_t)roi * kModelParams;
	const float x0 = th[0], y0 = th[1], z0 = th[2], I = th[3], bg = th[4];
	build_z_row(P, P.luts + (size_t)lut_of(P, roi) * P.planes * P.steps, z0, zi, dzi);
	__syncthreads();
	const float radialcov = P.steps / (P.maxr - P.minr);
	const float step = 1.0f / P.ovs;
	const int npx = P.width * P.height;
	// pixel p = y * width + x, stepped by the block size without a division per pixel
	const int dyp = kModelBlock / P.width, dxp = kModelBlock - dyp * P.width;
	int y = (int)threadIdx.x / P.width, x = (int)threadIdx.x - y * P.width;
	for (int p = threadIdx.x; p < npx; p += kModelBlock, x += dxp, y += dyp) {
		if (x >= P.width) { x -= P.width; y++; }
		const PixelG G = pixel_g<JAC, OVS>(P, x, y, x0, y0, radialcov, step, zi, dzi);
		if (P.mu) P.mu[(size_t)roi * npx + p] = bg + I * G.g;
		if (JAC) {
			float* J = P.jac + (size_t)roi * kModelParams * npx + p;
			J[0] = I * G.gx;
			J[npx] = I * G.gy;
			J[2 * npx] = I * G.gz;
			J[3 * npx] = G.g;
			J[4 * npx] = 1.0f;
		}
	}
}

// Sum of K per-lane float partials over the workgroup in double; the result is
// in red[0..K) for every lane after the call (red[24 + 24 w ..] hold wavefront w's sums).
template <int K>
__device__ __forceinline__ void block_sum_d(const float (&v)[K], double* red)
{
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	double d[K];
#pragma unroll
	for (int k = 0; k < K; k++) {
		d[k] = (double)v[k];
		for (int off = 32; off > 0; off >>= 1) d[k] += __shfl_xor(d[k], off);
	}
	__syncthreads();  // red may still be read from the previous call
	if (lane == 0)
#pragma unroll
		for (int k = 0; k < K; k++) red[24 + wid * 24 + k] = d[k];
	__syncthreads();
	if (threadIdx.x < K) {
		double s = 0;
		for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[24 + w * 24 + threadIdx.x];
		red[threadIdx.x] = s;
	}
	__syncthreads();
}

template <int OVS>
__global__ __launch_bounds__(kModelBlock) void lutmodel_backward_kernel(LutModelParams P)
{
	extern __shared__ __align__(16) unsigned char smem[];
	const int roi = blockIdx.x;
	const ModelLds L = model_lds(P, false);
	float* zi = (float*)(smem + L.zi);
	float* dzi = (float*)(smem + L.dzi);
	double* red = (double*)(smem + L.red);
	const float* th = P.theta + (size_t)roi * kModelParams;
	const float x0 = th[0], y0 = th[1], z0 = th[2], I = th[3];
	build_z_row(P, P.luts + (size_t)lut_of(P, roi) * P.planes * P.steps, z0, zi, dzi);
	__syncthreads();
	const float radialcov = P.steps / (P.maxr - P.minr);
	const float step = 1.0f / P.ovs;
	const int npx = P.width * P.height;
	float acc[kModelParams] = { 0, 0, 0, 0, 0 };
	const int dyp = kModelBlock / P.width, dxp = kModelBlock - dyp * P.width;
	int y = (int)threadIdx.x / P.width, x = (int)threadIdx.x - y * P.width;
	for (int p = threadIdx.x; p < npx; p += kModelBlock, x += dxp, y += dyp) {
		if (x >= P.width) { x -= P.width; y++; }
		const PixelG G = pixel_g<true, OVS>(P, x, y, x0, y0, radialcov, step, zi, dzi);
		const float gm = P.gmu[(size_t)roi * npx + p];
		acc[0] += gm * I * G.gx;
		acc[1] += gm * I * G.gy;
		acc[2] += gm * I * G.gz;
		acc[3] += gm * G.g;
		acc[4] += gm;
	}
	block_sum_d<kModelParams>(acc, red);
	if (threadIdx.x < kModelParams) P.grad[(size_t)roi * kModelParams + threadIdx.x] = (float)red[threadIdx.x];
}

// ------------------------------------------------------------ maximum likelihood

// Cholesky factorisation A = L L^T of a symmetric positive definite 5x5
// matrix, L stored in the lower triangle of a; false if A is not positive
// definite.  Fully unrolled so the matrix stays in registers.
__device__ __forceinline__ bool chol5(double (&a)[5][5])
{
#pragma unroll
	for (int j = 0; j < 5; j++) {
		double s = a[j][j];
#pragma unroll
		for (int k = 0; k < j; k++) s -= a[j][k] * a[j][k];
		if (!(s > 0)) return false;
		const double d = sqrt(s);
		a[j][j] = d;
#pragma unroll
		for (int i = j + 1; i < 5; i++) {
			double t = a[i][j];
#pragma unroll
			for (int k = 0; k < j; k++) t -= a[i][k] * a[j][k];
			a[i][j] = t / d;
		}
	}
	return true;
}

// Solves L L^T x = b in place.
__device__ __forceinline__ void chol5_solve(const double (&L)[5][5], double (&b)[5])
{
#pragma unroll
	for (int i = 0; i < 5; i++) {
		double s = b[i];
#pragma unroll
		for (int k = 0; k < i; k++) s -= L[i][k] * b[k];
		b[i] = s / L[i][i];
	}
#pragma unroll
	for (int i = 4; i >= 0; i--) {
		double s = b[i];
#pragma unroll
		for (int k = i + 1; k < 5; k++) s -= L[k][i] * b[k];
		b[i] = s / L[i][i];
	}
}

// Index of (i, j), i <= j, in the packed upper triangle of a 5x5 matrix.
__host__ __device__ constexpr int tri5(int i, int j) { return i * 5 - i * (i - 1) / 2 + (j - i); }

// Fisher-scoring Newton iterations on the Poisson negative log-likelihood
// sum(mu - d ln mu): gradient sum (1 - d/mu) J, Fisher matrix sum J J^T / mu
// (J = d mu / d theta), step F^-1 (-gradient) solved in FP64.  Steps are
// limited to 1 px in x/y and 2 planes in z per iteration; I and bg that would
// turn non-positive are halved instead; z stays within the LUT.  After the
// last update one more pass evaluates F, its FP64 inverse (the CRLB) and the
// negative log-likelihood at the fitted parameters.
// MAXB: the largest workgroup the instance is launched with (256, 512 or 1024; a workgroup of 1024
// lanes puts 4 wavefronts on each SIMD, which leaves 128 VGPRs per lane).
template <int OVS, int MAXB>
__global__ __launch_bounds__(MAXB) void lutmodel_fit_kernel(LutModelParams P)
{
	extern __shared__ __align__(16) unsigned char smem[];
	const int roi = blockIdx.x;
	const int nt = blockDim.x;
	const ModelLds L = model_lds(P, true, nt >> 6);
	float* zi = (float*)(smem + L.zi);
	float* dzi = (float*)(smem + L.dzi);
	float4* row = (float4*)(smem + L.row);
	float* dat = (float*)(smem + L.data);
	double* red = (double*)(smem + L.red);
	float* th = (float*)(smem + L.th);
	const int npx = P.width * P.height;
	const size_t d0 = (size_t)roi * npx;
	if (P.data_pdt == 1)
		for (int p = threadIdx.x; p < npx; p += nt) dat[p] = (float)((const uint16_t*)P.data)[d0 + p];
	else if (P.data_pdt == 0)
		for (int p = threadIdx.x; p < npx; p += nt) dat[p] = (float)((const uint8_t*)P.data)[d0 + p];
	else
		for (int p = threadIdx.x; p < npx; p += nt) dat[p] = ((const float*)P.data)[d0 + p];
	if (threadIdx.x < kModelParams) {
		float v;
		if (P.start_xy) {
			// (x, y, z, -1, -1) from the localization and the matched z
			const int k = threadIdx.x;
			const float lim = k == 0 ? (float)(P.width - 1) : k == 1 ? (float)(P.height - 1) : (float)(P.planes - 1);
			const float s = k < 2 ? P.start_xy[(size_t)roi * P.start_stride + k] : k == 2 ? P.start_z[roi] : -1.0f;
			v = k > 2 ? -1.0f : s == s ? fminf(fmaxf(s, 0.0f), lim) : 0.5f * lim;
		} else {
			v = P.theta[(size_t)roi * kModelParams + threadIdx.x];
		}
		th[threadIdx.x] = v;
	}
	const float* lut = P.luts + (size_t)lut_of(P, roi) * P.planes * P.steps;
	const float radialcov = P.steps / (P.maxr - P.minr);
	const float step = 1.0f / P.ovs;
	// pixel p = y * width + x stepped by the block size without a division per pixel
	// (dxp = nt mod width < width, so x wraps at most once per step)
	const int y1 = (int)threadIdx.x / P.width, x1 = (int)threadIdx.x - y1 * P.width;
	const int dyp = nt / P.width, dxp = nt - dyp * P.width;
	__syncthreads();

	if (!(th[3] > 0)) {
		// start: I and bg by linear least squares of d = bg + I g at the start position
		build_z_row(P, lut, th[2], zi, dzi);
		__syncthreads();
		build_row_table(P, zi, dzi, row);
		__syncthreads();
		float v[4] = { 0, 0, 0, 0 };
		int x = x1, y = y1;
		for (int p = threadIdx.x; p < npx; p += nt, x += dxp, y += dyp) {
			if (x >= P.width) { x -= P.width; y++; }
			const PixelG G = pixel_g_fast<OVS>(P, x, y, th[0], th[1], radialcov, step, row);
			const float d = dat[p];
			v[0] += G.g;
			v[1] += G.g * G.g;
			v[2] += G.g * d;
			v[3] += d;
		}
		block_sum_d<4>(v, red);
		if (threadIdx.x == 0) {
			const double sg = red[0], sgg = red[1], sgd = red[2], sd = red[3], N = (double)npx;
			const double det = sgg * N - sg * sg;
			double I = det > 0 ? (N * sgd - sg * sd) / det : 0.0;
			double bg = det > 0 ? (sgg * sd - sg * sgd) / det : sd / N;
			if (!(bg > 1e-3)) bg = 1e-3;
			if (!(I > 0)) I = sg > 0 ? fmax((sd - bg * N) / sg, 1.0) : 1.0;
			th[3] = (float)I;
			th[4] = (float)bg;
		}
		__syncthreads();
	}

	for (int it = 0; it <= P.iters; it++) {
		const float x0 = th[0], y0 = th[1], z0 = th[2], I = th[3], bg = th[4];
		build_z_row(P, lut, z0, zi, dzi);
		__syncthreads();
		build_row_table(P, zi, dzi, row);
		__syncthreads();
		const bool last = it == P.iters;
		float v[21];
#pragma unroll
		for (int k = 0; k < 21; k++) v[k] = 0;
		int x = x1, y = y1;
		for (int p = threadIdx.x; p < npx; p += nt, x += dxp, y += dyp) {
			if (x >= P.width) { x -= P.width; y++; }
			const PixelG G = pixel_g_fast<OVS>(P, x, y, x0, y0, radialcov, step, row);
			float mu = __builtin_fmaf(I, G.g, bg);
			mu = mu > 1e-6f ? mu : 1e-6f;
			const float d = dat[p];
			const float J[5] = { I * G.gx, I * G.gy, I * G.gz, G.g, 1.0f };
			const float fw = __builtin_amdgcn_rcpf(mu);
			const float gw = __builtin_fmaf(-d, fw, 1.0f);
#pragma unroll
			for (int i = 0; i < 5; i++) {
				v[i] = __builtin_fmaf(gw, J[i], v[i]);
				const float wJ = fw * J[i];
#pragma unroll
				for (int j = i; j < 5; j++) v[5 + tri5(i, j)] = __builtin_fmaf(wJ, J[j], v[5 + tri5(i, j)]);
			}
			if (last) v[20] += mu - d * logf(mu);
		}
		block_sum_d<21>(v, red);
		if (threadIdx.x == 0) {
			const double* s = red;
			double F[5][5], g[5];
			for (int i = 0; i < 5; i++) {
				g[i] = -s[i];
				for (int j = i; j < 5; j++) F[i][j] = F[j][i] = s[5 + tri5(i, j)];
			}
			double Lc[5][5];
#pragma unroll
			for (int i = 0; i < 5; i++)
#pragma unroll
				for (int j = 0; j < 5; j++) Lc[i][j] = F[i][j];
			const bool spd = chol5(Lc);
			if (!last) {
				bool finite = spd;
				if (spd) {
					chol5_solve(Lc, g);
#pragma unroll
					for (int k = 0; k < 5; k++) finite = finite && isfinite(g[k]);
				}
				if (finite) {
					const float lim[3] = { 1.0f, 1.0f, 2.0f };
					for (int k = 0; k < 3; k++) {
						const float dk = (float)g[k];
						th[k] += dk > lim[k] ? lim[k] : (dk < -lim[k] ? -lim[k] : dk);
					}
					const float zmax = (float)(P.planes - 1);
					th[2] = th[2] < 0 ? 0 : (th[2] > zmax ? zmax : th[2]);
					const float In = th[3] + (float)g[3], bn = th[4] + (float)g[4];
					th[3] = In > 0 ? In : th[3] * 0.5f;
					th[4] = bn > 0 ? bn : th[4] * 0.5f;
				}
			} else {
				float* to = P.theta_out + (size_t)roi * kModelParams;
				for (int k = 0; k < 5; k++) to[k] = th[k];
				if (P.fisher)
					for (int i = 0; i < 5; i++)
						for (int j = 0; j < 5; j++) P.fisher[(size_t)roi * 25 + i * 5 + j] = F[i][j];
				if (P.nll) P.nll[roi] = (float)s[20];
				if (P.crlb) {
					// column k of F^-1 solves F c = e_k; the CRLB is the square root of its k-th entry
#pragma unroll
					for (int k = 0; k < 5; k++) {
						double e[5] = { 0, 0, 0, 0, 0 };
						e[k] = 1;
						if (spd) chol5_solve(Lc, e);
						P.crlb[(size_t)roi * kModelParams + k] = spd && e[k] > 0 ? (float)sqrt(e[k]) : __builtin_nanf("");
					}
				}
			}
		}
		__syncthreads();
	}
}

// ------------------------------------------------------------ launches

static hipError_t set_lds(const void* k, size_t lds)
{
	if (lds <= 64 * 1024) return hipSuccess;
	return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <bool JAC, int OVS>
static hipError_t launch_forward(const LutModelParams& p, size_t lds, hipStream_t stream)
{
	hipError_t e = set_lds((const void*)lutmodel_forward_kernel<JAC, OVS>, lds);
	if (e != hipSuccess) return e;
	lutmodel_forward_kernel<JAC, OVS><<<p.n, kModelBlock, lds, stream>>>(p);
	return hipGetLastError();
}

template <bool JAC>
static hipError_t launch_forward_ovs(const LutModelParams& p, size_t lds, hipStream_t stream)
{
	switch (p.ovs) {
	case 1: return launch_forward<JAC, 1>(p, lds, stream);
	case 2: return launch_forward<JAC, 2>(p, lds, stream);
	case 3: return launch_forward<JAC, 3>(p, lds, stream);
	case 4: return launch_forward<JAC, 4>(p, lds, stream);
	case 5: return launch_forward<JAC, 5>(p, lds, stream);
	case 6: return launch_forward<JAC, 6>(p, lds, stream);
	case 7: return launch_forward<JAC, 7>(p, lds, stream);
	case 8: return launch_forward<JAC, 8>(p, lds, stream);
	default: return launch_forward<JAC, 0>(p, lds, stream);
	}
}

hipError_t launch_lutmodel_forward(const LutModelParams& p, hipStream_t stream)
{
	if (p.n <= 0) return hipSuccess;
	const size_t lds = lutmodel_lds_bytes(p, false);
	return p.jac ? launch_forward_ovs<true>(p, lds, stream) : launch_forward_ovs<false>(p, lds, stream);
}

template <int OVS>
static hipError_t launch_backward(const LutModelParams& p, size_t lds, hipStream_t stream)
{
	hipError_t e = set_lds((const void*)lutmodel_backward_kernel<OVS>, lds);
	if (e != hipSuccess) return e;
	lutmodel_backward_kernel<OVS><<<p.n, kModelBlock, lds, stream>>>(p);
	return hipGetLastError();
}

hipError_t launch_lutmodel_backward(const LutModelParams& p, hipStream_t stream)
{
	if (p.n <= 0) return hipSuccess;
	const size_t lds = lutmodel_lds_bytes(p, false);
	switch (p.ovs) {
	case 1: return launch_backward<1>(p, lds, stream);
	case 2: return launch_backward<2>(p, lds, stream);
	case 3: return launch_backward<3>(p, lds, stream);
	case 4: return launch_backward<4>(p, lds, stream);
	case 5: return launch_backward<5>(p, lds, stream);
	case 6: return launch_backward<6>(p, lds, stream);
	case 7: return launch_backward<7>(p, lds, stream);
	case 8: return launch_backward<8>(p, lds, stream);
	default: return launch_backward<0>(p, lds, stream);
	}
}

template <int OVS, int MAXB>
static hipError_t launch_fit(const LutModelParams& p, size_t lds, hipStream_t stream, int block)
{
	static_assert(OVS >= 1 && OVS <= kMaxFitOvs, "sub-sample count");
	hipError_t e = set_lds((const void*)lutmodel_fit_kernel<OVS, MAXB>, lds);
	if (e != hipSuccess) return e;
	lutmodel_fit_kernel<OVS, MAXB><<<p.n, block, lds, stream>>>(p);
	return hipGetLastError();
}

template <int MAXB>
static hipError_t launch_fit_ovs(const LutModelParams& p, size_t lds, hipStream_t stream, int block)
{
	switch (p.ovs) {
	case 1: return launch_fit<1, MAXB>(p, lds, stream, block);
	case 2: return launch_fit<2, MAXB>(p, lds, stream, block);
	case 3: return launch_fit<3, MAXB>(p, lds, stream, block);
	case 4: return launch_fit<4, MAXB>(p, lds, stream, block);
	case 5: return launch_fit<5, MAXB>(p, lds, stream, block);
	default: return launch_fit<6, MAXB>(p, lds, stream, block);
	}
}

bool lutmodel_fit_block_ok(int block, int ovs)
{
	if (block == 64 || block == 128 || block == 256) return true;
	// 8 wavefronts per workgroup leave 256 VGPRs per lane, 16 leave 128: the 1024-lane instance of
	// 4 x 4 sub-samples spills 23 VGPRs around the FP64 solve, outside the pixel loop; larger ones
	// would spill inside it
	return (block == 512 && ovs <= kMaxWideFitOvs) || (block == 1024 && ovs <= kMax1024FitOvs);
}

hipError_t launch_lutmodel_fit(const LutModelParams& p, hipStream_t stream, int block)
{
	if (p.n <= 0) return hipSuccess;
	if (p.ovs < 1 || p.ovs > kMaxFitOvs || p.steps < 2) return hipErrorInvalidValue;
	if (!lutmodel_fit_block_ok(block, p.ovs)) return hipErrorInvalidValue;
	const size_t lds = lutmodel_lds_bytes(p, true, block);
	if (block == 1024) {
		switch (p.ovs) {
		case 1: return launch_fit<1, 1024>(p, lds, stream, block);
		case 2: return launch_fit<2, 1024>(p, lds, stream, block);
		case 3: return launch_fit<3, 1024>(p, lds, stream, block);
		default: return launch_fit<4, 1024>(p, lds, stream, block);
		}
	}
	if (block == 512) return launch_fit_ovs<512>(p, lds, stream, block);
	switch (p.ovs) {
	case 7: return launch_fit<7, 256>(p, lds, stream, block);
	case 8: return launch_fit<8, 256>(p, lds, stream, block);
	default: return launch_fit_ovs<256>(p, lds, stream, block);
	}
}

}  // namespace qtrk
