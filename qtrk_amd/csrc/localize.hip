// localize.hip -- fused per-ROI localization chain for gfx950.
//
// One workgroup (4 wavefronts, 256 lanes) localizes one ROI entirely out of
// LDS: pixel load/convert (+ calibration) -> mean/COM -> QI iterations
// (quadrant profiles, kissfft-ordered radix-4/2 FFT autoconvolution, 7-point
// peak fit) -> radial profile (+ normalisation) -> Z-LUT scores, best plane
// and 7-point FP64 fit.  Nothing but the ROI, the bead's LUT rows and the
// 52-byte result touches HBM.
//
// The arithmetic follows the reference CPU tracker operation for operation
// (this file is built with -ffp-contract=off and IEEE division/sqrt):
//   * quadrant profiles: one lane per (quadrant, radius) summing its angular
//     samples in the reference order in float (cputrack/cpu_tracker.cpp:631-677);
//   * the QI FFTs reproduce kissfft's decimation-in-time butterflies
//     (cputrack/kissfft.h:78-152) with one lane per butterfly;
//   * the peak fits are the reference's weighted least-squares solve
//     (cputrack/LsqQuadraticFit.h:18-24, 64-102, 132-176);
//   * LUT scores: one lane per plane, the reference's float difference and
//     double accumulation in radius order (cputrack/cpu_tracker.cpp:829-839).
// Double sums over many pixels/angles (COM moments, radial-profile bins) are
// reduced in a fixed tree instead of left to right; for integer pixel data
// the mean/variance sums are exact either way.
//
// The 2D Gaussian MLE alone on ROIs of up to 1,024 pixels runs instead in
// gauss_wave_kernel, one wavefront per ROI (the same computation, every sum
// inside the wavefront).  Its exp and erf are qmath.h's, the same source the
// CPU oracle compiles, so the model terms round as they do there.
#include "fft_device.hpp"
#include "localize.hpp"
#include "peakfit.hpp"
#define QM_FN __device__ static inline
#include "qmath.h"

#include <cmath>
#include <type_traits>

namespace qtrk {

// LDS map.  The ROI, the radial profile and the reduction scratch live for the
// whole kernel; the QI (or XCor1D) tables/buffers, the radial-profile partial
// sums, the QuadrantAlign buffers and the LUT scores are used in disjoint phases
// and share one region.
struct LdsLayout {
	int img, zprof, red, st;
	int qdir, qprof, fft, tw, perm;               // QI phase
	int xprof, xfft, xtw, xperm, xtmp;            // XCor1D phase
	int zdir, zpart, zcnt;                        // radial-profile phase
	int adir, aprof, alut, afft, atw, aperm, atmp;  // QuadrantAlign phase
	int fbuf, fcol, ftmp, ftwx, ftwy, fpx, fpy;     // 2D power-spectrum phase
	int scores, zw;                               // LUT phase (zw: adjusted weights)
	int gcol;                                     // 2D Gaussian MLE: per-column and per-row terms
	int lutfull;                                  // 1: the whole LUT fits over the image and phases
	int total;
};

// Optional modes the parameters ask for (never in the LUT-build variant).
__host__ __device__ inline bool mode_xcor(const LocalizeParams& P) { return !P.build && (P.mode & kModeXCor1D); }
__host__ __device__ inline bool mode_gauss(const LocalizeParams& P)
{
	return !P.build && !(P.mode & kModeXCor1D) && !(P.mode & kModeQI) && (P.mode & kModeGauss2D);
}
__host__ __device__ inline bool mode_align(const LocalizeParams& P)
{
	return !P.build && (P.mode & kModeZLUTAlign) && (P.mode & kModeLocalizeZ);
}
__host__ __device__ inline bool mode_zweighted(const LocalizeParams& P)
{
	return !P.build && (P.mode & kModeLocalizeZWeighted) && (P.mode & kModeLocalizeZ);
}
__host__ __device__ inline bool mode_fourier(const LocalizeParams& P)
{
	return P.build ? P.fourier != 0 : ((P.mode & kModeFourierLUT) && (P.mode & kModeLocalizeZ));
}
__host__ __device__ inline bool mode_ext(const LocalizeParams& P)
{
	return mode_xcor(P) || mode_align(P) || mode_zweighted(P) || mode_fourier(P);
}

// Bytes per ROI pixel as staged in LDS: the source format when it converts to
// float exactly and no calibration rewrites it (u8, u16), float otherwise.
__host__ __device__ inline int lds_pix_bytes(const LocalizeParams& P) { return P.lds_pix == 0 ? 1 : P.lds_pix == 1 ? 2 : 4; }

__host__ __device__ inline int align16(int x) { return (x + 15) & ~15; }

__host__ __device__ inline int z_chunks(const LocalizeParams& P)
{
	if (P.z_nb <= 0) return 1;
	int c = kBlock / P.z_nb;
	return c < 1 ? 1 : c;
}

// `ext` false leaves the optional-mode phases out (the kernel variant without them).
__host__ __device__ inline LdsLayout lds_layout(const LocalizeParams& P, bool ext = true)
{
	LdsLayout L{};
	int o = 0;
	const int N = P.qi_fft.n > 0 ? P.qi_fft.n : 0;
	// the image, then the phases' scratch (one region, reused phase by phase), then the radial
	// profile and the reduction scratch, which live through the LUT phase
	L.img = o;    o += align16(P.width * P.height * lds_pix_bytes(P));
	const int u = o;
	// QI phase: directions, quadrant profiles, 4 FFT buffers, twiddles, leaf order
	int q = u;
	L.qdir = q;   q += align16(4 * P.qi_A * 8);  // per quadrant, signs applied (instances without the offset table)
	L.qprof = q;  q += align16(4 * P.qi_nr * 4);
	L.fft = q;    q += align16(4 * N * 8);
	L.tw = q;     q += align16(2 * N * 8);
	L.perm = q;   q += align16(N * 4);
	// XCor1D phase: x and y profiles, 4 transforms (+ their generic-radix staging)
	int x = u;
	if (ext && mode_xcor(P)) {
		const int Nx = P.xc_fft.n;
		L.xprof = x;  x += align16(2 * Nx * 4);
		L.xfft = x;   x += align16(4 * Nx * 8);
		L.xtw = x;    x += align16(2 * Nx * 8);
		L.xperm = x;  x += align16(Nx * 4);
		L.xtmp = x;   x += align16(4 * Nx * 8);
	}
	// radial-profile phase
	int z = u;
	const int nch = z_chunks(P);
	L.zdir = z;   z += align16(P.z_A * 8);
	L.zpart = z;  z += align16(nch * P.z_nb * 8);
	L.zcnt = z;   z += align16(nch * P.z_nb * 4);
	// QuadrantAlign phase: directions, quadrant profiles, LUT row, 3 transforms
	int a = u;
	if (ext && mode_align(P)) {
		const int Na = P.qa_fft.n;
		L.adir = a;   a += align16(4 * P.qi_A * 8);
		L.aprof = a;  a += align16(4 * P.z_nb * 4);
		L.alut = a;   a += align16(Na * 4);
		L.afft = a;   a += align16(3 * Na * 8);
		L.atw = a;    a += align16(2 * Na * 8);
		L.aperm = a;  a += align16(Na * 4);
		L.atmp = a;   a += align16(3 * Na * 8);
	}
	// 2D power-spectrum phase: the complex ROI, per-wavefront column buffers and generic-radix
	// staging, the row and column twiddles and leaf orders
	int f = u;
	if (ext && mode_fourier(P)) {
		const int W = P.width, H = P.height, M = W > H ? W : H;
		L.fbuf = f;   f += align16(W * H * 8);
		L.fcol = f;   f += align16(4 * H * 8);
		L.ftmp = f;   f += align16(4 * M * 8);
		L.ftwx = f;   f += align16(W * 8);
		L.ftwy = f;   f += align16(H * 8);
		L.fpx = f;    f += align16(W * 4);
		L.fpy = f;    f += align16(H * 4);
	}
	// 2D Gaussian MLE phase: 4 terms per column, 3 per row
	int g = u;
	if (mode_gauss(P)) {
		L.gcol = g;   g += align16((4 * P.width + 3 * P.height) * 4);
	}
	// LUT phase
	L.scores = u;
	int s = u + align16(P.zplanes * 8);
	L.zw = s;
	if (ext && mode_zweighted(P)) s += align16(P.z_nb * 8);
	o = q > z ? q : z;
	o = o > s ? o : s;
	o = o > x ? o : x;
	o = o > a ? o : a;
	o = o > f ? o : f;
	o = o > g ? o : g;
	// The LUT phase scores every plane in one pass when the bead's whole LUT (rows padded to
	// z_nb + 1 floats) fits over the image and the phase scratch, both dead by then, at one
	// plane per thread -- and when any growth of the region this takes leaves as many
	// workgroups per CU; otherwise it goes through the image's space in tiles.
	const int tail = align16(P.z_nb * 4) + align16(64 * 8) + align16(16 * 4);
	const int need = P.zplanes * (P.z_nb + 1) * 4;
	L.lutfull = 0;
	if (!P.build && P.zluts && P.zplanes > 0 && P.zplanes <= kBlock && (P.z_nb & 3) == 0) {
		const int grown = align16(o > need ? o : need);
		if (kLdsPerCU / (grown + tail) >= kLdsPerCU / (o + tail)) {
			L.lutfull = 1;
			o = grown;
		}
	}
	L.zprof = o;  o += align16(P.z_nb * 4);
	L.red = o;    o += align16(64 * 8);
	L.st = o;     o += align16(16 * 4);
	L.total = o;
	return L;
}

size_t localize_lds_bytes(const LocalizeParams& p) { return (size_t)lds_layout(p).total; }


// Truncation toward zero with the hardware's defined saturation (out-of-range
// values clamp to INT_MIN/INT_MAX, NaN gives 0); unlike a C++ cast it has no
// undefined behaviour the optimiser could exploit.
__device__ __forceinline__ int trunc_sat(float x)
{
	int r;
	asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
	return r;
}

// a * b + c for 0 <= a, b < 2^24, as one instruction; the compiler otherwise
// scales b * a and c separately when the sum indexes a float array.  b is
// wavefront-uniform.
__device__ __forceinline__ int mad_u24(int a, int b, int c)
{
	int r;
	asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
	return r;
}

// A float2 of a table at a per-lane byte offset plus a wavefront-uniform one, as one buffer
// load: the uniform part is the scalar offset (a global load would form a 64-bit address per
// lane and load); reads past the table's end return zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const void* base, int bytes)
{
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ float2 qi_offset(__amdgpu_buffer_rsrc_t table, unsigned lane_off, int soff)
{
	const auto v = __builtin_amdgcn_raw_buffer_load_b64(table, lane_off, __builtin_amdgcn_readfirstlane(soff), 0);
	return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}

// Bilinear samples with the reference's "outside" rule (cputrack/utils.h:39-58):
// a sample is inside iff trunc(x) is in [0, W-2] and trunc(y) in [0, H-2] (the
// saturating conversion keeps huge coordinates out of range); the caller passes
// ok = false for a NaN centre, whose samples the reference's x86 conversion also
// puts outside.  Outside samples read pixel (0,0) and are masked to +0.0, which
// leaves the sum unchanged (it starts at +0.0, so it never holds -0.0) and keeps
// every load unconditional; the additions run in the reference's order.
template <typename T, typename Acc, int U>
__device__ __forceinline__ void sample_n_acc(const T* img, int W, float Wf, float Hf, unsigned Wm2, unsigned Hm2,
                                             bool ok, const float (&x)[U], const float (&y)[U], Acc& sum, int& cnt)
{
	int in[U], rx[U], ry[U];
	float v00[U], v10[U], v01[U], v11[U];
#pragma unroll
	for (int u = 0; u < U; u++) {
		rx[u] = trunc_sat(x[u]);
		ry[u] = trunc_sat(y[u]);
		in[u] = ok && (unsigned)rx[u] <= Wm2 && (unsigned)ry[u] <= Hm2;
	}
#pragma unroll
	for (int u = 0; u < U; u++) {
		// ROIs are far below 2^24 pixels: a 24-bit multiply-add forms the index, which
		// is then masked to 0 outside (no branch; the product of out-of-range
		// coordinates is discarded)
		const T* p = img + (mad_u24(ry[u], W, rx[u]) & -in[u]);
		v00[u] = (float)p[0];
		v10[u] = (float)p[1];
		v01[u] = (float)p[W];
		v11[u] = (float)p[W + 1];
	}
#pragma unroll
	for (int u = 0; u < U; u++) {
		const float fx = x[u] - (float)rx[u];
		const float a0 = v00[u] + (v10[u] - v00[u]) * fx;
		const float a1 = v01[u] + (v11[u] - v01[u]) * fx;
		const float v = a0 + (a1 - a0) * (y[u] - (float)ry[u]);
		sum += (Acc)__int_as_float(__float_as_int(v) & -in[u]);
		cnt += in[u];
	}
}

// The same samples when every one is known to be inside (the caller checked the circle
// against the ROI): no range test, no masking; the sum runs in the same order.  Inside,
// x, y >= 1, where x - (float)(int)x is x - floor(x), exactly the hardware fract (the
// difference is a multiple of ulp(x) below 1, so it is representable), and the pixel
// index is one 24-bit multiply-add.
template <typename T, typename Acc, int U>
__device__ __forceinline__ void sample_n_inside(const T* img, int W, const float (&x)[U], const float (&y)[U], Acc& sum)
{
	int idx[U];
	float v00[U], v10[U], v01[U], v11[U];
#pragma unroll
	for (int u = 0; u < U; u++) idx[u] = mad_u24(trunc_sat(y[u]), W, trunc_sat(x[u]));
#pragma unroll
	for (int u = 0; u < U; u++) {
		const T* p = img + idx[u];
		v00[u] = (float)p[0];
		v10[u] = (float)p[1];
		v01[u] = (float)p[W];
		v11[u] = (float)p[W + 1];
	}
#pragma unroll
	for (int u = 0; u < U; u++) {
		const float fx = __builtin_amdgcn_fractf(x[u]);
		const float a0 = v00[u] + (v10[u] - v00[u]) * fx;
		const float a1 = v01[u] + (v11[u] - v01[u]) * fx;
		sum += (Acc)(a0 + (a1 - a0) * __builtin_amdgcn_fractf(y[u]));
	}
}

// ROI staging.  A contiguous ROI (rows `pitch` = W pixels apart) at a 16-byte
// aligned address without calibration is read in 16-byte chunks, up to 8 per
// thread in flight, and written to LDS in 16-byte stores; otherwise pixels are
// read one by one, 8 per thread in flight.  Either way each pixel's value is
// the reference's: the source value as float, 0 for the first four with
// ClearFirstFourPixels, then offset/gain with their unit factors.
template <typename PT, typename SPix>
__device__ __forceinline__ void stage_roi(const uint8_t* base, int pitch, int W, int npx, const float* off,
                                          const float* gain, bool clear4, SPix* img, int tid)
{
	constexpr int K = 8;
	if constexpr (sizeof(SPix) >= sizeof(PT)) {  // (the LDS format is never narrower than the source)
	if (!off && !gain && pitch == W * (int)sizeof(PT) && ((uintptr_t)base & 15) == 0) {
		constexpr int PPC = 16 / sizeof(PT);  // pixels per chunk
		const int nchunk = npx / PPC;
		const uint4* src = (const uint4*)base;
		for (int c0 = tid; c0 < nchunk; c0 += K * kBlock) {
			uint4 raw[K];
#pragma unroll
			for (int k = 0; k < K; k++) {
				const int c = c0 + k * kBlock;
				raw[k] = c < nchunk ? src[c] : make_uint4(0, 0, 0, 0);
			}
#pragma unroll
			for (int k = 0; k < K; k++) {
				const int c = c0 + k * kBlock;
				if (c < nchunk) {
					PT px[PPC];
					__builtin_memcpy(px, &raw[k], 16);
					SPix out[PPC];
#pragma unroll
					for (int j = 0; j < PPC; j++) {
						float v = (float)px[j];
						if (clear4 && c * PPC + j < 4) v = 0.0f;
						out[j] = (SPix)v;  // exact: SPix is float or the source type
					}
					uint4 o[PPC * sizeof(SPix) / 16];
					__builtin_memcpy(o, out, sizeof(o));
					uint4* dst = (uint4*)(img + c * PPC);
#pragma unroll
					for (int q = 0; q < (int)(PPC * sizeof(SPix) / 16); q++) dst[q] = o[q];
				}
			}
		}
		for (int i = nchunk * PPC + tid; i < npx; i += kBlock) {
			float v = (float)((const PT*)base)[i];
			if (clear4 && i < 4) v = 0.0f;
			img[i] = (SPix)v;
		}
		return;
	}
	}
	const float invW = 1.0f / (float)W;
	for (int i0 = tid; i0 < npx; i0 += K * kBlock) {
		PT v[K];
		float o[K], g[K];
#pragma unroll
		for (int k = 0; k < K; k++) {
			// unconditional loads (past the end: the last pixel again), all in flight together
			const int i = min(i0 + k * kBlock, npx - 1);
			// row = i / W: the float estimate is within one row, then corrected
			int y = (int)((float)i * invW);
			int x = i - y * W;
			const int lo = x < 0, hi = x >= W;
			y += hi - lo;
			x += (lo - hi) * W;
			v[k] = ((const PT*)(base + (int64_t)y * pitch))[x];
			o[k] = off ? off[i] : 0.0f;
			g[k] = gain ? gain[i] : 0.0f;
		}
#pragma unroll
		for (int k = 0; k < K; k++) {
			const int i = i0 + k * kBlock;
			if (i < npx) {
				float t = (float)v[k];
				if (clear4 && i < 4) t = 0.0f;
				if (off && !gain) t = t + o[k] * 1.0f;
				else if (gain && !off) t = t * g[k] * 1.0f;
				else if (gain && off) t = (t + o[k] * 1.0f) * g[k] * 1.0f;
				img[i] = (SPix)t;  // exact: SPix is float whenever t is not an integer of the source range
			}
		}
	}
}

// True when every sample cx + d.x r, cy + d.y r (|d.x|, |d.y| <= 1, float arithmetic) of a
// circle of radius |r| truncates into [0, W-2] x [0, H-2]: |fl(d r)| <= |r| and rounding is
// monotonic, so the extremes are fl(c -+ |r|), kept at least 1 px from the low edges and
// within W-2, H-2 at the high ones.  False for a NaN centre or radius.
__device__ __forceinline__ bool circle_inside(float cx, float cy, float r, float Wf, float Hf)
{
	const float a = fabsf(r);
	return cx - a >= 1.0f && cx + a <= Wf - 2.0f && cy - a >= 1.0f && cy + a <= Hf - 2.0f;
}

// Interpolate (cputrack/utils.h:40-58): the reference's int conversion truncates toward zero
// (and gives INT_MIN, so "outside", for NaN and out-of-range values); 0 outside the ROI.
template <typename SPix>
__device__ __forceinline__ float interpolate_ref(const SPix* img, int W, int H, float x, float y)
{
	if (!(x > -2147483648.0f && x < 2147483648.0f && y > -2147483648.0f && y < 2147483648.0f)) return 0.0f;
	const int rx = (int)x, ry = (int)y;
	if (rx < 0 || ry < 0 || rx + 1 >= W || ry + 1 >= H) return 0.0f;
	const float v00 = (float)img[W * ry + rx], v10 = (float)img[W * ry + rx + 1];
	const float v01 = (float)img[W * (ry + 1) + rx], v11 = (float)img[W * (ry + 1) + rx + 1];
	const float fx = x - (float)rx, fy = y - (float)ry;
	const float v0 = v00 + (v10 - v00) * fx, v1 = v01 + (v11 - v01) * fx;
	return v0 + (v1 - v0) * fy;
}

__device__ __forceinline__ void wave_argmax(float& v, int& i)
{
	for (int off = 32; off > 0; off >>= 1) {
		const float ov = __shfl_xor(v, off);
		const int oi = __shfl_xor(i, off);
		if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
	}
}
__device__ __forceinline__ void wave_argmax(double& v, int& i)
{
	for (int off = 32; off > 0; off >>= 1) {
		const double ov = __shfl_xor(v, off);
		const int oi = __shfl_xor(i, off);
		if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
	}
}

// Sum of K doubles over the workgroup; result in red[32..32+K) for all threads.  The K
// butterfly reductions advance level by level, so their shuffles are in flight together.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* red, int lane, int wid)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
#pragma unroll
		for (int k = 0; k < K; k++) v[k] += __shfl_xor(v[k], off);
	if (lane == 0)
		for (int k = 0; k < K; k++) red[wid * K + k] = v[k];
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int k = 0; k < K; k++) {
			double s = red[k];
			for (int w = 1; w < kBlock / 64; w++) s += red[w * K + k];
			red[32 + k] = s;
		}
	}
	__syncthreads();
}

// kissfft's stages from the deepest (smallest) to the outermost, one lane per
// butterfly, `nbufs` transforms of length F.n side by side (one per wavefront).
// MIXED adds the radix-3, radix-5 and generic butterflies
// (cputrack/kissfft.h:154-290); without it only radix 4 and 2 occur (QI's
// power-of-two lengths).  A generic radix-p stage writes each output, a length-p
// DFT of the group's inputs, into `tmp` and copies the stage back afterwards.
template <bool MIXED>
__device__ __forceinline__ void fft_stages(float2* base, float2* tmp, int nbufs, const FFTPlan& F, const float2* tw,
                                           float nii, int wid, int lane)
{
	const int N = F.n;
#pragma unroll
	for (int si = 0; si < kMaxFFTStages; si++) {
		const int s = kMaxFFTStages - 1 - si;
		if (s >= F.nst) continue;  // uniform: every thread skips together
		const int r = F.radix[s], m = F.m[s], fs = F.fs[s];
		const bool generic = MIXED && r != 2 && r != 3 && r != 4 && r != 5;
		if (wid < nbufs) {
			float2* X = base + wid * N;
			if (!generic) {
				const int nbf = N / r;
				for (int b = lane; b < nbf; b += 64) {
					const int g = b / m, k = b - g * m;
					float2* G = X + g * r * m;
					if (r == 4) {
						const float2 s0 = cmul(G[k + m], tw[k * fs]);
						const float2 s1 = cmul(G[k + 2 * m], tw[k * fs * 2]);
						const float2 s2 = cmul(G[k + 3 * m], tw[k * fs * 3]);
						float2 f0 = G[k];
						const float2 s5 = csub(f0, s1);
						f0 = cadd(f0, s1);
						const float2 s3 = cadd(s0, s2);
						float2 s4 = csub(s0, s2);
						s4 = make_float2(s4.y * nii, -s4.x * nii);
						G[k + 2 * m] = csub(f0, s3);
						G[k] = cadd(f0, s3);
						G[k + m] = cadd(s5, s4);
						G[k + 3 * m] = csub(s5, s4);
					} else if (!MIXED || r == 2) {
						const float2 t = cmul(G[m + k], tw[k * fs]);
						const float2 f0 = G[k];
						G[m + k] = csub(f0, t);
						G[k] = cadd(f0, t);
					} else if (r == 3) {
						// HALF_OF's product with the double 0.5 rounds like a product with 0.5f
						const float2 e = tw[fs * m];
						const float2 s1 = cmul(G[k + m], tw[k * fs]);
						const float2 s2 = cmul(G[k + 2 * m], tw[2 * k * fs]);
						const float2 s3 = cadd(s1, s2);
						float2 s0 = csub(s1, s2);
						const float2 f0 = G[k];
						const float2 fm = make_float2(f0.x - s3.x * 0.5f, f0.y - s3.y * 0.5f);
						s0 = make_float2(s0.x * e.y, s0.y * e.y);
						G[k] = cadd(f0, s3);
						G[k + 2 * m] = make_float2(fm.x + s0.y, fm.y - s0.x);
						G[k + m] = cadd(fm, make_float2(-s0.y, s0.x));
					} else {
						const float2 ya = tw[fs * m], yb = tw[fs * 2 * m];
						const float2 s0 = G[k];
						const float2 s1 = cmul(G[k + m], tw[k * fs]);
						const float2 s2 = cmul(G[k + 2 * m], tw[2 * k * fs]);
						const float2 s3 = cmul(G[k + 3 * m], tw[3 * k * fs]);
						const float2 s4 = cmul(G[k + 4 * m], tw[4 * k * fs]);
						const float2 s7 = cadd(s1, s4), s10 = csub(s1, s4), s8 = cadd(s2, s3), s9 = csub(s2, s3);
						G[k] = cadd(cadd(s0, s7), s8);
						const float2 s5 = cadd(s0, make_float2(s7.x * ya.x + s8.x * yb.x, s7.y * ya.x + s8.y * yb.x));
						const float2 s6 = make_float2(s10.y * ya.y + s9.y * yb.y, -(s10.x * ya.y) - s9.x * yb.y);
						G[k + m] = csub(s5, s6);
						G[k + 4 * m] = cadd(s5, s6);
						const float2 s11 = cadd(s0, make_float2(s7.x * yb.x + s8.x * ya.x, s7.y * yb.x + s8.y * ya.x));
						const float2 s12 = make_float2(-(s10.y * yb.y) + s9.y * ya.y, s10.x * yb.y - s9.x * ya.y);
						G[k + 2 * m] = cadd(s11, s12);
						G[k + 3 * m] = csub(s11, s12);
					}
				}
			} else {
				float2* T = tmp + wid * N;
				const int gl = r * m;
				for (int t = lane; t < N; t += 64) {
					const int g = t / gl, kk = t - g * gl, u = kk % m;
					const float2* G = X + g * gl;
					float2 acc = G[u];
					int ti = 0;
					for (int q = 1; q < r; q++) {
						ti += fs * kk;
						if (ti >= N) ti -= N;
						acc = cadd(acc, cmul(G[u + q * m], tw[ti]));
					}
					T[t] = acc;
				}
			}
		}
		__syncthreads();
		if (generic) {
			if (wid < nbufs)
				for (int t = lane; t < N; t += 64) base[wid * N + t] = tmp[wid * N + t];
			__syncthreads();
		}
	}
}

// Orders the LDS accesses of one wavefront: what its lanes wrote before is seen by its
// lanes after (a wavefront's LDS instructions execute in order; this keeps the compiler
// from moving them across).
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One power-of-two transform (QI's lengths: radices 4 and 2, group and twiddle indices by
// shifts and masks) by one wavefront in place, its stages ordered by wave_sync: no
// workgroup barrier.  The butterflies, and so the rounding, are those of fft_stages.
__device__ __forceinline__ void fft_pow2_wave(float2* X, const FFTPlan& F, const float2* tw, float nii, int lane)
{
	const int ln = __builtin_ctz(F.n);
#pragma unroll
	for (int si = 0; si < kMaxFFTStages; si++) {
		const int s = kMaxFFTStages - 1 - si;
		if (s >= F.nst) continue;  // uniform
		const int r = F.radix[s];
		const int lr = r == 4 ? 2 : 1, lm = __builtin_ctz(F.m[s]), lf = __builtin_ctz(F.fs[s]);
		const int nb = 1 << (ln - lr);  // butterflies
		for (int e = lane; e < nb; e += 64) {
			const int g = e >> lm, k = e & ((1 << lm) - 1);
			bfly_pow2(X + (g << (lm + lr)), r, k, 1 << lm, k << lf, tw, nii);
		}
		wave_sync();
	}
}

// NormalizeRadialProfile (cputrack/utils.cpp:230-261) by one thread: double sums,
// float mean and deviations, 1/sqrt of the double mean square.
// The sums run in index order; the values are read 8 at a time ahead of them,
// so that the serial additions do not wait on each LDS read.
__device__ __forceinline__ void normalize_profile(float* p, int n)
{
	constexpr int K = 8;
	double s = 0.0;
	for (int i0 = 0; i0 < n; i0 += K) {
		float v[K];
#pragma unroll
		for (int j = 0; j < K; j++) v[j] = p[min(i0 + j, n - 1)];
#pragma unroll
		for (int j = 0; j < K; j++)
			if (i0 + j < n) s += (double)v[j];
	}
	const float pm = (float)(s / (double)n);
	double s2 = 0.0;
	for (int i0 = 0; i0 < n; i0 += K) {
		float v[K];
#pragma unroll
		for (int j = 0; j < K; j++) v[j] = p[min(i0 + j, n - 1)] - pm;
#pragma unroll
		for (int j = 0; j < K; j++)
			if (i0 + j < n) {
				p[i0 + j] = v[j];
				s2 += (double)(float)(v[j] * v[j]);
			}
	}
	const double inv = (double)1.0f / sqrt(s2 / (double)n);
	for (int i = 0; i < n; i++) p[i] = (float)((double)p[i] * inv);
}

// Whether a double sum of the floats v (n of them, lg >= log2 n) is exact in any order: every
// partial sum is a multiple of the smallest ulp among them and below n times twice the largest,
// which fits a double's 53 bits when the exponents span at most 29 - lg (one bit of margin
// here).  emin/emax: biased exponents of the nonzero values (denormals count as 1); fin: all
// finite.  Then a tree sum equals the reference's sequential one bit for bit.
__device__ __forceinline__ bool exact_any_order(int emin, int emax, bool fin, int lg)
{
	return fin && (emax < emin || emax - emin <= 28 - lg);
}
__device__ __forceinline__ void exp_range(float v, int& emin, int& emax, bool& fin)
{
	const int e = (int)((__float_as_uint(v) >> 23) & 255);
	fin = fin && e != 255;
	if (v != 0.0f) {
		emin = min(emin, e > 0 ? e : 1);
		emax = max(emax, e);
	}
}
__device__ __forceinline__ double wave_sum(double v)
{
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
	return v;
}
__device__ __forceinline__ int wave_min(int v)
{
	for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
	return v;
}
__device__ __forceinline__ int wave_max(int v)
{
	for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
	return v;
}

// NormalizeRadialProfile by one wavefront, bit-identical to normalize_profile: the mean's double
// sum as a tree when exact_any_order holds (else lane 0 adds in index order), the deviations
// and their float squares per lane, the squares' double sum likewise, the scaling per lane.
__device__ __forceinline__ void normalize_profile_wave(float* p, int n, int lane)
{
	int lg = 0;
	while ((1 << lg) < n) lg++;
	int emin = 255, emax = 0;
	bool fin = true;
	double part = 0.0;
	for (int i = lane; i < n; i += 64) {
		const float v = p[i];
		exp_range(v, emin, emax, fin);
		part += (double)v;
	}
	emin = wave_min(emin);
	emax = wave_max(emax);
	fin = __all(fin);
	double sum;
	if (exact_any_order(emin, emax, fin, lg)) {
		sum = wave_sum(part);
	} else {
		sum = 0.0;
		if (lane == 0)
			for (int i = 0; i < n; i++) sum += (double)p[i];
		sum = __shfl(sum, 0);
	}
	const float pm = (float)(sum / (double)n);
	emin = 255;
	emax = 0;
	fin = true;
	part = 0.0;
	for (int i = lane; i < n; i += 64) {
		const float d = p[i] - pm;
		const float q = d * d;
		p[i] = d;
		exp_range(q, emin, emax, fin);
		part += (double)q;
	}
	emin = wave_min(emin);
	emax = wave_max(emax);
	fin = __all(fin);
	double s2;
	if (exact_any_order(emin, emax, fin, lg)) {
		s2 = wave_sum(part);
	} else {
		wave_sync();
		s2 = 0.0;
		if (lane == 0) {
			constexpr int K = 8;
			for (int i0 = 0; i0 < n; i0 += K) {
				float v[K];
#pragma unroll
				for (int j = 0; j < K; j++) v[j] = p[min(i0 + j, n - 1)];
#pragma unroll
				for (int j = 0; j < K; j++)
					if (i0 + j < n) s2 += (double)(float)(v[j] * v[j]);
			}
		}
		s2 = __shfl(s2, 0);
	}
	const double inv = (double)1.0f / sqrt(s2 / (double)n);
	for (int i = lane; i < n; i += 64) p[i] = (float)((double)p[i] * inv);
}

// QI profile concatenations (cputrack/cpu_tracker.cpp:327-343):
// X = [rev(q1+q2), q0+q3], Y = [rev(q2+q3), q0+q1].
__device__ __forceinline__ float qi_input(const float* qp, int nr, int axis, int n)
{
	const float* q0 = qp;
	const float* q1 = qp + nr;
	const float* q2 = qp + 2 * nr;
	const float* q3 = qp + 3 * nr;
	if (n < nr) {
		const int r = nr - 1 - n;
		return axis == 0 ? q1[r] + q2[r] : q2[r] + q3[r];
	}
	const int r = n - nr;
	return axis == 0 ? q0[r] + q3[r] : q0[r] + q1[r];
}

// Spectrum products a * conj(b) for two inverse transforms, written in leaf order
// into buffers 0 and 1: buffer pairs (a0, b0) and (a1, b1).  Every thread reads
// before any writes (at most 2N / kBlock <= 4 values per thread in registers).
__device__ __forceinline__ void spectrum_products(float2* fft, const int* perm, int N, int a0, int b0, int a1, int b1,
                                                  int tid)
{
	float2 prod[4];
#pragma unroll
	for (int i = 0; i < 4; i++) {
		const int t = tid + i * kBlock;
		if (t < 2 * N) {
			const int ax = t >= N ? 1 : 0;
			const int s = perm[t - ax * N];
			const float2 a = fft[(ax ? a1 : a0) * N + s];
			const float2 b = fft[(ax ? b1 : b0) * N + s];
			prod[i] = cmul(a, make_float2(b.x, -b.y));
		}
	}
	__syncthreads();
#pragma unroll
	for (int i = 0; i < 4; i++) {
		const int t = tid + i * kBlock;
		if (t < 2 * N) fft[t] = prod[i];
	}
	__syncthreads();
}

// The real part of G rotated by `shift` (out[x] = G[(x + shift) % N].x), its first
// maximum over the wavefront and the 7-point QI peak fit around it
// (cputrack/LsqQuadraticFit.h:132-176).  The result is valid in lane 0.
__device__ __forceinline__ float rotated_peak(const float2* G, int N, int shift, int lane)
{
	auto at = [&](int x) {
		int i = x + shift;
		if (i >= N) i -= N;
		return G[i].x;
	};
	float bv = 0.0f;
	int bi = 0x7fffffff;
	for (int x = lane; x < N; x += 64) {
		const float v = at(x);
		if (bi == 0x7fffffff || v > bv) { bv = v; bi = x; }
	}
	if (bi == 0x7fffffff) bv = -INFINITY;
	wave_argmax(bv, bi);
	if (bi == 0x7fffffff) bi = 0;  // only when every value is NaN: keep the window in bounds
	return lane == 0 ? peak_fit<float>(bi, N, kQIW, at) : 0.0f;
}

// One term of a LUT score (cputrack/cpu_tracker.cpp:829-836): d is the float difference in double;
// s - (w d)^2 with radial weights w, else s - d^2 as one fused multiply-add, which rounds as the
// separate product and sum do because d^2 (a 24-bit significand squared) is exact in double.
__device__ __forceinline__ double zcmp_step(double s, double d, const float* w, int r)
{
	if (w) {
		d *= (double)w[r];
		d = -d * d;
		return s + d;
	}
	return __builtin_fma(-d, d, s);
}

// Best plane (first maximum) of the LUT scores over the workgroup and the FP64
// 7-point fit around it (cputrack/cpu_tracker.cpp:851-861); valid in thread 0.
__device__ __forceinline__ float scores_peak(const double* scores, int planes, double* red, float* st, int tid,
                                             int lane, int wid, int& best)
{
	double bv = 0.0;
	int bi = 0x7fffffff;
	for (int k = tid; k < planes; k += kBlock) {
		const double v = scores[k];
		if (bi == 0x7fffffff || v > bv) { bv = v; bi = k; }
	}
	if (bi == 0x7fffffff) bv = -INFINITY;
	wave_argmax(bv, bi);
	if (lane == 0) {
		red[2 * wid] = bv;
		((int*)st)[4 + wid] = bi;
	}
	__syncthreads();
	float z = 0.0f;
	if (tid == 0) {
		double v = red[0];
		int i = ((int*)st)[4];
		for (int w = 1; w < kBlock / 64; w++) {
			const double ov = red[2 * w];
			const int oi = ((int*)st)[4 + w];
			if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
		}
		z = (float)peak_fit<double>(i, planes, kZW, [&](int k) { return scores[k]; });
		best = i;
	}
	return z;
}

// ComputeXCorInterpolated (cputrack/cpu_tracker.cpp:190-270, XCorScale = 1).  Per
// iteration an x profile (xc_len samples along x, each the mean of the in-bounds
// samples over xc_width rows around cy) and a y profile, both taken from the
// centre at the start of the iteration; each is autocorrelated with its reverse
// (XCorFFTHelper, :140-158) on four wavefronts like QI, the 7-point fit finds the
// peak, and the centre moves by (offset - 1) / 2.  A profile longer than the ROI
// leaves the start in place with the boundary flag set.
template <typename SPix>
__device__ __forceinline__ void xcor1d_phase(const LocalizeParams& P, const LdsLayout& L, unsigned char* lds, const SPix* img,
                             float& cx, float& cy, int& boundary, int tid, int lane, int wid)
{
	const int W = P.width, H = P.height, N = P.xc_len;
	if (N > W || N > H) {
		boundary = 1;
		return;
	}
	const int pw = P.xc_width < N ? P.xc_width : N;
	float* xprof = (float*)(lds + L.xprof);
	float2* fft = (float2*)(lds + L.xfft);
	float2* tw = (float2*)(lds + L.xtw);
	int* perm = (int*)(lds + L.xperm);
	float2* tmp = (float2*)(lds + L.xtmp);
	float* st = (float*)(lds + L.st);
	const float Wf = (float)W, Hf = (float)H;
	const unsigned Wm2 = (unsigned)(W - 2), Hm2 = (unsigned)(H - 2);
	for (int i = tid; i < N; i += kBlock) {
		tw[i] = P.xc_fft.tw_fwd[i];
		tw[N + i] = P.xc_fft.tw_inv[i];
		perm[i] = P.xc_fft.perm[i];
	}
	boundary = 0;
	const float half = (1.0f * (float)N) / 2;
	const int ph = pw / 2;
	for (int k = 0; k < P.xc_iters; k++) {
		boundary = (cx + half >= Wf) || (cx - half < 0.0f) || (cy + half >= Hf) || (cy - half < 0.0f);
		const float xmin = cx - half, ymin = cy - half;
		const bool ok = !isnan(cx) && !isnan(cy);
		// slot t < N: x profile sample t; t >= N: y profile sample t - N
		for (int t = tid; t < 2 * N; t += kBlock) {
			const int axis = t >= N ? 1 : 0, i = t - axis * N;
			const float along = (float)i * 1.0f + (axis ? ymin : xmin);
			const float c = axis ? cx : cy;
			float sum = 0.0f;
			int cnt = 0;
			int j = 0;
			for (; j + 4 <= pw; j += 4) {
				float x[4], y[4];
#pragma unroll
				for (int u = 0; u < 4; u++) {
					const float across = c + 1.0f * (float)(j + u - ph);
					x[u] = axis ? across : along;
					y[u] = axis ? along : across;
				}
				sample_n_acc<SPix, float, 4>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
			}
			for (; j < pw; j++) {
				const float across = c + 1.0f * (float)(j - ph);
				const float x[1] = { axis ? across : along }, y[1] = { axis ? along : across };
				sample_n_acc<SPix, float, 1>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
			}
			xprof[t] = cnt > 0 ? sum / (float)cnt : 0.0f;
		}
		__syncthreads();
		// leaves of the 4 forward transforms: x, reversed x, y, reversed y
		{
			float2* F = fft + wid * N;
			const float* src = xprof + (wid >> 1) * N;
			for (int j = lane; j < N; j += 64) {
				int s = perm[j];
				if (wid & 1) s = N - 1 - s;
				F[j] = make_float2(src[s], 0.0f);
			}
		}
		__syncthreads();
		fft_stages<true>(fft, tmp, 4, P.xc_fft, tw, 1.0f, wid, lane);
		spectrum_products(fft, perm, N, 0, 1, 2, 3, tid);
		fft_stages<true>(fft, tmp, 2, P.xc_fft, tw + N, -1.0f, wid, lane);
		if (wid < 2) {
			const float maxPos = rotated_peak(fft + wid * N, N, N / 2, lane);
			if (lane == 0) {
				const float off = maxPos - (float)N / 2;
				st[wid] = (wid == 0 ? cx : cy) + (off - 1) * 1.0f * 0.5f;
			}
		}
		__syncthreads();
		cx = st[0];
		cy = st[1];
		__syncthreads();
	}
}

// QuadrantAlign (cputrack/cpu_tracker.cpp:449-512, offsets :409-438): the bead's LUT
// rows around z interpolated in float into [rev(L), L], the four quadrant profiles
// on the Z-LUT's radii (QI's direction table, no radial weights), each normalised,
// then X = [rev(q1+q2), q0+q3] and Y = [rev(q2+q3), q0+q1] correlated against the
// LUT row: three forward transforms on three wavefronts, two inverse.  The centre
// moves by the 7-point peak offset times pi/4 and the Z-LUT's bin width.
template <typename SPix>
__device__ __forceinline__ void quadrant_align_phase(const LocalizeParams& P, const LdsLayout& L, unsigned char* lds, const SPix* img,
                                     float mean, const float* lut, float z, float& cx, float& cy, int tid, int lane,
                                     int wid)
{
	const int W = P.width, H = P.height, nb = P.z_nb, N = P.qa_fft.n, A = P.qi_A;
	float2* qdir = (float2*)(lds + L.adir);
	float* qprof = (float*)(lds + L.aprof);
	float* lutc = (float*)(lds + L.alut);
	float2* fft = (float2*)(lds + L.afft);
	float2* tw = (float2*)(lds + L.atw);
	int* perm = (int*)(lds + L.aperm);
	float2* tmp = (float2*)(lds + L.atmp);
	float* st = (float*)(lds + L.st);
	const float Wf = (float)W, Hf = (float)H;
	const unsigned Wm2 = (unsigned)(W - 2), Hm2 = (unsigned)(H - 2);
	// rows clamp((int)z) and clamp((int)z + 1); frac = z - (int)z
	const int iz = trunc_sat(z);
	const int izc = iz < -1 ? -1 : (iz > P.zplanes ? P.zplanes : iz);
	const int zp0 = izc < 0 ? 0 : (izc > P.zplanes - 1 ? P.zplanes - 1 : izc);
	const int zp1 = izc + 1 < 0 ? 0 : (izc + 1 > P.zplanes - 1 ? P.zplanes - 1 : izc + 1);
	const float frac = z - (float)iz;
	const float* l0 = lut + (size_t)zp0 * nb;
	const float* l1 = lut + (size_t)zp1 * nb;
	for (int j = tid; j < N; j += kBlock) {
		const int r = j < nb ? nb - 1 - j : j - nb;
		lutc[j] = l0[r] + (l1[r] - l0[r]) * frac;
		tw[j] = P.qa_fft.tw_fwd[j];
		tw[N + j] = P.qa_fft.tw_inv[j];
		perm[j] = P.qa_fft.perm[j];
	}
	for (int j = tid; j < 4 * A; j += kBlock) {
		const int q = j / A;
		const float2 d = P.qi_dirs[j - q * A];
		const float sx = (q == 0 || q == 3) ? 1.0f : -1.0f;
		const float sy = (q < 2) ? 1.0f : -1.0f;
		qdir[j] = make_float2(sx * d.x, sy * d.y);
	}
	__syncthreads();
	const float rstep = (P.z_maxr - P.z_minr) / (float)nb;
	const int steps = P.qa_steps, stride = P.qa_stride;
	const bool ok = !isnan(cx) && !isnan(cy);
	for (int base = 0; base < 4 * nb; base += kBlock) {
		const int p = base + tid;
		if (p < 4 * nb) {
			const int q = p / nb, ri = p - q * nb;
			const float2* qd = qdir + q * A;
			const float r = P.z_minr + rstep * (float)ri;
			float sum = 0.0f;
			int cnt = 0;
			int a = 0;
			for (; a + 4 <= steps; a += 4) {
				float x[4], y[4];
#pragma unroll
				for (int u = 0; u < 4; u++) {
					const float2 d = qd[stride * (a + u)];
					x[u] = cx + d.x * r;
					y[u] = cy + d.y * r;
				}
				sample_n_acc<SPix, float, 4>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
			}
			for (; a < steps; a++) {
				const float2 d = qd[stride * a];
				const float x[1] = { cx + d.x * r }, y[1] = { cy + d.y * r };
				sample_n_acc<SPix, float, 1>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
			}
			qprof[p] = cnt >= 4 ? sum / (float)cnt : mean;
		}
	}
	__syncthreads();
	if (tid < 4) normalize_profile(qprof + tid * nb, nb);
	__syncthreads();
	if (wid < 3) {
		float2* F = fft + wid * N;
		for (int j = lane; j < N; j += 64) {
			const int s = perm[j];
			F[j] = make_float2(wid < 2 ? qi_input(qprof, nb, wid, s) : lutc[s], 0.0f);
		}
	}
	__syncthreads();
	fft_stages<true>(fft, tmp, 3, P.qa_fft, tw, 1.0f, wid, lane);
	spectrum_products(fft, perm, N, 0, 2, 1, 2, tid);
	fft_stages<true>(fft, tmp, 2, P.qa_fft, tw + N, -1.0f, wid, lane);
	const float ppl = (P.z_maxr - P.z_minr) / (float)nb;
	if (wid < 2) {
		const float maxPos = rotated_peak(fft + wid * N, N, nb, lane);
		if (lane == 0) {
			const float off = (maxPos - (float)nb) * (3.14159265359f / 4);
			st[wid] = (wid == 0 ? cx : cy) + off * ppl;
		}
	}
	__syncthreads();
	cx = st[0];
	cy = st[1];
	__syncthreads();
}

// FFT2D::Apply (cputrack/cpu_tracker.cpp:920-952) on the ROI in LDS: kissfft of every row
// (four rows at a time, one per wavefront, written in leaf order straight into the complex
// buffer), then of every column (gathered into per-wavefront buffers in leaf order and
// scattered back), then the squared magnitude written over the ROI with the zero frequency
// moved to (W/2, H/2).
__device__ __forceinline__ void power_spectrum_phase(const LocalizeParams& P, const LdsLayout& L, unsigned char* lds,
                                                     float* img, int tid, int lane, int wid)
{
	const int W = P.width, H = P.height;
	float2* cb = (float2*)(lds + L.fbuf);
	float2* col = (float2*)(lds + L.fcol);
	float2* tmp = (float2*)(lds + L.ftmp);
	float2* twx = (float2*)(lds + L.ftwx);
	float2* twy = (float2*)(lds + L.ftwy);
	int* px = (int*)(lds + L.fpx);
	int* py = (int*)(lds + L.fpy);
	for (int i = tid; i < W; i += kBlock) {
		twx[i] = P.fx_fft.tw_fwd[i];
		px[i] = P.fx_fft.perm[i];
	}
	for (int i = tid; i < H; i += kBlock) {
		twy[i] = P.fy_fft.tw_fwd[i];
		py[i] = P.fy_fft.perm[i];
	}
	__syncthreads();
	for (int y0 = 0; y0 < H; y0 += 4) {
		const int nr = H - y0 < 4 ? H - y0 : 4;
		if (wid < nr) {
			float2* F = cb + (y0 + wid) * W;
			const float* src = img + (y0 + wid) * W;
			for (int j = lane; j < W; j += 64) F[j] = make_float2(src[px[j]], 0.0f);
		}
		__syncthreads();
		fft_stages<true>(cb + y0 * W, tmp, nr, P.fx_fft, twx, 1.0f, wid, lane);
	}
	for (int x0 = 0; x0 < W; x0 += 4) {
		const int nc = W - x0 < 4 ? W - x0 : 4;
		if (wid < nc) {
			float2* F = col + wid * H;
			for (int j = lane; j < H; j += 64) F[j] = cb[py[j] * W + x0 + wid];
		}
		__syncthreads();
		fft_stages<true>(col, tmp, nc, P.fy_fft, twy, 1.0f, wid, lane);
		if (wid < nc)
			for (int y = lane; y < H; y += 64) cb[y * W + x0 + wid] = col[wid * H + y];
		__syncthreads();
	}
	for (int i = tid; i < W * H; i += kBlock) {
		const int y = i / W, x = i - y * W;
		const float2 v = cb[i];
		img[(x + W / 2) % W + ((y + H / 2) % H) * W] = v.x * v.x + v.y * v.y;
	}
	__syncthreads();
}

// EXT compiles in the optional modes (LT_XCor1D, LT_ZLUTAlign, LT_LocalizeZWeighted,
// LT_FourierLUT); the variant without them is the one the common modes run.
// CW, CH: the ROI size as compile-time constants (0: from P), see launch_variant
template <typename SPix, bool EXT, int CW = 0, int CH = 0>
__global__ void __launch_bounds__(kBlock) localize_kernel(LocalizeParams P)
{
	extern __shared__ __align__(16) unsigned char lds[];
	const LdsLayout L = lds_layout(P, EXT);
	SPix* img = (SPix*)(lds + L.img);
	float2* qdir = (float2*)(lds + L.qdir);
	float2* zdir = (float2*)(lds + L.zdir);
	float* qprof = (float*)(lds + L.qprof);
	float2* fft = (float2*)(lds + L.fft);
	float2* tw = (float2*)(lds + L.tw);
	int* perm = (int*)(lds + L.perm);
	float* zprof = (float*)(lds + L.zprof);
	double* zpart = (double*)(lds + L.zpart);
	int* zcnt = (int*)(lds + L.zcnt);
	double* scores = (double*)(lds + L.scores);
	double* red = (double*)(lds + L.red);
	float* st = (float*)(lds + L.st);

	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	// instances for ROIs of 96 px and more read the QI sample offsets from P.qi_offs (buffer loads,
	// prefetched a group of angles ahead) instead of forming them from LDS directions: fewer VALU and
	// LDS instructions per sample for 10 more VGPRs, which these instances have to spare (their LDS
	// already holds a CU to three ROIs), while smaller ones would lose a wavefront per SIMD
	constexpr bool kQiTable = CW >= 96;
	// In the same instances the wavefronts issue at priority 1 from the end of the COM on, except in
	// the two sampling loops (QI and radial profile); the ROI staging, the COM and the sampling loops
	// run at 0.  When a SIMD holds a wavefront of one ROI's dependent
	// stretch (transforms, peak fit, reductions, LUT scores) and one of another ROI's sampling, the
	// first goes first, since the sampling of the other ROI has work to spare and its own ROI waits
	// for it only at the end of the loop.  The other instances compiled for a size (five ROIs per CU)
	// do so only between the QI sampling loops and run at 0 elsewhere: priority 1 outside the QI
	// iterations too measured slower for their consecutive batches.
	constexpr bool kChainPrio = CW >= 96;
	constexpr bool kQiPrio = CW > 0;
	const int roi = blockIdx.x;
	const int W = CW ? CW : P.width, H = CH ? CH : P.height, npx = W * H;
	const bool build = P.build != 0;
	const DevJob job = P.jobs ? P.jobs[roi] : DevJob{};
	const int bead = build ? roi : job.zlutIndex;

	// ---- ROI -> LDS as float (cputrack/cpu_tracker.h:123-136), ClearFirstFourPixels
	// (cputrack/QueuedCPUTracker.cpp:280-282), per-bead offset/gain with unit
	// factors (cputrack/QueuedCPUTracker.cpp:258-271, cputrack/cpu_tracker.cpp:98-112).
	{
		const uint8_t* base = P.src + (P.roi_offset ? P.roi_offset[roi] : (int64_t)roi * H * P.pitch);
		const bool calib = (P.calib_offset || P.calib_gain) && bead >= 0 && bead < P.zcount;
		const float* off = calib && P.calib_offset ? P.calib_offset + (size_t)bead * npx : nullptr;
		const float* gain = calib && P.calib_gain ? P.calib_gain + (size_t)bead * npx : nullptr;
		const bool clear4 = (P.mode & kModeClearFirstFour) && !build;
		if (P.pdt == 1) stage_roi<uint16_t>(base, P.pitch, W, npx, off, gain, clear4, img, tid);
		else if (P.pdt == 0) stage_roi<uint8_t>(base, P.pitch, W, npx, off, gain, clear4, img, tid);
		else stage_roi<float>(base, P.pitch, W, npx, off, gain, clear4, img, tid);
	}
	__syncthreads();
	const float Wf = (float)W, Hf = (float)H;
	const unsigned Wm2 = (unsigned)(W - 2), Hm2 = (unsigned)(H - 2);

	// ---- mean, stdev and background-corrected COM (cputrack/cpu_tracker.cpp:681-729)
	float mean = 0.0f, comx, comy;
	const bool known = build && P.known_pos;
	if (known) {
		comx = P.known_pos[2 * roi];
		comy = P.known_pos[2 * roi + 1];
	} else {
		double m1[2] = { 0.0, 0.0 };
		for (int i = tid; i < npx; i += kBlock) {
			const float v = (float)img[i];
			m1[0] += (double)v;
			m1[1] += (double)(float)(v * v);
		}
		block_sum<2>(m1, red, lane, wid);
		const float invN = 1.0f / (float)npx;
		mean = (float)(red[32] * (double)invN);
		const double var = red[33] * (double)invN - (double)(mean * mean);
		const float stdev = sqrtf((float)var);
		const float bgs = (build ? 0.0f : P.com_bgcorrection) * stdev;
		__syncthreads();
		double m2[3] = { 0.0, 0.0, 0.0 };
		for (int y = wid; y < H; y += kBlock / 64) {
			for (int x = lane; x < W; x += 64) {
				const float d = fabsf((float)img[y * W + x] - mean) - bgs;
				const float v = (0.0f < d) ? d : 0.0f;
				m2[0] += (double)v;
				m2[1] += (double)x * (double)v;
				m2[2] += (double)y * (double)v;
			}
		}
		block_sum<3>(m2, red, lane, wid);
		comx = (float)(red[33] / red[32]);
		comy = (float)(red[34] / red[32]);
		// NaN fallback to the ROI centre (cputrack/QueuedCPUTracker.cpp:294-295); BuildLUT has none.
		if (!build && (isnan(comx) || isnan(comy))) {
			comx = (float)(W / 2);
			comy = (float)(H / 2);
		}
	}

	if constexpr (kChainPrio) __builtin_amdgcn_s_setprio(1);  // staging and COM ran at 0
	// ---- XCor1D, which takes precedence over QI (cputrack/QueuedCPUTracker.cpp:299-303)
	float cx = comx, cy = comy;
	int boundary = 0;
	const bool xcor = EXT && mode_xcor(P);
	if constexpr (EXT) {
		if (xcor) xcor1d_phase<SPix>(P, L, lds, img, cx, cy, boundary, tid, lane, wid);
	}

	// ---- QI (cputrack/cpu_tracker.cpp:285-406)
	const bool do_qi = ((P.mode & kModeQI) && !xcor) || (build && !known);
	if (do_qi && P.qi_iters > 0) {
		const int nr = P.qi_nr, N = P.qi_fft.n;
		const float maxr = P.qi_maxr, minr = P.qi_minr;
		const float rstep = (maxr - minr) / (float)nr;
		const float* window = build ? nullptr : P.qi_window;
		for (int i = tid; i < N; i += kBlock) {
			tw[i] = P.qi_fft.tw_fwd[i];
			tw[N + i] = P.qi_fft.tw_inv[i];
			perm[i] = P.qi_fft.perm[i];
		}
		const int A = P.qi_A;
		const __amdgpu_buffer_rsrc_t qtab = table_rsrc(P.qi_offs, 4 * A * nr * 8);
		if constexpr (!kQiTable) {
			// direction table angle-major, the 4 quadrants' entries of an angle adjacent (a wavefront's
			// 4 quadrants read 4 neighbouring words, in different banks), each with the quadrant's
			// signs applied: (mx*dx, my*dy), exact sign flips of the reference's mx*quadrantDirs[i].x
			// (cputrack/cpu_tracker.cpp:658-659)
			for (int j = tid; j < 4 * P.qi_A; j += kBlock) {
				const int a = j >> 2, q = j & 3;
				const float2 d = P.qi_dirs[a];
				const float sx = (q == 0 || q == 3) ? 1.0f : -1.0f;
				const float sy = (q < 2) ? 1.0f : -1.0f;
				qdir[j] = make_float2(sx * d.x, sy * d.y);
			}
		}
		__syncthreads();
		for (int k = 0; k < P.qi_iters; k++) {
			boundary = (cx + maxr >= (float)W) || (cx - maxr < 0.0f) || (cy + maxr >= (float)H) || (cy - maxr < 0.0f);
			const int steps = P.qi_iter_steps[k], stride = P.qi_iter_stride[k];
			const bool ok = !isnan(cx) && !isnan(cy);
			// uniform loop over (quadrant, radius) slots: the angle index is the same in every lane
			// (with the offset table, a half-wavefront's offsets are 256 contiguous bytes of it).
			// (quadrant, radius) slots in segments of 32 consecutive radii of one quadrant: each
			// half-wavefront samples along one ray (the LDS serves a half-wavefront at a time, and
			// neighbouring radii read neighbouring pixels), and a wavefront's two segments share
			// their radius band, so wavefronts of inner radii, whose circles lie inside the ROI,
			// take the unmasked loop
			if constexpr (kQiPrio) __builtin_amdgcn_s_setprio(0);
			for (int base = 0; base < 4 * nr; base += kBlock) {
				const int p = base + tid;
				int q, ri;
				if ((nr & 31) == 0) {
					const int sg = p >> 5;
					q = sg & 3;
					ri = ((sg >> 2) << 5) + (p & 31);
				} else {
					q = p / nr;
					ri = p - q * nr;
				}
				const float r = minr + rstep * (float)ri;
				const bool valid = p < 4 * nr;
				const bool clear = !valid || (ok && circle_inside(cx, cy, r, Wf, Hf));
				const bool fast = __ballot(!clear) == 0;  // uniform over the wavefront
				if (valid) {
					float sum = 0.0f;
					int cnt = 0;
					if constexpr (kQiTable) {
						// this lane's column of the offset table (angle j at element j * nr); the uniform
						// part of each address advances in scalar registers
						const int jstep8 = stride * nr * 8;
						const unsigned lane_off = (unsigned)(((q * A) * nr + ri) * 8);
						// one vector offset per angle of a group: the group's start is the only scalar
						const unsigned voff[4] = { lane_off, lane_off + jstep8, lane_off + 2 * jstep8,
						                           lane_off + 3 * jstep8 };
						int a = 0;
						const int n4 = steps & ~3;
						// groups of 4 angles; the next group's offsets load while the current group is
						// sampled (two register sets in turn, the scheduling barriers keep each group's
						// loads ahead of the previous group's samples; the loop's last load repeats a
						// group instead of reading past the table)
						float2 o0[4], o1[4];
						auto load4 = [&](float2 (&o)[4], int a0) {
#pragma unroll
							for (int u = 0; u < 4; u++) o[u] = qi_offset(qtab, voff[u], a0 * jstep8);
						};
						auto groups = [&](auto&& sample4) {
							if (n4 == 0) return;
							load4(o0, 0);
							int g = 0;
							for (; g + 8 <= n4; g += 8) {
								load4(o1, g + 4);
								__builtin_amdgcn_sched_barrier(0);
								sample4(o0);
								load4(o0, g + 8 < n4 ? g + 8 : g + 4);
								__builtin_amdgcn_sched_barrier(0);
								sample4(o1);
							}
							if (g < n4) sample4(o0);  // an odd number of groups: the last one
						};
						auto xy4 = [&](const float2 (&o)[4], float (&x)[4], float (&y)[4]) {
#pragma unroll
							for (int u = 0; u < 4; u++) {
								x[u] = cx + o[u].x;
								y[u] = cy + o[u].y;
							}
						};
						if (fast) {
							groups([&](const float2 (&o)[4]) {
								float x[4], y[4];
								xy4(o, x, y);
								sample_n_inside<SPix, float, 4>(img, W, x, y, sum);
							});
							for (a = n4; a < steps; a++) {
								const float2 d = qi_offset(qtab, lane_off, a * jstep8);
								const float x[1] = { cx + d.x }, y[1] = { cy + d.y };
								sample_n_inside<SPix, float, 1>(img, W, x, y, sum);
							}
							cnt = steps;
						} else {
							groups([&](const float2 (&o)[4]) {
								float x[4], y[4];
								xy4(o, x, y);
								sample_n_acc<SPix, float, 4>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
							});
							for (a = n4; a < steps; a++) {
								const float2 d = qi_offset(qtab, lane_off, a * jstep8);
								const float x[1] = { cx + d.x }, y[1] = { cy + d.y };
								sample_n_acc<SPix, float, 1>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
							}
						}
					} else {
						const float2* qd = qdir + q;
						const int st4 = 4 * stride;
						int a = 0;
						if (fast) {
							for (; a + 4 <= steps; a += 4) {
								float x[4], y[4];
#pragma unroll
								for (int u = 0; u < 4; u++) {
									const float2 d = qd[st4 * (a + u)];
									x[u] = cx + d.x * r;
									y[u] = cy + d.y * r;
								}
								sample_n_inside<SPix, float, 4>(img, W, x, y, sum);
							}
							for (; a < steps; a++) {
								const float2 d = qd[st4 * a];
								const float x[1] = { cx + d.x * r }, y[1] = { cy + d.y * r };
								sample_n_inside<SPix, float, 1>(img, W, x, y, sum);
							}
							cnt = steps;
						} else {
							for (; a + 4 <= steps; a += 4) {
								float x[4], y[4];
#pragma unroll
								for (int u = 0; u < 4; u++) {
									const float2 d = qd[st4 * (a + u)];
									x[u] = cx + d.x * r;
									y[u] = cy + d.y * r;
								}
								sample_n_acc<SPix, float, 4>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
							}
							for (; a < steps; a++) {
								const float2 d = qd[st4 * a];
								const float x[1] = { cx + d.x * r }, y[1] = { cy + d.y * r };
								sample_n_acc<SPix, float, 1>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, sum, cnt);
							}
						}
					}
					float val = cnt >= 4 ? sum / (float)cnt : mean;
					if (window) val *= window[ri];
					qprof[q * nr + ri] = val;
				}
			}
			if constexpr (kQiPrio) __builtin_amdgcn_s_setprio(1);
			__syncthreads();
			// wavefront w: the leaves of transform w (X, reversed X, Y, reversed Y) into buffer
			// w, then its forward transform
			{
				float2* F = fft + wid * N;
				const int axis = wid >> 1;
				for (int j = lane; j < N; j += 64) {
					int s = perm[j];
					if (wid & 1) s = N - 1 - s;
					F[j] = make_float2(qi_input(qprof, nr, axis, s), 0.0f);
				}
				wave_sync();
				fft_pow2_wave(F, P.qi_fft, tw, 1.0f, lane);
			}
			__syncthreads();
			// wavefront 0 takes X (buffers 0, 1), wavefront 1 Y (buffers 2, 3): F(p) * conj(F(rev p))
			// read into registers, written in leaf order into the first buffer, inverse
			// transform there, then the peak
			if (wid < 2) {
				float2* A = fft + (2 * wid) * N;
				const float2* B = A + N;
				constexpr int KP = kMaxFFTLen / 64;
				float2 prod[KP];
#pragma unroll
				for (int i = 0; i < KP; i++) {
					const int t = lane + 64 * i;
					if (t < N) {
						const int s = perm[t];
						const float2 b = B[s];
						prod[i] = cmul(A[s], make_float2(b.x, -b.y));
					}
				}
				wave_sync();
#pragma unroll
				for (int i = 0; i < KP; i++) {
					const int t = lane + 64 * i;
					if (t < N) A[t] = prod[i];
				}
				wave_sync();
				fft_pow2_wave(A, P.qi_fft, tw + N, -1.0f, lane);
				const float2* G = A;
				float bv = 0.0f;
				int bi = 0x7fffffff;
				for (int x = lane; x < N; x += 64) {
					const float v = G[(x + nr) & (N - 1)].x;
					if (bi == 0x7fffffff || v > bv) { bv = v; bi = x; }
				}
				if (bi == 0x7fffffff) bv = -INFINITY;
				wave_argmax(bv, bi);
				if (lane == 0) {
					const float maxPos = peak_fit<float>(bi, N, kQIW, [&](int x) { return G[(x + nr) & (N - 1)].x; });
					const float offs = (maxPos - (float)nr) * (3.14159265359f / 4);
					if (wid == 0) st[0] = cx + offs * rstep;
					else st[1] = cy + offs * rstep;
				}
			}
			__syncthreads();
			cx = st[0];
			cy = st[1];
			__syncthreads();
		}
		if constexpr (kQiPrio && !kChainPrio) __builtin_amdgcn_s_setprio(0);
	}

	// ---- 2D Gaussian MLE when QI is off (cputrack/cpu_tracker.cpp:514-589, selected at
	// cputrack/QueuedCPUTracker.cpp:309-313): pixel-integrated Gaussian, independent Newton
	// steps for x, y, I0 and bg, float model terms summed in double over the ROI.
	if (mode_gauss(P)) {
		const float sigma = P.gauss_sigma;
		float px = comx, py = comy;
		float I0 = mean * 0.5f * (float)W * (float)H;
		float bg = mean * 0.5f;
		const float s2 = 1.0f / (sqrtf(2.0f) * sigma);
		const float s2pi = 1.0f / (sqrtf(2 * 3.14159265359f) * sigma);
		const float s2pi3 = 1.0f / (sqrtf(2 * 3.14159265359f) * sigma * sigma * sigma);
		// The terms that depend on x alone (or y alone) are computed once per column (row)
		// and iteration: D = 0.5 erf(e0) - 0.5 erf(e1), C1 = (I0 s2pi)(exp(-e1^2) - exp(-e0^2)),
		// C2 = (I0 s2pi3)((c - p - 0.5) exp(-e1^2) - (c - p + 0.5) exp(-e0^2)), and I0 D for
		// columns; each pixel then forms the reference's left-to-right products from them
		// (I0 DX DY = (I0 DX) DY, I0 s2pi (eX1 - eX0) DY = C1x DY, ...), so its terms round as
		// the per-pixel formulas do.  Pixels are spread over all threads.
		float* gx = (float*)(lds + L.gcol);  // [4][W]: DX, C1x, C2x, I0 DX
		float* gy = gx + 4 * W;              // [3][H]: DY, C1y, C2y
		const int y1 = tid / W, x1 = tid - y1 * W, dyp = kBlock / W, dxp = kBlock - dyp * W;
		for (int it = 0; it < P.gauss_iters; it++) {
			const float Is = I0 * s2pi, Is3 = I0 * s2pi3;
			for (int c = tid; c < W + H; c += kBlock) {
				const bool row = c >= W;
				const int i = row ? c - W : c;
				const float q = row ? py : px;
				const float e0 = ((float)i - q + 0.5f) * s2;
				const float e1 = ((float)i - q - 0.5f) * s2;
				const float D = 0.5f * qm_erff(e0) - 0.5f * qm_erff(e1);
				const float x1e = qm_expf(-e1 * e1), x0e = qm_expf(-e0 * e0);
				const float C1 = Is * (x1e - x0e);
				const float C2 = Is3 * (((float)i - q - 0.5f) * x1e - ((float)i - q + 0.5f) * x0e);
				if (row) {
					gy[i] = D;
					gy[H + i] = C1;
					gy[2 * H + i] = C2;
				} else {
					gx[i] = D;
					gx[W + i] = C1;
					gx[2 * W + i] = C2;
					gx[3 * W + i] = I0 * D;
				}
			}
			__syncthreads();
			double acc[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
			int x = x1, y = y1;
			for (int p = tid; p < npx; p += kBlock, x += dxp, y += dyp) {
				if (x >= W) { x -= W; y++; }
				const float DX = gx[x], DY = gy[y];
				const float mu = bg + gx[3 * W + x] * DY;
				const float dmx = gx[W + x] * DY;
				const float dmy = gy[H + y] * DX;
				const float dmI = DX * DY;
				const float smp = (float)img[p];
				const float f = smp / mu - 1;
				acc[0] += (double)(dmx * f);
				acc[1] += (double)(dmy * f);
				acc[2] += (double)(dmI * f);
				acc[3] += (double)f;
				const float d2x = gx[2 * W + x] * DY;
				const float d2y = gy[2 * H + y] * DX;
				acc[4] += (double)(d2x * f - dmx * dmx * smp / (mu * mu));
				acc[5] += (double)(d2y * f - dmy * dmy * smp / (mu * mu));
				acc[6] += (double)(-dmI * dmI * smp / (mu * mu));
				acc[7] += (double)(-smp / (mu * mu));
			}
			block_sum<8>(acc, red, lane, wid);
			px = (float)((double)px - red[32] / red[36]);
			py = (float)((double)py - red[33] / red[37]);
			I0 = (float)((double)I0 - red[34] / red[38]);
			bg = (float)((double)bg - red[35] / red[39]);
			__syncthreads();
		}
		cx = px;
		cy = py;
	}

	// ---- BUILDLUT_IMAGELUT (cputrack/QueuedCPUTracker.cpp:606-624): the ROI sampled 1 px apart
	// on an ilut_w x ilut_h grid centred on the build position, minus the mean, added into the
	// bead's plane; before the profile, while the ROI is in LDS unchanged
	if (build && P.image_lut) {
		const int iw = P.ilut_w, ih = P.ilut_h;
		float* dst = P.image_lut + ((size_t)roi * P.ilut_planes + P.build_plane) * ih * iw;
		const float x0 = cx - (float)(iw / 2) * 1.0f, y0 = cy - (float)(ih / 2) * 1.0f;
		for (int i = tid; i < iw * ih; i += kBlock) {
			const int y = i / iw, x = i - y * iw;
			dst[i] += interpolate_ref(img, W, H, x0 + (float)x * 1.0f, y0 + (float)y * 1.0f) - mean;
		}
	}

	// ---- radial profile (cputrack/utils.cpp:275-320) and normalisation (:230-261), LUT
	// compare; with LT_ZLUTAlign 5 passes, each profiling at the current centre,
	// quadrant-aligning it from the second pass on and scoring the pass's profile, then
	// the adjusted-weight compare (cputrack/QueuedCPUTracker.cpp:319-344)
	float z = 0.0f;
	const bool lz = (P.mode & kModeLocalizeZ) || build;
	if (lz && P.z_nb > 0) {
		const int nb = P.z_nb, nch = z_chunks(P);
		// LT_FourierLUT / BUILDLUT_FOURIER: the profile of the ROI's power spectrum about (W/2, H/2)
		// over radii 1 .. 0.3 W, fallback 0, in logarithm (cputrack/cpu_tracker.cpp:954-963)
		const bool fourier = EXT && mode_fourier(P);
		const float pminr = fourier ? 1.0f : P.z_minr;
		const float pmaxr = fourier ? (float)W * 0.3f : P.z_maxr;
		const float rstep = (pmaxr - pminr) / (float)nb;
		const bool has_lut = !build && P.zluts && bead >= 0 && bead < P.zcount && P.zplanes > 0;
		const float* L0 = has_lut ? P.zluts + (size_t)bead * P.zplanes * nb : nullptr;
		const int passes = (EXT && mode_align(P) && has_lut) ? 5 : 1;
		// The LUT scores read the bead's LUT through LDS tiles in the image's place (below);
		// its first NA * kBlock 16-byte chunks load into registers here, so that they arrive
		// while the radial profile is sampled.  Only in the instances for ROIs of 96 px and
		// more, whose LDS holds a CU to three ROIs or fewer: there the 19 more VGPRs cost no
		// wavefronts (smaller ROIs fit five per CU, and five wavefronts per SIMD need <= 96
		// VGPRs; the variant with the optional modes has registers to spare neither).
		const int pitchL = nb + 1;
		const int tmax = min(kBlock, (W * H * (int)sizeof(SPix)) / (pitchL * 4));
		const bool lutfull = !EXT && L.lutfull;  // all planes in one tile (lds_layout)
		const bool tiled = has_lut && !(EXT && mode_align(P)) && (nb & 3) == 0 && ((uintptr_t)L0 & 15) == 0 &&
		                   (tmax >= 32 || lutfull);
		constexpr int NA = (!EXT && CW >= 96) ? kLutAhead : 0;
		const int qpr = nb >> 2;  // 16-byte chunks per LUT row
		const float4* L4 = (const float4*)L0;
		float4 lahead[NA > 0 ? NA : 1];
		if (NA > 0 && tiled) {
			const int nqall = P.zplanes * qpr;
#pragma unroll
			for (int u = 0; u < NA; u++) {
				const int c = tid + u * kBlock;
				if (c < nqall) lahead[u] = L4[c];
			}
		}
		for (int pass = 0; pass < passes; pass++) {
			if constexpr (EXT && std::is_same<SPix, float>::value) {
				if (fourier) power_spectrum_phase(P, L, lds, img, tid, lane, wid);
			}
			const float pcx = fourier ? (float)(W / 2) : cx, pcy = fourier ? (float)(H / 2) : cy;
			const float pfall = fourier ? 0.0f : mean;
			const bool ok = !isnan(pcx) && !isnan(pcy);
			for (int j = tid; j < P.z_A; j += kBlock) zdir[j] = P.z_dirs[j];
			__syncthreads();
			// slots bin-major (lane = nch chunks x bins), so that wavefronts of inner bins take
			// the unmasked loop
			if constexpr (kChainPrio) __builtin_amdgcn_s_setprio(0);
			for (int base = 0; base < nch * nb; base += kBlock) {
				const int p = base + tid;
				const int bin = p / nch, ch = p - bin * nch;
				const float r = pminr + rstep * (float)bin;
				const bool valid = p < nch * nb;
				const bool clear = !valid || (ok && circle_inside(pcx, pcy, r, Wf, Hf));
				const bool fast = __ballot(!clear) == 0;  // uniform over the wavefront
				if (valid) {
					const int a0 = (ch * P.z_A) / nch, a1 = ((ch + 1) * P.z_A) / nch;
					double s = 0.0;
					int c = 0;
					int a = a0;
					if (fast) {
						for (; a + 4 <= a1; a += 4) {
							float x[4], y[4];
#pragma unroll
							for (int u = 0; u < 4; u++) {
								const float2 d = zdir[a + u];
								x[u] = pcx + d.x * r;
								y[u] = pcy + d.y * r;
							}
							sample_n_inside<SPix, double, 4>(img, W, x, y, s);
						}
						for (; a < a1; a++) {
							const float2 d = zdir[a];
							const float x[1] = { pcx + d.x * r }, y[1] = { pcy + d.y * r };
							sample_n_inside<SPix, double, 1>(img, W, x, y, s);
						}
						c = a1 - a0;
					} else {
						for (; a + 4 <= a1; a += 4) {
							float x[4], y[4];
#pragma unroll
							for (int u = 0; u < 4; u++) {
								const float2 d = zdir[a + u];
								x[u] = pcx + d.x * r;
								y[u] = pcy + d.y * r;
							}
							sample_n_acc<SPix, double, 4>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, s, c);
						}
						for (; a < a1; a++) {
							const float2 d = zdir[a];
							const float x[1] = { pcx + d.x * r }, y[1] = { pcy + d.y * r };
							sample_n_acc<SPix, double, 1>(img, W, Wf, Hf, Wm2, Hm2, ok, x, y, s, c);
						}
					}
					zpart[ch * nb + bin] = s;
					zcnt[ch * nb + bin] = c;
				}
			}
			if constexpr (kChainPrio) __builtin_amdgcn_s_setprio(1);
			__syncthreads();
			for (int bin = tid; bin < nb; bin += kBlock) {
				double s = zpart[bin];
				int c = zcnt[bin];
				for (int ch = 1; ch < nch; ch++) {
					s += zpart[ch * nb + bin];
					c += zcnt[ch * nb + bin];
				}
				const float v = c > 4 ? (float)(s / (double)c) : pfall;
				zprof[bin] = fourier ? logf(v) : v;
			}
			__syncthreads();
			if (!fourier) {
				if ((P.mode & kModeNormalizeProfile) && wid == 0) normalize_profile_wave(zprof, nb, lane);
				__syncthreads();
				boundary = (cx + P.z_maxr >= (float)W) || (cx - P.z_maxr < 0.0f) || (cy + P.z_maxr >= (float)H) ||
				           (cy - P.z_maxr < 0.0f);
			}
			if (P.prof_out && !build)
				for (int bin = tid; bin < nb; bin += kBlock) P.prof_out[(size_t)roi * nb + bin] = zprof[bin];
			if constexpr (EXT) {
				// the centre moves; the profile of this pass stays the one scored, and
				// QuadrantAlign leaves the boundary flag cleared (cputrack/cpu_tracker.cpp:484)
				if (pass > 0) {
					quadrant_align_phase<SPix>(P, L, lds, img, mean, L0, z, cx, cy, tid, lane, wid);
					boundary = 0;
				}
			}

			if (build) {
				float* dst = P.build_lut + ((size_t)roi * P.zplanes + P.build_plane) * nb;
				for (int bin = tid; bin < nb; bin += kBlock) dst[bin] += zprof[bin];
			} else if (has_lut) {
				// ---- LUT scores, best plane, FP64 peak fit (cputrack/cpu_tracker.cpp:791-868)
				// One thread per plane sums its bins in order.  Unless QuadrantAlign samples the
				// ROI again, the LUT goes through LDS tiles in the image's place (one tile over
				// the image and the phase scratch when it fits: lutfull): coalesced 16-byte loads
				// (the first ones loaded ahead, above), rows padded to nb + 1 floats so that the
				// threads' reads of a bin fall in different banks.
				if (tiled) {
					float* tile = (float*)(lds + L.img);
					const int ntile = lutfull ? 1 : (P.zplanes + tmax - 1) / tmax;
					const int tp = (P.zplanes + ntile - 1) / ntile;  // planes per tile, balanced
					auto put = [&](int cr, const float4& v) {   // chunk cr of the tile
						const int kk = cr / qpr;
						float* d = tile + kk * pitchL + ((cr - kk * qpr) << 2);
						d[0] = v.x;
						d[1] = v.y;
						d[2] = v.z;
						d[3] = v.w;
					};
					for (int k0 = 0; k0 < P.zplanes; k0 += tp) {
						const int np = min(tp, P.zplanes - k0);
						const int cb = k0 * qpr, ce = cb + np * qpr;  // the tile's chunks
#pragma unroll
						for (int u = 0; u < NA; u++) {
							const int c = tid + u * kBlock;
							if (c >= cb && c < ce) put(c - cb, lahead[u]);
						}
						// chunks past those loaded ahead, 4 loads in flight
						for (int c0 = max(cb, NA * kBlock) + tid; c0 < ce; c0 += 4 * kBlock) {
							float4 v[4];
#pragma unroll
							for (int u = 0; u < 4; u++) {
								const int c = c0 + u * kBlock;
								v[u] = L4[c < ce ? c : ce - 1];
							}
#pragma unroll
							for (int u = 0; u < 4; u++) {
								const int c = c0 + u * kBlock;
								if (c < ce) put(c - cb, v[u]);
							}
						}
						__syncthreads();
						double s = 0.0;
						if (tid < np) {
							// bins read 4 at a time ahead of the serial sum (nb is a multiple of 4)
							const float* row = tile + tid * pitchL;
							for (int r0 = 0; r0 < nb; r0 += 4) {
								float lv[4], pv[4];
#pragma unroll
								for (int j = 0; j < 4; j++) {
									lv[j] = row[r0 + j];
									pv[j] = zprof[r0 + j];
								}
#pragma unroll
								for (int j = 0; j < 4; j++) {
									const double d = (double)(pv[j] - lv[j]);
									s = zcmp_step(s, d, P.zcmp, r0 + j);
								}
							}
						}
						if (lutfull) __syncthreads();  // the LUT is read: the scores go over it
						if (tid < np) {
							scores[k0 + tid] = s;
							if (P.cmp_out) P.cmp_out[(size_t)bead * P.zplanes + k0 + tid] = (float)s;
						}
						__syncthreads();
					}
				} else {
					for (int k = tid; k < P.zplanes; k += kBlock) {
						const float* Lk = L0 + (size_t)k * nb;
						double s = 0.0;
						for (int r = 0; r < nb; r++) s = zcmp_step(s, (double)(zprof[r] - Lk[r]), P.zcmp, r);
						scores[k] = s;
						if (P.cmp_out) P.cmp_out[(size_t)bead * P.zplanes + k] = (float)s;
					}
					__syncthreads();
				}
				int best = 0;
				z = scores_peak(scores, P.zplanes, red, st, tid, lane, wid, best);
				if (tid == 0 && P.best_out) P.best_out[roi] = best;
				if constexpr (EXT) {
					// every thread needs z for the next alignment or the adjusted weights
					if (tid == 0) st[2] = z;
					__syncthreads();
					z = st[2];
				}
			}
		}
		if constexpr (EXT) {
			if (mode_zweighted(P) && has_lut) {
				// ---- LUTProfileCompareAdjustedWeights (cputrack/cpu_tracker.cpp:871-901): radius
				// weights |L[zi+1] - L[zi]| * r_i around the plane of z, double differences.  A z
				// below 0 (or NaN) uses plane 0; the reference would index before the LUT there.
				double* zw = (double*)(lds + L.zw);
				int zi = trunc_sat(z);
				if (zi > P.zplanes - 2) zi = P.zplanes - 2;
				if (zi < 0) zi = 0;
				const int zi1 = zi + 1 < P.zplanes ? zi + 1 : zi;
				const float* z0 = L0 + (size_t)zi * nb;
				const float* z1 = L0 + (size_t)zi1 * nb;
				__syncthreads();  // the last pass's scores are read; zw follows them in LDS
				for (int r = tid; r < nb; r += kBlock)
					zw[r] = fabs((double)z1[r] - (double)z0[r]) *
					        (double)(P.z_minr + (P.z_maxr - P.z_minr) * (float)r / (float)nb);
				__syncthreads();
				for (int k = tid; k < P.zplanes; k += kBlock) {
					const float* Lk = L0 + (size_t)k * nb;
					double s = 0.0;
					for (int r = 0; r < nb; r++) {
						double d = (double)zprof[r] - (double)Lk[r];
						d = -d * d;
						d *= zw[r];
						s += d;
					}
					scores[k] = s;
				}
				__syncthreads();
				int best = 0;
				z = scores_peak(scores, P.zplanes, red, st, tid, lane, wid, best);
			}
		}
	}

	if (tid == 0 && !build) {
		if (P.best_out && !(lz && P.zluts && bead >= 0 && bead < P.zcount && P.zplanes > 0)) P.best_out[roi] = -1;
		DevResult R;
		R.job = job;
		R.x = cx;
		R.y = cy;
		R.z = z;
		R.comx = comx;
		R.comy = comy;
		R.error = boundary ? 1u : 0u;
		R.imageMean = mean;
		P.results[roi] = R;
	}
}

// ---- 2D Gaussian MLE alone on small ROIs: one wavefront per ROI
//
// The same computation as localize_kernel's staging, mean/COM and Gaussian MLE phases
// (cputrack/QueuedCPUTracker.cpp:258-313, cputrack/cpu_tracker.cpp:514-589, 681-729)
// with every sum kept inside the wavefront: each lane sums its pixels (p = lane + 64 k),
// then a butterfly over the wavefront.  No workgroup barrier, and a fourth of the
// shuffles per ROI of the four-wavefront form.  The double sums group the pixels
// differently from the workgroup form (exact for integer-valued pixels).
constexpr int kGaussWaveMaxPx = 1024;

__host__ __device__ inline int gauss_wave_lds_floats(const LocalizeParams& P)
{
	return (P.width * P.height + 4 * P.width + 3 * P.height + 3) & ~3;
}
inline bool gauss_wave_eligible(const LocalizeParams& P)
{
	return mode_gauss(P) && !(P.mode & kModeLocalizeZ) && P.width * P.height <= kGaussWaveMaxPx && P.width >= 2 &&
	       P.height >= 2 && P.pdt >= 0 && P.pdt <= 2;
}

template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K])
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
#pragma unroll
		for (int k = 0; k < K; k++) v[k] += __shfl_xor(v[k], off);
}

__global__ void __launch_bounds__(kBlock) gauss_wave_kernel(LocalizeParams P)
{
	extern __shared__ __align__(16) unsigned char lds[];
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const int roi = blockIdx.x * (kBlock / 64) + wid;
	if (roi >= P.n) return;  // whole wavefronts; nothing below waits on other wavefronts
	const int W = P.width, H = P.height, npx = W * H;
	float* img = (float*)lds + (size_t)wid * gauss_wave_lds_floats(P);
	float* gx = img + npx;     // [4][W]: DX, C1x, C2x, I0 DX
	float* gy = gx + 4 * W;    // [3][H]: DY, C1y, C2y
	const DevJob job = P.jobs ? P.jobs[roi] : DevJob{};
	const int bead = job.zlutIndex;

	// ROI -> LDS as float, ClearFirstFourPixels, offset/gain with unit factors
	{
		const uint8_t* base = P.src + (P.roi_offset ? P.roi_offset[roi] : (int64_t)roi * H * P.pitch);
		const bool calib = (P.calib_offset || P.calib_gain) && bead >= 0 && bead < P.zcount;
		const float* off = calib && P.calib_offset ? P.calib_offset + (size_t)bead * npx : nullptr;
		const float* gain = calib && P.calib_gain ? P.calib_gain + (size_t)bead * npx : nullptr;
		const bool clear4 = (P.mode & kModeClearFirstFour) != 0;
		const int y1 = lane / W, x1 = lane - y1 * W, dyp = 64 / W, dxp = 64 - dyp * W;
		int x = x1, y = y1;
		for (int i = lane; i < npx; i += 64, x += dxp, y += dyp) {
			if (x >= W) { x -= W; y++; }
			const uint8_t* row = base + (int64_t)y * P.pitch;
			float v;
			if (P.pdt == 1) v = (float)((const uint16_t*)row)[x];
			else if (P.pdt == 0) v = (float)row[x];
			else v = ((const float*)row)[x];
			if (clear4 && i < 4) v = 0.0f;
			if (off && !gain) v = v + off[i] * 1.0f;
			else if (gain && !off) v = v * gain[i] * 1.0f;
			else if (gain && off) v = (v + off[i] * 1.0f) * gain[i] * 1.0f;
			img[i] = v;
		}
	}
	wave_sync();

	// mean, stdev and background-corrected COM
	double m1[2] = { 0.0, 0.0 };
	for (int i = lane; i < npx; i += 64) {
		const float v = img[i];
		m1[0] += (double)v;
		m1[1] += (double)(float)(v * v);
	}
	wave_sum<2>(m1);
	const float invN = 1.0f / (float)npx;
	const float mean = (float)(m1[0] * (double)invN);
	const double var = m1[1] * (double)invN - (double)(mean * mean);
	const float stdev = sqrtf((float)var);
	const float bgs = P.com_bgcorrection * stdev;
	const int y1 = lane / W, x1 = lane - y1 * W, dyp = 64 / W, dxp = 64 - dyp * W;
	double m2[3] = { 0.0, 0.0, 0.0 };
	{
		int x = x1, y = y1;
		for (int i = lane; i < npx; i += 64, x += dxp, y += dyp) {
			if (x >= W) { x -= W; y++; }
			const float d = fabsf(img[i] - mean) - bgs;
			const float v = (0.0f < d) ? d : 0.0f;
			m2[0] += (double)v;
			m2[1] += (double)x * (double)v;
			m2[2] += (double)y * (double)v;
		}
	}
	wave_sum<3>(m2);
	float comx = (float)(m2[1] / m2[0]);
	float comy = (float)(m2[2] / m2[0]);
	if (isnan(comx) || isnan(comy)) {
		comx = (float)(W / 2);
		comy = (float)(H / 2);
	}

	// Gaussian MLE, as in localize_kernel: per-column and per-row terms, then per pixel
	const float sigma = P.gauss_sigma;
	float px = comx, py = comy;
	float I0 = mean * 0.5f * (float)W * (float)H;
	float bg = mean * 0.5f;
	const float s2 = 1.0f / (sqrtf(2.0f) * sigma);
	const float s2pi = 1.0f / (sqrtf(2 * 3.14159265359f) * sigma);
	const float s2pi3 = 1.0f / (sqrtf(2 * 3.14159265359f) * sigma * sigma * sigma);
	for (int it = 0; it < P.gauss_iters; it++) {
		const float Is = I0 * s2pi, Is3 = I0 * s2pi3;
		for (int c = lane; c < W + H; c += 64) {
			const bool row = c >= W;
			const int i = row ? c - W : c;
			const float q = row ? py : px;
			const float e0 = ((float)i - q + 0.5f) * s2;
			const float e1 = ((float)i - q - 0.5f) * s2;
			const float D = 0.5f * qm_erff(e0) - 0.5f * qm_erff(e1);
			const float x1e = qm_expf(-e1 * e1), x0e = qm_expf(-e0 * e0);
			const float C1 = Is * (x1e - x0e);
			const float C2 = Is3 * (((float)i - q - 0.5f) * x1e - ((float)i - q + 0.5f) * x0e);
			if (row) {
				gy[i] = D;
				gy[H + i] = C1;
				gy[2 * H + i] = C2;
			} else {
				gx[i] = D;
				gx[W + i] = C1;
				gx[2 * W + i] = C2;
				gx[3 * W + i] = I0 * D;
			}
		}
		wave_sync();
		double acc[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
		int x = x1, y = y1;
		for (int p = lane; p < npx; p += 64, x += dxp, y += dyp) {
			if (x >= W) { x -= W; y++; }
			const float DX = gx[x], DY = gy[y];
			const float mu = bg + gx[3 * W + x] * DY;
			const float dmx = gx[W + x] * DY;
			const float dmy = gy[H + y] * DX;
			const float dmI = DX * DY;
			const float smp = img[p];
			const float f = smp / mu - 1;
			acc[0] += (double)(dmx * f);
			acc[1] += (double)(dmy * f);
			acc[2] += (double)(dmI * f);
			acc[3] += (double)f;
			const float d2x = gx[2 * W + x] * DY;
			const float d2y = gy[2 * H + y] * DX;
			acc[4] += (double)(d2x * f - dmx * dmx * smp / (mu * mu));
			acc[5] += (double)(d2y * f - dmy * dmy * smp / (mu * mu));
			acc[6] += (double)(-dmI * dmI * smp / (mu * mu));
			acc[7] += (double)(-smp / (mu * mu));
		}
		wave_sum<8>(acc);
		px = (float)((double)px - acc[0] / acc[4]);
		py = (float)((double)py - acc[1] / acc[5]);
		I0 = (float)((double)I0 - acc[2] / acc[6]);
		bg = (float)((double)bg - acc[3] / acc[7]);
		wave_sync();  // every lane has read this iteration's terms before they are rewritten
	}

	if (lane == 0) {
		if (P.best_out) P.best_out[roi] = -1;
		DevResult R;
		R.job = job;
		R.x = px;
		R.y = py;
		R.z = 0.0f;
		R.comx = comx;
		R.comy = comy;
		R.error = 0u;
		R.imageMean = mean;
		P.results[roi] = R;
	}
}

// Launches above the default 64 KB of dynamic LDS raise the kernel's limit first (on the
// current device, every time: the setting is per device).
template <typename SPix, bool EXT, int CW, int CH>
static hipError_t launch_sized(const LocalizeParams& p, size_t lds, hipStream_t stream)
{
	if (lds > 64 * 1024) {
		const hipError_t e = hipFuncSetAttribute((const void*)localize_kernel<SPix, EXT, CW, CH>,
		                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
		if (e != hipSuccess) return e;
	}
	hipLaunchKernelGGL((localize_kernel<SPix, EXT, CW, CH>), dim3(p.n), dim3(kBlock), lds, stream, p);
	return hipGetLastError();
}

// Square ROIs of common sizes, staged as float, run an instance compiled with the size as a
// constant: a bilinear sample's second-row offset (W and W + 1 floats) then folds into the LDS
// instruction's immediate offsets instead of costing an address addition per sample, and the loops
// over rows and columns have constant bounds (realtime100, 100x100: 0.747 -> 0.728 ms per launch).
template <typename SPix, bool EXT>
static hipError_t launch_variant(const LocalizeParams& p, size_t lds, hipStream_t stream)
{
	if constexpr (sizeof(SPix) == sizeof(float) && !EXT) {
		if (p.width == p.height) {
			switch (p.width) {
			case 32: return launch_sized<SPix, EXT, 32, 32>(p, lds, stream);
			case 48: return launch_sized<SPix, EXT, 48, 48>(p, lds, stream);
			case 64: return launch_sized<SPix, EXT, 64, 64>(p, lds, stream);
			case 80: return launch_sized<SPix, EXT, 80, 80>(p, lds, stream);
			case 96: return launch_sized<SPix, EXT, 96, 96>(p, lds, stream);
			case 100: return launch_sized<SPix, EXT, 100, 100>(p, lds, stream);
			case 128: return launch_sized<SPix, EXT, 128, 128>(p, lds, stream);
			default: break;
			}
		}
	}
	return launch_sized<SPix, EXT, 0, 0>(p, lds, stream);
}

hipError_t launch_localize(const LocalizeParams& p, hipStream_t stream)
{
	if (p.n <= 0) return hipSuccess;
	const size_t lds = localize_lds_bytes(p);
	const bool ext = mode_ext(p);
	if (ext && mode_xcor(p) && (p.xc_fft.n < 2 || p.xc_fft.n > kMaxFFTLen)) return hipErrorInvalidValue;
	if (ext && mode_align(p) && (p.qa_fft.n != 2 * p.z_nb || p.qa_fft.n > kMaxFFTLen)) return hipErrorInvalidValue;
	if (ext && mode_fourier(p) && (p.lds_pix != 2 || p.fx_fft.n != p.width || p.fy_fft.n != p.height))
		return hipErrorInvalidValue;
	if (!p.build && gauss_wave_eligible(p)) {
		const size_t wl = (size_t)(kBlock / 64) * gauss_wave_lds_floats(p) * sizeof(float);
		hipLaunchKernelGGL(gauss_wave_kernel, dim3((p.n + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), wl, stream, p);
		return hipGetLastError();
	}
	if (lds > 160 * 1024) return hipErrorInvalidValue;
	if (p.lds_pix == 0) return ext ? launch_variant<uint8_t, true>(p, lds, stream) : launch_variant<uint8_t, false>(p, lds, stream);
	if (p.lds_pix == 1) return ext ? launch_variant<uint16_t, true>(p, lds, stream) : launch_variant<uint16_t, false>(p, lds, stream);
	return ext ? launch_variant<float, true>(p, lds, stream) : launch_variant<float, false>(p, lds, stream);
}

// NormalizeRadialProfile on each row (FinalizeLUT, cputrack/QueuedCPUTracker.cpp:648-659).
__global__ void normalize_rows_kernel(float* data, int rows, int len)
{
	const int row = blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= rows) return;
	float* p = data + (size_t)row * len;
	double s = 0.0;
	for (int i = 0; i < len; i++) s += (double)p[i];
	const float m = (float)(s / (double)len);
	double s2 = 0.0;
	for (int i = 0; i < len; i++) {
		const float v = p[i] - m;
		p[i] = v;
		s2 += (double)(float)(v * v);
	}
	const double inv = (double)1.0f / sqrt(s2 / (double)len);
	for (int i = 0; i < len; i++) p[i] = (float)((double)p[i] * inv);
}

// The image LUT's first and second differences across planes (cputrack/QueuedCPUTracker.cpp:
// 661-696): plane z of bead b, pixel k, for 0 < z < planes - 1: dz = 0.5 (U - L) / 1 and
// dz2 = (U - 2 C + L) / (1 * 1) with L, C, U the planes z - 1, z, z + 1; the first and last planes
// take those of their neighbour.  One thread per (bead, pixel), planes in order.
__global__ void image_lut_diff_kernel(const float* lut, float* dz, float* dz2, int beads, int planes, int npx)
{
	const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= (size_t)beads * npx) return;
	const size_t b = g / npx, k = g - b * npx;
	const size_t base = b * planes * (size_t)npx + k;
	for (int z = 1; z < planes - 1; z++) {
		const float L = lut[base + (size_t)(z - 1) * npx], C = lut[base + (size_t)z * npx],
		            U = lut[base + (size_t)(z + 1) * npx];
		dz[base + (size_t)z * npx] = 0.5f * (U - L) / 1.0f;
		dz2[base + (size_t)z * npx] = (U - 2 * C + L) / (1.0f * 1.0f);
	}
	dz[base] = dz[base + npx];
	dz2[base] = dz2[base + npx];
	dz[base + (size_t)(planes - 1) * npx] = dz[base + (size_t)(planes - 2) * npx];
	dz2[base + (size_t)(planes - 1) * npx] = dz2[base + (size_t)(planes - 2) * npx];
}

hipError_t launch_image_lut_diff(const float* lut, float* dz, float* dz2, int beads, int planes, int npx,
                                 hipStream_t stream)
{
	const size_t n = (size_t)beads * npx;
	if (n == 0 || planes < 3) return hipSuccess;
	hipLaunchKernelGGL(image_lut_diff_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, lut, dz, dz2,
	                   beads, planes, npx);
	return hipGetLastError();
}

hipError_t launch_normalize_rows(float* data, int rows, int len, hipStream_t stream)
{
	if (rows <= 0 || len <= 0) return hipSuccess;
	hipLaunchKernelGGL(normalize_rows_kernel, dim3((rows + 63) / 64), dim3(64), 0, stream, data, rows, len);
	return hipGetLastError();
}

}  // namespace qtrk
