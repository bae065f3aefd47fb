// localize.hpp -- parameter block of the fused localization kernel, shared by
// the host runtime (hip_tracker.cpp) and the kernels (localize.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace qtrk {

// 24-byte job and 52-byte result, field for field the packed LocalizationJob /
// LocalizationResult of the C ABI (all fields are 4 bytes, so the natural
// layout has no padding and 4-byte alignment is safe on the GPU).
struct DevJob {
	uint32_t frame, timestamp;
	int32_t zlutIndex;
	float ix, iy, iz;
};
struct DevResult {
	DevJob job;
	float x, y, z;
	float comx, comy;
	uint32_t error;
	float imageMean;
};
static_assert(sizeof(DevJob) == 24, "job layout");
static_assert(sizeof(DevResult) == 52, "result layout");

constexpr int kMaxFFTStages = 8;
constexpr int kBlock = 256;  // threads per ROI workgroup (4 wavefronts)
constexpr int kLdsPerCU = 160 * 1024;  // LDS bytes of a gfx950 CU
// 16-byte LUT chunks per thread loaded into registers before the radial profile is sampled: a
// 200-plane x 64-step LUT whole (52 VGPRs; the workgroup's LDS already limits a CU to three ROIs,
// i.e. three wavefronts per SIMD, which leaves that many registers unused)
constexpr int kLutAhead = 13;

enum : int {
	kModeXCor1D = 1, kModeQI = 2, kModeGauss2D = 4, kModeZLUTAlign = 8, kModeLocalizeZ = 16,
	kModeNormalizeProfile = 64, kModeClearFirstFour = 128, kModeFourierLUT = 256, kModeLocalizeZWeighted = 512
};

// kissfft's plan for one transform length (cputrack/kissfft.h:11-47): stage radices
// ("4s, then 2s, then 3, 5, 7, ..."), remainders, twiddle strides, the input index
// read at each leaf, and the forward/inverse twiddles in GPU memory.
struct FFTPlan {
	int n, nst;
	int radix[kMaxFFTStages], m[kMaxFFTStages], fs[kMaxFFTStages];
	const int* perm;            // [n]
	const float2* tw_fwd;       // [n]
	const float2* tw_inv;       // [n]
};

// Largest transform the kernels stage: the spectrum products of two transforms are
// held in registers, at most 4 per thread.
constexpr int kMaxFFTLen = 2 * kBlock;

struct LocalizeParams {
	// ROI source: ROI i starts at src + roi_offset[i] (or src + i*height*pitch
	// when roi_offset is null); rows are `pitch` bytes apart.
	const uint8_t* src;
	const int64_t* roi_offset;
	int pitch;
	int pdt;  // 0 u8, 1 u16, 2 f32
	int lds_pix;  // pixel format staged in LDS (same codes): the source format unless calibration is on
	int n;
	const DevJob* jobs;
	DevResult* results;
	int width, height;
	int mode;
	float com_bgcorrection;
	int gauss_iters;      // LT_Gaussian2D: Newton iterations and fixed sigma
	float gauss_sigma;

	// LUT-build variant (QueuedCPUTracker::BuildLUT): COM without background
	// correction, QI without the radial window, profile accumulated into
	// build_lut[(roi*planes + build_plane)*z_nb + r], no LUT compare.
	int build;
	int build_plane;
	const float* known_pos;  // [n][2] or null
	float* build_lut;
	// BUILDLUT_IMAGELUT: image_lut[roi][ilut_planes][ilut_h][ilut_w], plane build_plane
	// accumulated; null when the flag is off
	float* image_lut;
	int ilut_planes, ilut_h, ilut_w;

	// QI
	int qi_iters, qi_nr, qi_A;
	float qi_minr, qi_maxr;
	const int* qi_iter_steps;   // [qi_iters] angular samples per quadrant
	const int* qi_iter_stride;  // [qi_iters] step through the direction table
	const float2* qi_dirs;      // [qi_A]
	// [4][qi_A][qi_nr]: sample offsets (fl(mx*dx_j * r_i), fl(my*dy_j * r_i)) of quadrant q, angle j,
	// radius r_i = minr + ((maxr - minr) / nr) * i, the float products the reference adds to the
	// centre (cputrack/cpu_tracker.cpp:658-659); the same for every ROI
	const float2* qi_offs;
	const float* qi_window;     // [qi_nr] or null

	FFTPlan qi_fft;             // length 2*qi_nr (a power of two)

	// LT_XCor1D (cputrack/cpu_tracker.cpp:190-270): profile length, width and iterations
	int xc_len, xc_width, xc_iters;
	FFTPlan xc_fft;             // length xc_len

	// LT_ZLUTAlign (cputrack/cpu_tracker.cpp:409-512): angular samples per quadrant and
	// their stride through qi_dirs, and the transform of length 2*z_nb
	int qa_steps, qa_stride;
	FFTPlan qa_fft;

	// LT_FourierLUT (cputrack/cpu_tracker.cpp:920-963): transforms of the ROI's rows (length
	// width) and columns (length height); `fourier` selects BUILDLUT_FOURIER in the build variant
	FFTPlan fx_fft, fy_fft;
	int fourier;

	// radial profile + Z-LUT
	int z_nb, z_A;
	float z_minr, z_maxr;
	const float2* z_dirs;       // [z_A]
	const float* zluts;         // [zcount][zplanes][z_nb] or null
	int zcount, zplanes;
	const float* zcmp;          // [z_nb] or null
	float* cmp_out;             // [zcount][zplanes] or null

	// per-bead calibration images [zcount][height][width]
	const float* calib_offset;
	const float* calib_gain;

	// optional diagnostics per ROI: best LUT plane (-1 without a compare) and the radial profile
	int* best_out;              // [n] or null
	float* prof_out;            // [n][z_nb] or null
};

// Dynamic LDS bytes needed by localize_kernel for these parameters.
size_t localize_lds_bytes(const LocalizeParams& p);
// Launches one workgroup per ROI on `stream`.
hipError_t launch_localize(const LocalizeParams& p, hipStream_t stream);
// Normalizes `rows` profiles of `len` values in place (FinalizeLUT).
hipError_t launch_normalize_rows(float* data, int rows, int len, hipStream_t stream);
// The image LUT's plane differences of FinalizeLUT ([beads][planes][npx] each; planes >= 3).
hipError_t launch_image_lut_diff(const float* lut, float* dz, float* dz2, int beads, int planes, int npx,
                                 hipStream_t stream);

}  // namespace qtrk
