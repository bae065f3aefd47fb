// lutmatch.hpp -- grouped radial Z-LUT matching: many profiles of each bead
// against that bead's LUT, as one launch over all beads.
//
// The reference scores one profile at a time (CPUTracker::LUTProfileCompare,
// cputrack/cpu_tracker.cpp:791-868):
//     score[k] = sum_r -(w_r * (float)(prof[r] - L[k][r]))^2     (double sum)
// followed by the 7-point FP64 fit around the best plane.  Here the profiles
// of one bead are matched together:
//   * exact path: every plane scored in exactly that arithmetic, with each
//     LUT row read once per 8 profiles;
//   * matrix-core path: -score = |p|^2_w - 2 p.L_w + |L|^2_w, and the cross
//     term for all (profile, plane) pairs is computed on the matrix cores.
//     Each tile of 32 planes is stored as its mean row M_t plus BF16/FP16
//     residuals D = L - M_t: p.L = p.M_t (per profile and tile, from M_t and
//     p split into high and low 16-bit parts, three MFMA products: about
//     16 bits more than one part) + p.D (MFMA GEMM, FP32 accumulation).
//     Neighbouring planes of a LUT are similar, so |D| is a fraction of |L|
//     and so is the rounding error of the
//     approximate score -- what decides whether the right planes reach the
//     candidate list.  The 8 planes with the highest approximate score are
//     rescored in the reference arithmetic, the best of them is the best plane,
//     and the fit window around it is scored in the reference arithmetic too.
//     Whenever the reference's best plane is among the 8 candidates, the
//     result (best plane and z) is identical to the reference's.
#pragma once

#include "devbuf.hpp"

#include <hip/hip_runtime.h>
#include <cstdint>

namespace qtrk {

enum : int { kMatchFP32 = 0, kMatchBF16 = 1, kMatchFP16 = 2 };
// A device batch's fourth form: the profile kernel and the grouped match in kMatchFP32 (the
// grouped match itself knows only the three precisions above).
constexpr int kMatchFP32Grouped = 3;
// the grouped match's precision for a batch form
inline int match_kernel_precision(int form) { return form == kMatchFP32Grouped ? kMatchFP32 : form; }
constexpr int kMatchTopK = 8;
constexpr int kMatchMaxRes = 256;     // radial steps supported by the matrix-core path
constexpr int kMatchExactTile = 8;    // profiles per workgroup, exact path (4 for few per bead)
constexpr int kMatchCUs = 256;        // compute units of an MI355X

struct LutMatchParams {
	int beads, planes, res;
	int ptiles, ksteps;       // ceil(planes / 32), ceil(res / 16)
	int idx_bits;             // plane-index bits embedded in candidate keys
	const float* lut;         // [beads][planes][res] float32
	const uint4* frag;        // [beads][ptiles][ksteps][64 lanes] x 8 16-bit residuals (MFMA A-operand order)
	const float* norm;        // [beads][ptiles * 32]: sum_r w_r^2 L^2 (float)
	const uint4* mfrag;       // [beads][mtiles][ksteps][64 lanes] x 8 16-bit tile means (MFMA A-operand order):
	                          // rows 0-15 of mean tile m hold the high parts of M_t for t = 16m ... 16m+15,
	                          // rows 16-31 the low parts M_t - high (0 past planes and res)
	int mtiles;               // ceil(ptiles / 16)
	const float* w;           // [res] radial weights, or null
	// profiles, grouped by bead: slots start[b] .. start[b+1]-1 belong to bead b;
	// slot s reads profile row rowidx[s] (s when rowidx is null) of prof[rows][res]
	const float* prof;
	const int* start;         // [beads + 1]
	const int* rowidx;
	int chunks;               // profile tiles per bead in the grid
	// outputs, indexed by profile row
	float* z;                 // z of row i at z[i * zstride] (zstride 0 or 1: dense)
	int zstride;
	int* best;                // or null
	int* cand;                // [rows][8] candidate planes (matrix-core path), or null
};

// Splits the FP32 LUT into tile means (raw_mean: [beads][ptiles][mres], mres = res
// rounded up to 8) and residual MFMA A-operand fragments of `prec` (kMatchBF16 or
// kMatchFP16), writes the means as high/low fragments (mfrag) of the same
// precision, and computes the weighted plane norms.
hipError_t launch_lutmatch_prepare(const LutMatchParams& p, int prec, uint4* frag, uint4* mfrag, float* norm,
                                   float* raw_mean, int mres, hipStream_t stream);
// One grouped match of all beads' profiles (prec: kMatchFP32 / BF16 / FP16);
// max_per_bead bounds start[b+1] - start[b].  frame_tile: profiles per
// matrix-core workgroup, 32 or 64 (0: 32, which with half the registers and
// LDS of 64 holds more workgroups per CU and was faster on every shape measured,
// although a bead's LUT tiles are then read once per 32 profiles).
hipError_t launch_lutmatch(const LutMatchParams& p, int prec, int max_per_bead, hipStream_t stream,
                           int frame_tile = 0, bool loop_tiles = false);
// The launch's grid has ceil(max_per_bead / tile) workgroups per bead.  max_per_bead must bound
// every bead's count unless loop_tiles, where a workgroup also takes the bead's later tiles (every
// ceil(max_per_bead / tile)-th): then any max_per_bead is correct, a smaller one only slower.
// Dynamic LDS bytes of that launch.
size_t lutmatch_lds_bytes(int prec, int res, int planes, int ksteps, int max_per_bead, int frame_tile = 0,
                          int beads = 1);

// Groups rows 0..n-1 by bead (bead of row i at bead_idx[i * stride]; rows of other indices are left
// out): counts [beads] (scratch, ends as fill cursors), start [beads + 1], rowidx [n].
hipError_t launch_group_by_bead(const int* bead_idx, int stride, int n, int beads, int* counts, int* start,
                                int* rowidx, hipStream_t stream);

// Device-resident LUT of a grouped match in every precision's format: the FP32
// LUT and weights, and per precision (built on first use) the tile means,
// residual fragments and plane norms of the matrix-core path.
struct LutMatchTables {
	int beads = 0, planes = 0, res = 0, ptiles = 0, ksteps = 0, mres = 0, mtiles = 0, idx_bits = 0;
	DeviceBuffer lut, w, norm, raw_mean, frag[3], mfrag[3];
	bool have_w = false, prepared[3] = {};

	LutMatchTables(int beads, int planes, int res);
	// luts [beads][planes][res] and weights [res] (or null), host or device memory
	void set(const float* luts, const float* weights, bool device_src, hipStream_t stream);
	void prepare(int prec, hipStream_t stream);
	// parameters with the LUT fields filled (profiles, grouping and outputs left null)
	LutMatchParams params(int prec) const;
};

}  // namespace qtrk
