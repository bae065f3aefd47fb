// lutmatch.hip -- grouped radial Z-LUT matching for gfx950 (exact path and
// BF16/FP16 matrix-core path; the scoring rule is stated in lutmatch.hpp).
//
// Matrix-core path, one workgroup (4 wavefronts) per bead and 32 (or 64) profiles:
//   * the bead's profiles are staged in LDS (float) and turned into MFMA
//     B operands once: B[k=r][n=profile] = w_r^2 * prof[r], rounded to BF16/FP16
//     (its rounding remainder is a second set of B operands);
//   * the LUT was split once (launch_lutmatch_prepare) into tile means and
//     residual A-operand fragments, [bead][32-plane tile][16-radius step][lane]
//     x 8 values, so each wavefront's operand load is one contiguous, fully
//     coalesced 1 KB read;
//   * the tile-mean term p.M_t of every (tile, profile) pair comes from the
//     matrix cores too: the means' high and low parts against B and its
//     remainder (hi.hi + lo.hi + hi.lo + lo.lo, FP32 accumulation);
//   * v_mfma_f32_32x32x16_{bf16,f16} accumulates p.D_w for a 32-plane x
//     32-profile tile in FP32; the epilogue adds the tile-mean term
//     2 p.M_t, subtracts |L|^2_w and folds the 16 values each lane holds into a
//     per-profile top-8 kept in registers (sorting networks on keys that carry
//     the plane index in their low bits);
//   * the top-8 lists of the two half-waves and the four wavefronts are merged,
//     the 8 candidates and the fit window are rescored in the reference
//     arithmetic, and the FP64 7-point fit gives z.
#include "lutmatch.hpp"
#include "peakfit.hpp"

#include <cfloat>
#include <type_traits>

namespace qtrk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int PREC> struct MfmaOp;
template <> struct MfmaOp<kMatchBF16> {
	typedef bf16x8 V;
	typedef __bf16 E;
	static __device__ __forceinline__ f32x16 mma(V a, V b, f32x16 c)
	{
		return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
	}
};
template <> struct MfmaOp<kMatchFP16> {
	typedef f16x8 V;
	typedef _Float16 E;
	static __device__ __forceinline__ f32x16 mma(V a, V b, f32x16 c)
	{
		return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
	}
};

// Workgroup -> (bead, profile tile) such that all tiles of one bead run on the
// same XCD (workgroups are dispatched round-robin over the 8 XCDs), so the
// bead's LUT is fetched into one L2 only.  The grid has ceil(beads/8)*8*chunks
// workgroups; those mapped past the last bead exit.
__device__ __forceinline__ void xcd_map(int id, int chunks, int& bead, int& chunk)
{
	const int x = id & 7, q = id >> 3;
	chunk = q % chunks;
	bead = (q / chunks) * 8 + x;
}

// One reference score: float difference, optional weight, negated square and
// sum in double, radius by radius (cputrack/cpu_tracker.cpp:829-839).  pr and
// w are in LDS; the LUT row L is read from global memory 32 radii at a time
// (eight 16-byte loads in flight) when R is a multiple of 4, which keeps the
// row 16-byte aligned.
// Without weights the negated square of the float difference is exact in double
// (24-bit significand squared, far inside the exponent range), so one fused
// multiply-add rounds exactly as the separate product and sum do.
__device__ __forceinline__ void score_step(double& s, float p, float l, const float* w, int r)
{
	double d = (double)(p - l);
	if (w) {
		d *= (double)w[r];
		d = -d * d;
		s += d;
	} else {
		s = __builtin_fma(-d, d, s);
	}
}
__device__ __forceinline__ double exact_score(const float* pr, const float* L, const float* w, int R)
{
	double s = 0;
	int r = 0;
	if ((R & 3) == 0) {
		const float4* L4 = (const float4*)L;
		for (; r + 32 <= R; r += 32) {
			float4 v[8];
#pragma unroll
			for (int j = 0; j < 8; j++) v[j] = L4[(r >> 2) + j];
#pragma unroll
			for (int j = 0; j < 8; j++) {
				const int q = r + 4 * j;
				score_step(s, pr[q], v[j].x, w, q);
				score_step(s, pr[q + 1], v[j].y, w, q + 1);
				score_step(s, pr[q + 2], v[j].z, w, q + 2);
				score_step(s, pr[q + 3], v[j].w, w, q + 3);
			}
		}
	}
	for (; r < R; r++) score_step(s, pr[r], L[r], w, r);
	return s;
}

// ------------------------------------------------------------ top-8 selection

__device__ __forceinline__ void cx(float& a, float& b)
{
	const float hi = fmaxf(a, b), lo = fminf(a, b);
	a = hi;
	b = lo;
}

// Descending sort of 8 values: Batcher's odd-even merge sort (19 comparators).
__device__ __forceinline__ void sort8(float (&v)[8])
{
	cx(v[0], v[1]); cx(v[2], v[3]); cx(v[4], v[5]); cx(v[6], v[7]);
	cx(v[0], v[2]); cx(v[1], v[3]); cx(v[4], v[6]); cx(v[5], v[7]);
	cx(v[1], v[2]); cx(v[5], v[6]);
	cx(v[0], v[4]); cx(v[1], v[5]); cx(v[2], v[6]); cx(v[3], v[7]);
	cx(v[2], v[4]); cx(v[3], v[5]);
	cx(v[1], v[2]); cx(v[3], v[4]); cx(v[5], v[6]);
}

// a <- the 8 largest of two descending lists, descending: max(a_i, b_{7-i}) is
// bitonic and holds exactly those 8; a bitonic merger sorts it.
__device__ __forceinline__ void merge8(float (&a)[8], const float (&b)[8])
{
	float c[8];
#pragma unroll
	for (int i = 0; i < 8; i++) c[i] = fmaxf(a[i], b[7 - i]);
	cx(c[0], c[4]); cx(c[1], c[5]); cx(c[2], c[6]); cx(c[3], c[7]);
	cx(c[0], c[2]); cx(c[1], c[3]); cx(c[4], c[6]); cx(c[5], c[7]);
	cx(c[0], c[1]); cx(c[2], c[3]); cx(c[4], c[5]); cx(c[6], c[7]);
#pragma unroll
	for (int i = 0; i < 8; i++) a[i] = c[i];
}

// Selection key: the approximate score with its low idx_bits mantissa bits
// replaced by the plane index (a relative perturbation below 2^(idx_bits-23),
// far under the BF16/FP16 rounding of the score itself).  -FLT_MAX marks "no
// plane"; a NaN score gives a NaN key, which fmaxf never selects.
__device__ __forceinline__ float make_key(float s, unsigned p, unsigned imask)
{
	return __uint_as_float((__float_as_uint(s) & ~imask) | p);
}

// ------------------------------------------------------------ LDS maps

// Profile rows in LDS: 16-byte aligned, zero-padded past R, and at least the
// padded tile-mean length (4*ceil(R/4) + 4 >= 8*ceil(R/8)).
__host__ __device__ inline int prof_stride(int R) { return ((R + 3) & ~3) + 4; }

struct MfmaLds {
	int prof, w, w2, row, cm, bq, keys, cand, ex, blo, total;
};
__host__ __device__ inline int a16(int x) { return (x + 15) & ~15; }
__host__ __device__ inline MfmaLds mfma_lds(int FT, int R, int KS, int ptiles)
{
	MfmaLds L;
	const int F = FT * 32;
	int o = 0;
	L.prof = o;  o += a16(F * prof_stride(R) * 4);
	L.w = o;     o += a16(KS * 16 * 4);
	L.w2 = o;    o += a16(KS * 16 * 4);
	L.row = o;   o += a16(F * 4);
	L.cm = o;    o += a16(ptiles * F * 4);  // p . M_t for every (tile, profile)
	// the B fragments live until the tile loop ends; the selection and
	// rescoring scratch reuses their space afterwards
	L.bq = o;
	const int ebq = o + a16(FT * KS * 64 * 16);
	L.keys = o;  o += a16(4 * F * 8 * 4);
	L.cand = o;  o += a16(F * 8 * 4);
	L.ex = o;    o += a16(F * 8 * 8);
	// the B fragments' rounding remainders, used before the tile loop only, lie past the B
	// fragments (the selection scratch above is used after the loop only)
	L.blo = ebq;
	const int eblo = ebq + a16(FT * KS * 64 * 16);
	L.total = eblo > o ? eblo : o;
	return L;
}

struct ExactLds {
	int prof, w, sc, row, total;
};
__host__ __device__ inline ExactLds exact_lds(int R, int planes, int tile = kMatchExactTile)
{
	ExactLds L;
	const int R4 = (R + 3) & ~3;
	int o = 0;
	L.prof = o;  o += a16(tile * R4 * 4);
	L.w = o;     o += a16(R4 * 4);
	L.sc = o;    o += a16(tile * planes * 8);
	L.row = o;   o += a16(tile * 4);
	L.total = o;
	return L;
}

// ------------------------------------------------------------ matrix-core path

// One tile of up to FT * 32 profiles of one bead (slots start[bead] + chunk * FT * 32 ...).
template <int PREC, int FT, int KSMAX>
__device__ __forceinline__ void mfma_chunk(const LutMatchParams& P, int bead, int chunk, unsigned char* smem)
{
	typedef typename MfmaOp<PREC>::V V;
	typedef typename MfmaOp<PREC>::E E;
	constexpr int F = FT * 32;  // profiles per workgroup
	const int s0 = P.start[bead] + chunk * FT * 32;
	const int rem = P.start[bead + 1] - s0;
	const int nf = rem < FT * 32 ? rem : FT * 32;
	if (nf <= 0) return;

	const int R = P.res, RS = prof_stride(R), KS = P.ksteps, NP = P.planes;
	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	const int n = lane & 31, h = lane >> 5;
	const MfmaLds L = mfma_lds(FT, R, KS, P.ptiles);
	float* prof_s = (float*)(smem + L.prof);
	float* w_s = (float*)(smem + L.w);
	float* w2_s = (float*)(smem + L.w2);
	int* row_s = (int*)(smem + L.row);
	float* cm_s = (float*)(smem + L.cm);
	uint4* bq_s = (uint4*)(smem + L.bq);
	uint4* blo_s = (uint4*)(smem + L.blo);
	float* keys_s = (float*)(smem + L.keys);
	int* cand_s = (int*)(smem + L.cand);
	double* ex_s = (double*)(smem + L.ex);

	// this wavefront's mean tile (mean tile m, profile tile ft for job m * FT + ft) and, with up to
	// 8 k-steps, its first LUT tile load while the profiles are staged (with 16 k-steps both
	// operand sets would not fit in the registers together: the first tile loads after the
	// tile-mean terms)
	const int wu = __builtin_amdgcn_readfirstlane(wid);
	const int MJ = P.mtiles * FT;
	uint4 am[KSMAX];
	if (wu < MJ) {
		const uint4* mb = P.mfrag + ((size_t)bead * P.mtiles + wu / FT) * KS * 64 + lane;
#pragma unroll
		for (int ks = 0; ks < KSMAX; ks++)
			if (ks < KS) am[ks] = mb[(size_t)ks * 64];
	}
	const uint4* fb = P.frag + (size_t)bead * P.ptiles * KS * 64 + lane;
	uint4 a[KSMAX];
	auto load_first_tile = [&]() {
		if (wu < P.ptiles) {
#pragma unroll
			for (int ks = 0; ks < KSMAX; ks++)
				if (ks < KS) a[ks] = fb[((size_t)wu * KS + ks) * 64];
		}
	};
	if constexpr (KSMAX <= 8) load_first_tile();
	for (int i = tid; i < FT * 32; i += 256) row_s[i] = i < nf ? (P.rowidx ? P.rowidx[s0 + i] : s0 + i) : -1;
	for (int k = tid; k < KS * 16; k += 256) {
		const float wk = k < R ? (P.w ? P.w[k] : 1.0f) : 0.0f;
		w_s[k] = wk;
		w2_s[k] = wk * wk;
	}
	__syncthreads();
	// stage the profiles, zero-padded rows; 16-byte loads, 8 in flight, when rows are 16-byte aligned
	if ((R & 3) == 0) {
		const int q4 = R >> 2, qs = RS >> 2, total = FT * 32 * qs;
		for (int e0 = tid; e0 < total; e0 += 8 * 256) {
			float4 v[8];
#pragma unroll
			for (int j = 0; j < 8; j++) {
				const int e = e0 + j * 256, f = e / qs, q = e - f * qs;
				v[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
				if (e < total && f < nf && q < q4) v[j] = ((const float4*)(P.prof + (size_t)row_s[f] * R))[q];
			}
#pragma unroll
			for (int j = 0; j < 8; j++)
				if (e0 + j * 256 < total) ((float4*)prof_s)[e0 + j * 256] = v[j];
		}
	} else {
		for (int e = tid; e < FT * 32 * RS; e += 256) {
			const int f = e / RS, r = e - f * RS;
			prof_s[e] = (f < nf && r < R) ? P.prof[(size_t)row_s[f] * R + r] : 0.0f;
		}
	}
	__syncthreads();
	const float* wl = P.w ? w_s : nullptr;  // weights for the reference-arithmetic scores

	// B operands in MFMA order, shared by the four wavefronts: fragment (ft, ks)
	// lane l holds profile column l%32 of tile ft, radii 16*ks + 8*(l/32) + [0, 8),
	// weighted by w^2 and rounded to E; the remainders (exact in float) rounded to E
	for (int i = tid; i < FT * KS * 64; i += 256) {
		const int l = i & 63, fk = i >> 6, ks = fk % KS, ft = fk / KS;
		const float* pr = prof_s + (ft * 32 + (l & 31)) * RS;
		const int k0 = ks * 16 + (l >> 5) * 8;
		V v, vl;
#pragma unroll
		for (int j = 0; j < 8; j++) {
			const int k = k0 + j;
			const float x = k < R ? pr[k] * w2_s[k] : 0.0f;
			v[j] = (E)x;
			vl[j] = (E)(x - (float)v[j]);
		}
		bq_s[i] = __builtin_bit_cast(uint4, v);
		blo_s[i] = __builtin_bit_cast(uint4, vl);
	}
	__syncthreads();
	// the tile-mean term p.M_t of every (tile, profile) pair: mean tile m's rows r < 16 (high
	// parts of M_{16m+r}) and 16 + r (low parts) against B and its remainder; lane (n, h) holds
	// rows 8*(i/4) + 4*h + i%4 in accumulator i, so row r's high and low products are
	// accumulators i and i + 8 of the same lane
	auto mean_job = [&](int j) {
		const int ft = j % FT, mt = j / FT;
		f32x16 hi, lo;
#pragma unroll
		for (int i = 0; i < 16; i++) {
			hi[i] = 0.0f;
			lo[i] = 0.0f;
		}
#pragma unroll
		for (int ks = 0; ks < KSMAX; ks++) {
			if (ks < KS) {
				const V av = __builtin_bit_cast(V, am[ks]);
				hi = MfmaOp<PREC>::mma(av, __builtin_bit_cast(V, bq_s[(ft * KS + ks) * 64 + lane]), hi);
				lo = MfmaOp<PREC>::mma(av, __builtin_bit_cast(V, blo_s[(ft * KS + ks) * 64 + lane]), lo);
			}
		}
#pragma unroll
		for (int i = 0; i < 8; i++) {
			const int t = mt * 16 + 8 * (i >> 2) + 4 * h + (i & 3);
			if (t < P.ptiles) cm_s[t * F + ft * 32 + n] = hi[i] + ((hi[i + 8] + lo[i]) + lo[i + 8]);
		}
	};
	if (wu < MJ) mean_job(wu);
	for (int j = wu + 4; j < MJ; j += 4) {  // more than 4 jobs: the operands load here
		const uint4* mb = P.mfrag + ((size_t)bead * P.mtiles + j / FT) * KS * 64 + lane;
#pragma unroll
		for (int ks = 0; ks < KSMAX; ks++)
			if (ks < KS) am[ks] = mb[(size_t)ks * 64];
		mean_job(j);
	}
	if constexpr (KSMAX > 8) load_first_tile();
	__syncthreads();

	const float* nb = P.norm + (size_t)bead * P.ptiles * 32 + 4 * h;
	const unsigned imask = (1u << P.idx_bits) - 1u;
	float top[FT][8];
#pragma unroll
	for (int ft = 0; ft < FT; ft++)
#pragma unroll
		for (int i = 0; i < 8; i++) top[ft][i] = -FLT_MAX;

	for (int t = wu; t < P.ptiles; t += 4) {
		// |L|^2_w of this lane's rows: 8*(i/4) + 4*h + i%4 of the tile
		float nv[16];
#pragma unroll
		for (int j = 0; j < 4; j++) {
			const float4 v = *(const float4*)(nb + t * 32 + 8 * j);
			nv[4 * j] = v.x; nv[4 * j + 1] = v.y; nv[4 * j + 2] = v.z; nv[4 * j + 3] = v.w;
		}
		f32x16 acc[FT];
#pragma unroll
		for (int ft = 0; ft < FT; ft++)
#pragma unroll
			for (int i = 0; i < 16; i++) acc[ft][i] = 0.0f;
#pragma unroll
		for (int ks = 0; ks < KSMAX; ks++) {
			if (ks < KS) {
				const V av = __builtin_bit_cast(V, a[ks]);
#pragma unroll
				for (int ft = 0; ft < FT; ft++)
					acc[ft] = MfmaOp<PREC>::mma(av, __builtin_bit_cast(V, bq_s[(ft * KS + ks) * 64 + lane]), acc[ft]);
			}
		}
		// the next tile's operands load while this tile's epilogue runs
		if (t + 4 < P.ptiles) {
#pragma unroll
			for (int ks = 0; ks < KSMAX; ks++)
				if (ks < KS) a[ks] = fb[((size_t)(t + 4) * KS + ks) * 64];
		}
		// lane (n, h) takes p.M_t of profile ft * 32 + n
		float cm[FT];
#pragma unroll
		for (int ft = 0; ft < FT; ft++) cm[ft] = 2.0f * cm_s[t * F + ft * 32 + n];
#pragma unroll
		for (int ft = 0; ft < FT; ft++) {
			float g0[8], g1[8];
#pragma unroll
			for (int i = 0; i < 16; i++) {
				const int p = t * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
				const float s = 2.0f * acc[ft][i] + cm[ft] - nv[i];
				const float key = p < NP ? make_key(s, (unsigned)p, imask) : -FLT_MAX;
				if (i < 8) g0[i] = key; else g1[i - 8] = key;
			}
			sort8(g0);
			sort8(g1);
			merge8(g0, g1);
			merge8(top[ft], g0);
		}
	}
	__syncthreads();  // the B fragments' space is reused from here on

	// lanes n and n+32 hold the same profile column: merge, then across wavefronts
#pragma unroll
	for (int ft = 0; ft < FT; ft++) {
		float o[8];
#pragma unroll
		for (int i = 0; i < 8; i++) o[i] = __shfl_xor(top[ft][i], 32);
		merge8(top[ft], o);
		if (h == 0)
#pragma unroll
			for (int i = 0; i < 8; i++) keys_s[((wid * FT + ft) * 32 + n) * 8 + i] = top[ft][i];
	}
	__syncthreads();
	if (tid < FT * 32) {
		float m[8];
#pragma unroll
		for (int i = 0; i < 8; i++) m[i] = keys_s[tid * 8 + i];
		for (int w = 1; w < 4; w++) {
			float o[8];
#pragma unroll
			for (int i = 0; i < 8; i++) o[i] = keys_s[(w * FT * 32 + tid) * 8 + i];
			merge8(m, o);
		}
#pragma unroll
		for (int i = 0; i < 8; i++) cand_s[tid * 8 + i] = m[i] > -FLT_MAX ? (int)(__float_as_uint(m[i]) & imask) : -1;
	}
	__syncthreads();

	// rescore the candidates in the reference arithmetic
	const float* Lb = P.lut + (size_t)bead * NP * R;
	for (int it = tid; it < FT * 32 * 8; it += 256) {
		const int f = it >> 3, p = cand_s[it];
		ex_s[it] = (f < nf && p >= 0) ? exact_score(prof_s + f * RS, Lb + (size_t)p * R, wl, R) : 0.0;
	}
	__syncthreads();

	if (tid < nf) {
		const int* cp = cand_s + tid * 8;
		const double* ep = ex_s + tid * 8;
		// best candidate: highest score, the lower plane on ties, NaN never (the
		// reference's strict '>' scan keeps plane 0 when every score is NaN)
		int bi = -1;
		double bv = 0;
		for (int c = 0; c < 8; c++) {
			const int p = cp[c];
			const double v = ep[c];
			if (p < 0 || v != v) continue;
			if (bi < 0 || v > bv || (v == bv && p < bi)) { bi = p; bv = v; }
		}
		if (bi < 0) bi = 0;
		const float* pr = prof_s + tid * RS;
		const double zf = peak_fit<double>(bi, NP, kZW, [&](int k) -> double {
			for (int c = 0; c < 8; c++)
				if (cp[c] == k) return ep[c];
			return exact_score(pr, Lb + (size_t)k * R, wl, R);
		});
		const int row = row_s[tid];
		P.z[(size_t)row * P.zstride] = (float)zf;
		if (P.best) P.best[row] = bi;
		if (P.cand)
			for (int c = 0; c < 8; c++) P.cand[(size_t)row * 8 + c] = cp[c];
	}
}

// Workgroup (bead, chunk) takes the bead's profile tile `chunk` (the grid sized for the largest
// bead) or, with LOOP, the tiles chunk, chunk + chunks, ... (a grid sized without knowing the
// largest bead).  The loop is a separate instantiation: around the tile it keeps 39 more VGPRs
// live (159 against 120), which costs a wavefront per SIMD.
template <int PREC, int FT, int KSMAX, bool LOOP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void lutmatch_mfma_kernel(LutMatchParams P)
{
	extern __shared__ __align__(16) unsigned char smem[];
	int bead, chunk;
	xcd_map(blockIdx.x, P.chunks, bead, chunk);
	if (bead >= P.beads) return;
	const int count = P.start[bead + 1] - P.start[bead];
	if (!LOOP) {
		if (chunk * FT * 32 < count) mfma_chunk<PREC, FT, KSMAX>(P, bead, chunk, smem);
		return;
	}
	for (; chunk * FT * 32 < count; chunk += P.chunks) {
		mfma_chunk<PREC, FT, KSMAX>(P, bead, chunk, smem);
		__syncthreads();
	}
}

// ------------------------------------------------------------ exact path

// Up to TILE profiles of one bead per workgroup; lane per plane, all TILE scores
// accumulated together so each LUT row is read once per TILE profiles.
template <bool WEIGHTED, int TILE>
__device__ __forceinline__ void exact_chunk(const LutMatchParams& P, int bead, int chunk, unsigned char* smem)
{
	const int s0 = P.start[bead] + chunk * TILE;
	const int rem = P.start[bead + 1] - s0;
	const int nf = rem < TILE ? rem : TILE;
	if (nf <= 0) return;

	const int R = P.res, R4 = (R + 3) & ~3, NP = P.planes;
	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	const ExactLds L = exact_lds(R, NP, TILE);
	float* prof_s = (float*)(smem + L.prof);
	float* w_s = (float*)(smem + L.w);
	double* sc_s = (double*)(smem + L.sc);
	int* row_s = (int*)(smem + L.row);

	if (tid < TILE) row_s[tid] = tid < nf ? (P.rowidx ? P.rowidx[s0 + tid] : s0 + tid) : -1;
	if (WEIGHTED)
		for (int r = tid; r < R4; r += 256) w_s[r] = r < R ? P.w[r] : 0.0f;
	__syncthreads();
	for (int e = tid; e < TILE * R4; e += 256) {
		const int f = e / R4, r = e - f * R4;
		prof_s[e] = (f < nf && r < R) ? P.prof[(size_t)row_s[f] * R + r] : 0.0f;
	}
	__syncthreads();

	const float* Lb = P.lut + (size_t)bead * NP * R;
	for (int p = tid; p < NP; p += 256) {
		const float* Lr = Lb + (size_t)p * R;
		double acc[TILE];
#pragma unroll
		for (int f = 0; f < TILE; f++) acc[f] = 0.0;
		// radius r for all 16 profiles: float difference, weight, negated square, double sum
		// (unweighted: one exact fused multiply-add, as in score_step)
		auto step = [&](int r, float l) {
			const double wr = WEIGHTED ? (double)w_s[r] : 1.0;
#pragma unroll
			for (int f = 0; f < TILE; f++) {
				double d = (double)(prof_s[f * R4 + r] - l);
				if (WEIGHTED) {
					d *= wr;
					d = -d * d;
					acc[f] += d;
				} else {
					acc[f] = __builtin_fma(-d, d, acc[f]);
				}
			}
		};
		int r = 0;
		if ((R & 7) == 0) {  // 16-byte aligned rows: the next 8 radii load while 8 are scored
			const float4* L4 = (const float4*)Lr;
			const int nq = R >> 2;
			float4 c0 = L4[0], c1 = L4[1];
			for (int q = 0; q < nq; q += 2) {
				float4 n0 = c0, n1 = c1;
				if (q + 2 < nq) {
					n0 = L4[q + 2];
					n1 = L4[q + 3];
				}
				const int rr = q * 4;
				step(rr, c0.x); step(rr + 1, c0.y); step(rr + 2, c0.z); step(rr + 3, c0.w);
				step(rr + 4, c1.x); step(rr + 5, c1.y); step(rr + 6, c1.z); step(rr + 7, c1.w);
				c0 = n0;
				c1 = n1;
			}
			r = R;
		}
		for (; r < R; r++) step(r, Lr[r]);
#pragma unroll
		for (int f = 0; f < TILE; f++)
			if (f < nf) sc_s[f * NP + p] = acc[f];
	}
	__syncthreads();

	// argmax with the reference's strict '>' scan from plane 0 (NaN scores never
	// win; a NaN at plane 0 keeps plane 0), then the FP64 fit
	for (int f = wid; f < nf; f += 4) {
		const double* sc = sc_s + f * NP;
		double bv = 0;
		int bi = -1;
		for (int p = lane; p < NP; p += 64) {
			const double v = sc[p];
			if (v != v) continue;
			if (bi < 0 || v > bv) { bv = v; bi = p; }
		}
		for (int off = 32; off > 0; off >>= 1) {
			const double ov = __shfl_xor(bv, off);
			const int oi = __shfl_xor(bi, off);
			if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) { bv = ov; bi = oi; }
		}
		if (lane == 0) {
			if (bi < 0 || sc[0] != sc[0]) bi = 0;
			const double zf = peak_fit<double>(bi, NP, kZW, [&](int k) { return sc[k]; });
			const int row = row_s[f];
			P.z[(size_t)row * P.zstride] = (float)zf;
			if (P.best) P.best[row] = bi;
		}
	}
}

template <bool WEIGHTED, bool LOOP, int TILE>
__global__ __launch_bounds__(256) void lutmatch_exact_kernel(LutMatchParams P)
{
	extern __shared__ __align__(16) unsigned char smem[];
	int bead, chunk;
	xcd_map(blockIdx.x, P.chunks, bead, chunk);
	if (bead >= P.beads) return;
	const int count = P.start[bead + 1] - P.start[bead];
	if (!LOOP) {
		if (chunk * TILE < count) exact_chunk<WEIGHTED, TILE>(P, bead, chunk, smem);
		return;
	}
	for (; chunk * TILE < count; chunk += P.chunks) {
		exact_chunk<WEIGHTED, TILE>(P, bead, chunk, smem);
		__syncthreads();
	}
}

// ------------------------------------------------------------ preparation

// Tile means: raw[b][t][r] = mean of the tile's valid rows (mres-long rows, zero past res).
__global__ void lutmatch_mean_kernel(LutMatchParams P, float* raw, int mres)
{
	const size_t total = (size_t)P.beads * P.ptiles * mres;
	for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
		const int r = (int)(g % mres);
		const size_t q = g / mres;
		const int t = (int)(q % P.ptiles);
		const size_t b = q / P.ptiles;
		const int p0 = t * 32;
		const int cnt = P.planes - p0 < 32 ? P.planes - p0 : 32;
		float s = 0.0f;
		if (r < P.res) {
			for (int i = 0; i < cnt; i++) s += P.lut[(b * P.planes + p0 + i) * P.res + r];
			s /= (float)cnt;
		}
		raw[g] = s;
	}
}

// Tile means as A-operand fragments: lane (row = lane % 32, radii 8*(lane/32) + [0, 8)) of each
// (mean tile m, 16-radius step) holds, for row j < 16, M_{16m+j} rounded to E, and for row j >= 16
// the remainder M_{16m+j-16} - (float)E(M) rounded to E; rows of tiles past ptiles are 0.
template <typename E>
__global__ void lutmatch_meanfrag_kernel(LutMatchParams P, uint4* mfrag, const float* raw, int mres)
{
	typedef E V8 __attribute__((ext_vector_type(8)));
	const size_t total = (size_t)P.beads * P.mtiles * P.ksteps * 64;
	for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
		const int lane = (int)(g & 63);
		size_t q = g >> 6;
		const int ks = (int)(q % P.ksteps);
		q /= P.ksteps;
		const int m = (int)(q % P.mtiles);
		const size_t b = q / P.mtiles;
		const int j = lane & 31, t = m * 16 + (j & 15);
		const int k0 = ks * 16 + (lane >> 5) * 8;
		V8 v;
#pragma unroll
		for (int i = 0; i < 8; i++) {
			const int k = k0 + i;
			const float x = (t < P.ptiles && k < P.res) ? raw[(b * P.ptiles + t) * mres + k] : 0.0f;
			const E hi = (E)x;
			v[i] = j < 16 ? hi : (E)(x - (float)hi);
		}
		mfrag[g] = __builtin_bit_cast(uint4, v);
	}
}

// Residual fragments: lane (row = lane % 32, radii 8*(lane/32) + [0, 8)) of each
// (tile, 16-radius step) holds L - M_t rounded to E; padded rows and radii are 0.
template <typename E>
__global__ void lutmatch_frag_kernel(LutMatchParams P, uint4* frag, const float* raw, int mres)
{
	typedef E V8 __attribute__((ext_vector_type(8)));
	const size_t total = (size_t)P.beads * P.ptiles * P.ksteps * 64;
	for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
		const int lane = (int)(g & 63);
		size_t q = g >> 6;
		const int ks = (int)(q % P.ksteps);
		q /= P.ksteps;
		const int t = (int)(q % P.ptiles);
		const size_t b = q / P.ptiles;
		const int row = t * 32 + (lane & 31);
		const int k0 = ks * 16 + (lane >> 5) * 8;
		const float* mt = raw + (b * P.ptiles + t) * mres;
		V8 v;
#pragma unroll
		for (int j = 0; j < 8; j++) {
			const int k = k0 + j;
			v[j] = (E)((row < P.planes && k < P.res) ? P.lut[(b * P.planes + row) * P.res + k] - mt[k] : 0.0f);
		}
		frag[g] = __builtin_bit_cast(uint4, v);
	}
}

__global__ void lutmatch_norm_kernel(LutMatchParams P, float* norm)
{
	const int PP = P.ptiles * 32;
	const size_t total = (size_t)P.beads * PP;
	for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
		const int p = (int)(g % PP);
		const size_t b = g / PP;
		float s = 0.0f;
		if (p < P.planes) {
			const float* Lr = P.lut + (b * P.planes + p) * P.res;
			int r = 0;
			if ((P.res & 3) == 0 && ((uintptr_t)Lr & 15) == 0) {
				// the row in 16-byte loads, 8 in flight; the sum keeps its order
				for (; r + 32 <= P.res; r += 32) {
					float4 v[8];
#pragma unroll
					for (int j = 0; j < 8; j++) v[j] = ((const float4*)(Lr + r))[j];
#pragma unroll
					for (int j = 0; j < 8; j++) {
						const float e[4] = { v[j].x, v[j].y, v[j].z, v[j].w };
#pragma unroll
						for (int k = 0; k < 4; k++) {
							const int rr = r + 4 * j + k;
							const float w2 = P.w ? P.w[rr] * P.w[rr] : 1.0f;
							s += w2 * e[k] * e[k];
						}
					}
				}
			}
			for (; r < P.res; r++) {
				const float w2 = P.w ? P.w[r] * P.w[r] : 1.0f;
				s += w2 * Lr[r] * Lr[r];
			}
		}
		norm[g] = s;
	}
}

static int grid_for(size_t n) { return (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536); }

hipError_t launch_lutmatch_prepare(const LutMatchParams& p, int prec, uint4* frag, uint4* mfrag, float* norm,
                                   float* raw_mean, int mres, hipStream_t stream)
{
	if (prec != kMatchBF16 && prec != kMatchFP16) return hipErrorInvalidValue;
	lutmatch_mean_kernel<<<grid_for((size_t)p.beads * p.ptiles * mres), 256, 0, stream>>>(p, raw_mean, mres);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	const size_t nfrag = (size_t)p.beads * p.ptiles * p.ksteps * 64;
	const size_t nmfrag = (size_t)p.beads * p.mtiles * p.ksteps * 64;
	if (prec == kMatchBF16) {
		lutmatch_frag_kernel<__bf16><<<grid_for(nfrag), 256, 0, stream>>>(p, frag, raw_mean, mres);
		lutmatch_meanfrag_kernel<__bf16><<<grid_for(nmfrag), 256, 0, stream>>>(p, mfrag, raw_mean, mres);
	} else {
		lutmatch_frag_kernel<_Float16><<<grid_for(nfrag), 256, 0, stream>>>(p, frag, raw_mean, mres);
		lutmatch_meanfrag_kernel<_Float16><<<grid_for(nmfrag), 256, 0, stream>>>(p, mfrag, raw_mean, mres);
	}
	e = hipGetLastError();
	if (e != hipSuccess) return e;
	lutmatch_norm_kernel<<<grid_for((size_t)p.beads * p.ptiles * 32), 256, 0, stream>>>(p, norm);
	return hipGetLastError();
}

template <int PREC, int FT, int KSMAX, bool LOOP>
static hipError_t launch_mfma(const LutMatchParams& q, int grid, hipStream_t stream)
{
	const size_t lds = (size_t)mfma_lds(FT, q.res, q.ksteps, q.ptiles).total;
	hipError_t e = hipFuncSetAttribute((const void*)lutmatch_mfma_kernel<PREC, FT, KSMAX, LOOP>,
	                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
	if (e != hipSuccess) return e;
	lutmatch_mfma_kernel<PREC, FT, KSMAX, LOOP><<<grid, 256, lds, stream>>>(q);
	return hipGetLastError();
}

template <int PREC, bool LOOP>
static hipError_t launch_mfma_loop(const LutMatchParams& q, int FT, int grid, hipStream_t stream)
{
	if (q.ksteps <= 8)
		return FT == 1 ? launch_mfma<PREC, 1, 8, LOOP>(q, grid, stream) : launch_mfma<PREC, 2, 8, LOOP>(q, grid, stream);
	return FT == 1 ? launch_mfma<PREC, 1, 16, LOOP>(q, grid, stream) : launch_mfma<PREC, 2, 16, LOOP>(q, grid, stream);
}

template <int PREC>
static hipError_t launch_mfma_prec(const LutMatchParams& q, int FT, int grid, hipStream_t stream, bool loop)
{
	return loop ? launch_mfma_loop<PREC, true>(q, FT, grid, stream) : launch_mfma_loop<PREC, false>(q, FT, grid, stream);
}

template <bool WEIGHTED, bool LOOP, int TILE>
static hipError_t launch_exact(const LutMatchParams& q, int grid, size_t lds, hipStream_t stream)
{
	hipError_t e = hipFuncSetAttribute((const void*)lutmatch_exact_kernel<WEIGHTED, LOOP, TILE>,
	                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
	if (e != hipSuccess) return e;
	lutmatch_exact_kernel<WEIGHTED, LOOP, TILE><<<grid, 256, lds, stream>>>(q);
	return hipGetLastError();
}

// Profiles per workgroup of the exact path: 8 (each LUT row read once per 8; measured on 1,024 beads x
// 500 planes x 128 steps x 64 profiles: 0.51 ms, against 0.54 ms with 16 (214 VGPRs, two wavefronts
// per SIMD) and 0.665 ms with 4), or 4 when 8 would leave fewer than two workgroups per CU (a few
// profiles per bead, as in one camera batch) or when 8 rows of scores would not fit the LDS (many
// planes).
static int exact_tile(int beads, int max_per_bead, int res, int planes)
{
	int tile = kMatchExactTile;
	while (tile > 4 && ((long long)beads * ((max_per_bead + tile - 1) / tile) < 2 * kMatchCUs ||
	                    exact_lds(res, planes, tile).total > 160 * 1024))
		tile /= 2;
	return tile;
}

hipError_t launch_lutmatch(const LutMatchParams& p, int prec, int max_per_bead, hipStream_t stream, int frame_tile,
                           bool loop_tiles)
{
	if (p.beads <= 0 || max_per_bead <= 0) return hipSuccess;
	LutMatchParams q = p;
	if (q.zstride <= 0) q.zstride = 1;
	if (prec == kMatchFP32) {
		const int tile = exact_tile(p.beads, max_per_bead, p.res, p.planes);
		q.chunks = (max_per_bead + tile - 1) / tile;
		const int grid = (p.beads + 7) / 8 * 8 * q.chunks;
		const size_t lds = (size_t)exact_lds(p.res, p.planes, tile).total;
		if (lds > 160 * 1024) return hipErrorInvalidValue;
		auto go = [&](auto T) -> hipError_t {
			constexpr int TL = decltype(T)::value;
			if (p.w)
				return loop_tiles ? launch_exact<true, true, TL>(q, grid, lds, stream)
				                  : launch_exact<true, false, TL>(q, grid, lds, stream);
			return loop_tiles ? launch_exact<false, true, TL>(q, grid, lds, stream)
			                  : launch_exact<false, false, TL>(q, grid, lds, stream);
		};
		if (tile == 8) return go(std::integral_constant<int, 8>());
		return go(std::integral_constant<int, 4>());
	}
	if (p.res > kMatchMaxRes) return hipErrorInvalidValue;
	if (frame_tile != 0 && frame_tile != 32 && frame_tile != 64) return hipErrorInvalidValue;
	const int FT = frame_tile ? frame_tile / 32 : 1;
	q.chunks = (max_per_bead + FT * 32 - 1) / (FT * 32);
	const int grid = (p.beads + 7) / 8 * 8 * q.chunks;
	if (prec == kMatchBF16) return launch_mfma_prec<kMatchBF16>(q, FT, grid, stream, loop_tiles);
	if (prec == kMatchFP16) return launch_mfma_prec<kMatchFP16>(q, FT, grid, stream, loop_tiles);
	return hipErrorInvalidValue;
}

size_t lutmatch_lds_bytes(int prec, int res, int planes, int ksteps, int max_per_bead, int frame_tile, int beads)
{
	if (prec == kMatchFP32) return (size_t)exact_lds(res, planes, exact_tile(beads, max_per_bead, res, planes)).total;
	const int FT = frame_tile ? frame_tile / 32 : 1;
	return (size_t)mfma_lds(FT, res, ksteps, (planes + 31) / 32).total;
}

// ------------------------------------------------------- grouping by bead

// Profile rows grouped by bead on the GPU: counts[b] = rows of bead b (index in [0, beads); others
// belong to no group), start = exclusive prefix of the counts, rowidx[start[b] ...] = bead b's rows.
// Rows are placed with atomics, so their order within a bead is arbitrary: every profile's match
// is computed on its own, so its result does not depend on its slot.
__global__ void group_zero_kernel(int* counts, int n)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) counts[i] = 0;
}

__global__ void group_count_kernel(const int* bead_idx, int stride, int n, int beads, int* counts)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int b = bead_idx[(size_t)i * stride];
	if (b >= 0 && b < beads) atomicAdd(&counts[b], 1);
}

// One workgroup: start[0..beads] and counts[] turned into each bead's fill cursor (= start[b]).
__global__ __launch_bounds__(1024) void group_scan_kernel(int* counts, int beads, int* start)
{
	__shared__ int part[1024];
	const int tid = threadIdx.x, per = (beads + 1023) / 1024, b0 = tid * per;
	int s = 0;
	for (int b = b0; b < b0 + per && b < beads; b++) s += counts[b];
	part[tid] = s;
	__syncthreads();
	for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the per-thread sums
		const int v = tid >= off ? part[tid - off] : 0;
		__syncthreads();
		part[tid] += v;
		__syncthreads();
	}
	int run = tid ? part[tid - 1] : 0;
	for (int b = b0; b < b0 + per && b < beads; b++) {
		const int c = counts[b];
		start[b] = run;
		counts[b] = run;
		run += c;
	}
	if (tid == 1023) start[beads] = part[1023];
}

__global__ void group_scatter_kernel(const int* bead_idx, int stride, int n, int beads, int* cursor, int* rowidx)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int b = bead_idx[(size_t)i * stride];
	if (b >= 0 && b < beads) rowidx[atomicAdd(&cursor[b], 1)] = i;
}

// The same grouping in one workgroup, the counts in LDS: one launch instead of four, for batches of
// up to kGroupSmallRows rows and kGroupSmallBeads beads.  Each thread's bead indices are loaded 8 at a
// time before their atomics (one memory round trip per 8 rows), and the bead counts are scanned with
// wavefront shuffles and one pass over the 16 wavefront totals.
constexpr int kGroupSmallRows = 65536, kGroupSmallBeads = 16384;

template <typename F>
__device__ __forceinline__ void for_rows_by_8(const int* bead_idx, int stride, int n, int beads, int tid, F&& f)
{
	constexpr int K = 8;
	for (int i0 = 0; i0 < n; i0 += 1024 * K) {
		int b[K];
#pragma unroll
		for (int k = 0; k < K; k++) {
			const int i = i0 + k * 1024 + tid;
			b[k] = i < n ? bead_idx[(size_t)i * stride] : -1;
		}
#pragma unroll
		for (int k = 0; k < K; k++)
			if (b[k] >= 0 && b[k] < beads) f(b[k], i0 + k * 1024 + tid);
	}
}

__global__ __launch_bounds__(1024) void group_small_kernel(const int* bead_idx, int stride, int n, int beads,
                                                           int* start, int* rowidx)
{
	extern __shared__ int cnt[];  // [beads], then [16] wavefront totals
	int* wsum = cnt + beads;
	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	for (int b = tid; b < beads; b += 1024) cnt[b] = 0;
	__syncthreads();
	for_rows_by_8(bead_idx, stride, n, beads, tid, [&](int b, int) { atomicAdd(&cnt[b], 1); });
	__syncthreads();
	// exclusive scan of the counts: thread t owns beads t*per .. t*per+per-1
	const int per = (beads + 1023) / 1024, b0 = tid * per;
	int sum = 0;
	for (int b = b0; b < b0 + per && b < beads; b++) sum += cnt[b];
	int incl = sum;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1) {
		const int v = __shfl_up(incl, off, 64);
		if (lane >= off) incl += v;
	}
	if (lane == 63) wsum[wid] = incl;
	__syncthreads();
	int wbase = 0;
	for (int w = 0; w < wid; w++) wbase += wsum[w];
	int run = wbase + incl - sum;
	for (int b = b0; b < b0 + per && b < beads; b++) {
		const int c = cnt[b];
		start[b] = run;
		cnt[b] = run;
		run += c;
	}
	if (tid == 1023) start[beads] = wbase + incl;
	__syncthreads();
	for_rows_by_8(bead_idx, stride, n, beads, tid, [&](int b, int i) { rowidx[atomicAdd(&cnt[b], 1)] = i; });
}

hipError_t launch_group_by_bead(const int* bead_idx, int stride, int n, int beads, int* counts, int* start,
                                int* rowidx, hipStream_t stream)
{
	if (beads <= 0) return hipSuccess;
	if (n <= kGroupSmallRows && beads <= kGroupSmallBeads) {
		const size_t lds = ((size_t)beads + 16) * sizeof(int);
		hipError_t e = hipFuncSetAttribute((const void*)group_small_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   (int)lds);
		if (e != hipSuccess) return e;
		group_small_kernel<<<1, 1024, lds, stream>>>(bead_idx, stride, n, beads, start, rowidx);
		return hipGetLastError();
	}
	group_zero_kernel<<<grid_for((size_t)beads), 256, 0, stream>>>(counts, beads);
	if (n > 0) group_count_kernel<<<(n + 255) / 256, 256, 0, stream>>>(bead_idx, stride, n, beads, counts);
	group_scan_kernel<<<1, 1024, 0, stream>>>(counts, beads, start);
	if (n > 0) group_scatter_kernel<<<(n + 255) / 256, 256, 0, stream>>>(bead_idx, stride, n, beads, counts, rowidx);
	return hipGetLastError();
}

}  // namespace qtrk
