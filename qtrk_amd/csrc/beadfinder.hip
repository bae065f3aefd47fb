// beadfinder.hip -- QTrkFindBeads: bead detection by cross-correlating a camera
// frame with a sample bead (cputrack/qtrk_c_api.cpp:33-58,
// cputrack/BeadFinder.cpp:163-245).
//
// On the GPU: the frame's mean (a float sum in index order, as ImageData::mean,
// cputrack/utils.h:103-108), the frame and the sample laid into complex arrays
// of the next power-of-two size, their 2D transforms (kissfft of every row,
// then of every column: FFT2D, cputrack/BeadFinder.cpp:45-79), the spectra's
// product, the inverse 2D transform (rows, then columns, as the reference), the
// real part and its maximum; then each candidate's re-centring sums
// (RecenterAndFilter, cputrack/BeadFinder.cpp:105-161), one workgroup per
// candidate, the sums in the reference's pixel order on one lane.
//
// On the host: the raster scan that starts a neighbourhood search at every
// pixel above the threshold and clears the neighbourhood (SearchArea,
// cputrack/BeadFinder.cpp:81-103), and the list filter that drops pairs of
// beads closer than the minimum distance.  Both depend on the order of the
// candidates found before them, which is sequential by definition.
//
// The transforms use the same butterflies, twiddles and leaf order as the
// localization kernel's (fft_device.hpp), so the correlation map equals the
// reference's bit for bit, and with it every threshold decision.
#include "ImageData.h"
#include "beadfinder.hpp"
#include "devbuf.hpp"
#include "errors.hpp"
#include "fft_device.hpp"
#include "hip_tracker.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <list>
#include <vector>

namespace qtrk {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 4096;  // floats staged per step of the ordered sum

// ImageData::mean over n floats: the sum in index order, then / n.  Waves 1-3
// stage the data into LDS one chunk ahead; lane 0 of wave 0 adds.
__global__ __launch_bounds__(kThreads) void ordered_mean_kernel(const float* data, int n, float* out)
{
	__shared__ __align__(16) float buf[2][kChunk];
	const int tid = threadIdx.x;
	const int nchunks = (n + kChunk - 1) / kChunk;
	if (tid >= 64)
		for (int i = tid - 64; i < kChunk && i < n; i += kThreads - 64) buf[0][i] = data[i];
	__syncthreads();
	float s = 0.0f;
	for (int c = 0; c < nchunks; c++) {
		const int nxt = (c + 1) * kChunk;
		if (tid >= 64 && c + 1 < nchunks) {
			const int len = min(kChunk, n - nxt);
			for (int i = tid - 64; i < len; i += kThreads - 64) buf[(c + 1) & 1][i] = data[nxt + i];
		}
		if (tid == 0) {
			// 16 values a step, the next 16 read while these are added: the additions wait on
			// each other only, not on the LDS
			const float* b = buf[c & 1];
			const float4* b4 = (const float4*)b;
			const int len = min(kChunk, n - c * kChunk), nq = len >> 4;
			float4 a0, a1, a2, a3;
			if (nq > 0) {
				a0 = b4[0];
				a1 = b4[1];
				a2 = b4[2];
				a3 = b4[3];
			}
			for (int q = 0; q < nq; q++) {
				const int nx = q + 1 < nq ? 4 * (q + 1) : 4 * q;
				const float4 c0 = b4[nx], c1 = b4[nx + 1], c2 = b4[nx + 2], c3 = b4[nx + 3];
				s += a0.x; s += a0.y; s += a0.z; s += a0.w;
				s += a1.x; s += a1.y; s += a1.z; s += a1.w;
				s += a2.x; s += a2.y; s += a2.z; s += a2.w;
				s += a3.x; s += a3.y; s += a3.z; s += a3.w;
				a0 = c0;
				a1 = c1;
				a2 = c2;
				a3 = c3;
			}
			for (int i = nq << 4; i < len; i++) s += b[i];
		}
		__syncthreads();
	}
	if (tid == 0) out[0] = s / (float)n;
}

// The frame minus its mean in the top-left w x h of a W2 x H2 complex array, zeros elsewhere
// (cputrack/BeadFinder.cpp:172-178).
__global__ void fill_frame_kernel(const float* img, int w, int h, const float* mean, float2* out, int W2, int H2)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (size_t)W2 * H2) return;
	const int x = (int)(i % W2), y = (int)(i / W2);
	out[i] = (x < w && y < h) ? make_float2(img[(size_t)y * w + x] - mean[0], 0.0f) : make_float2(0.0f, 0.0f);
}

// The sample minus its mean, pixel (x, y) at ((y - roi) mod H2, (x - roi) mod W2)
// (cputrack/BeadFinder.cpp:184-189); roi <= W2, H2, so each array pixel holds at most one.
__global__ void fill_sample_kernel(const float* smp, int roi, const float* smean, float2* out, int W2, int H2)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (size_t)W2 * H2) return;
	const int X = (int)(i % W2), Y = (int)(i / W2);
	const int x = (X + roi) & (W2 - 1), y = (Y + roi) & (H2 - 1);
	out[i] = (x < roi && y < roi) ? make_float2(smp[y * roi + x] - smean[0], 0.0f) : make_float2(0.0f, 0.0f);
}

// One kissfft transform per row of length n (a power of two): the row in LDS in leaf order,
// the stages from the deepest out, one thread per butterfly.
__global__ __launch_bounds__(kThreads) void fft_rows_kernel(float2* data, FFTPlan F, const float2* tw, float nii)
{
	extern __shared__ float2 X[];
	const int n = F.n, tid = threadIdx.x;
	float2* row = data + (size_t)blockIdx.x * n;
	for (int i = tid; i < n; i += kThreads) X[i] = row[F.perm[i]];
	__syncthreads();
	for (int s = F.nst - 1; s >= 0; s--) {
		const int r = F.radix[s], lr = r == 4 ? 2 : 1;
		const int lm = __builtin_ctz(F.m[s]), fs = F.fs[s];
		const int nb = n >> lr;
		for (int e = tid; e < nb; e += kThreads) {
			const int g = e >> lm, k = e & ((1 << lm) - 1);
			bfly_pow2(X + (g << (lm + lr)), r, k, 1 << lm, k * fs, tw, nii);
		}
		__syncthreads();
	}
	for (int i = tid; i < n; i += kThreads) row[i] = X[i];
}

// out[x][y] = in[y][x] for an H x W array (32 x 32 tiles through LDS).
__global__ __launch_bounds__(kThreads) void transpose_kernel(const float2* in, float2* out, int W, int H)
{
	__shared__ float2 t[32][33];
	const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for (int j = ty; j < 32; j += 8)
		if (by + j < H && bx + tx < W) t[j][tx] = in[(size_t)(by + j) * W + bx + tx];
	__syncthreads();
	for (int j = ty; j < 32; j += 8)
		if (bx + j < W && by + tx < H) out[(size_t)(bx + j) * H + by + tx] = t[tx][j];
}

__global__ void product_kernel(float2* a, const float2* b, size_t n)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = cmul(a[i], b[i]);
}

// map[y][x] = in[x][y].real for the transposed W x H result, and the maximum of
// std::max(0, v) over all of it: the reference's max starts at 0 and a NaN never
// replaces it; for floats >= 0 the bit patterns order like the values.
__global__ __launch_bounds__(kThreads) void real_map_kernel(const float2* in, float* map, int W, int H, unsigned* maxbits)
{
	__shared__ float t[32][33];
	const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx over y (rows of in), by over x
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	unsigned best = 0;
	for (int j = ty; j < 32; j += 8)
		if (by + j < W && bx + tx < H) {
			const float v = in[(size_t)(by + j) * H + bx + tx].x;
			t[j][tx] = v;
			if (v > 0.0f) best = max(best, __float_as_uint(v));
		}
	__syncthreads();
	for (int j = ty; j < 32; j += 8)
		if (bx + j < H && by + tx < W) map[(size_t)(bx + j) * W + by + tx] = t[tx][j];
	for (int off = 32; off > 0; off >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, off));
	if ((threadIdx.x & 63) == 0 && best) atomicMax(maxbits, best);
}

// RecenterAndFilter's per-bead part (cputrack/BeadFinder.cpp:114-145) for the candidates inside
// the frame: the roi x roi block in LDS, then on lane 0 the float mean and the squared-deviation
// moments in the reference's pixel order.  out[i] = the moved position, or (INT_MIN, 0) when the
// candidate is skipped (outside the frame, or no deviation).
__global__ __launch_bounds__(kThreads) void recenter_kernel(const float* img, int w, int h, int roi,
                                                            const int2* pts, int2* out)
{
	extern __shared__ float blk[];
	const int2 p = pts[blockIdx.x];
	const int tid = threadIdx.x;
	const bool inside = !(p.x < 0 || p.y < 0 || p.x > w - 1 - roi || p.y > h - 1 - roi);
	if (!inside) {
		if (tid == 0) out[blockIdx.x] = make_int2(INT_MIN, 0);
		return;
	}
	for (int i = tid; i < roi * roi; i += kThreads) {
		const int y = i / roi, x = i - y * roi;
		blk[i] = img[(size_t)(y + p.y) * w + x + p.x];
	}
	__syncthreads();
	if (tid != 0) return;
	const int n = roi * roi;
	float sum = 0.0f;
	for (int i = 0; i < n; i++) sum += blk[i];
	const float mean = sum / (float)n;
	float s = 0.0f, sx = 0.0f, sy = 0.0f;
	for (int y = 0; y < roi; y++)
		for (int x = 0; x < roi; x++) {
			float v = blk[y * roi + x] - mean;
			v = v * v;
			s += v;
			sx += v * (float)x;
			sy += v * (float)y;
		}
	if (s == 0.0f) {
		out[blockIdx.x] = make_int2(INT_MIN, 0);
		return;
	}
	const float half = (float)(roi / 2);
	// bp.x += sumX / sum - roi / 2: an int plus a float, truncated back to int
	out[blockIdx.x] = make_int2((int)((float)p.x + (sx / s - half)), (int)((float)p.y + (sy / s - half)));
}

int next_pow2(int x)
{
	int y = 1;
	while (y < x) y *= 2;
	return y;
}

unsigned grid1(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

std::vector<ROIPosition> find_beads(const float* img, int w, int h, int sx, int sy, int roi, float img_distance,
                                    float similarity, float* sample_out)
{
	if (!img) fail("null image");
	if (w < 1 || h < 1) fail("image width and height must be >= 1");
	if (roi < 1) fail("roi must be >= 1");
	if (roi > kMaxBeadFinderROI) fail("roi larger than " + std::to_string(kMaxBeadFinderROI) + " pixels");
	if (sx < 0 || sy < 0 || sx + roi > w || sy + roi > h) fail("the sample ROI is not inside the image");
	const int W2 = next_pow2(w), H2 = next_pow2(h);
	if (W2 > kMaxBeadFinderSide || H2 > kMaxBeadFinderSide)
		fail("image larger than " + std::to_string(kMaxBeadFinderSide) + " pixels a side");
	const size_t npx = (size_t)w * h, N2 = (size_t)W2 * H2;

	// the sample: ImageData::subimage of the frame (cputrack/utils.h:115-128)
	std::vector<float> smp((size_t)roi * roi);
	for (int y = 0; y < roi; y++)
		for (int x = 0; x < roi; x++) smp[(size_t)y * roi + x] = img[(size_t)(y + sy) * w + x + sx];
	if (sample_out) std::copy(smp.begin(), smp.end(), sample_out);

	DeviceBuffer d_img, d_smp, d_means, d_a, d_b, d_c, d_map, d_max, d_pts, d_out;
	d_img.ensure(npx * 4);
	d_smp.ensure(smp.size() * 4);
	d_means.ensure(8);
	d_a.ensure(N2 * 8);
	d_b.ensure(N2 * 8);
	d_c.ensure(N2 * 8);
	d_map.ensure(N2 * 4);
	d_max.ensure(4);
	HIP_CHECK(hipMemcpy(d_img.p, img, npx * 4, hipMemcpyHostToDevice));
	HIP_CHECK(hipMemcpy(d_smp.p, smp.data(), smp.size() * 4, hipMemcpyHostToDevice));
	float* means = d_means.as<float>();
	hipLaunchKernelGGL(ordered_mean_kernel, dim3(1), dim3(kThreads), 0, nullptr, d_img.as<float>(), (int)npx, means);
	hipLaunchKernelGGL(ordered_mean_kernel, dim3(1), dim3(kThreads), 0, nullptr, d_smp.as<float>(), roi * roi,
	                   means + 1);
	HIP_CHECK(hipGetLastError());

	HostFFT fx, fy;
	fx.build(W2);
	fy.build(H2);
	const FFTPlan px = fx.plan(), py = fy.plan();
	float2* A = d_a.as<float2>();  // the frame
	float2* B = d_b.as<float2>();  // the sample
	float2* C = d_c.as<float2>();  // transposed copies
	auto rows = [&](float2* data, int nrows, const FFTPlan& P, bool inverse) {
		hipLaunchKernelGGL(fft_rows_kernel, dim3(nrows), dim3(kThreads), (size_t)P.n * sizeof(float2), nullptr, data, P,
		                   inverse ? P.tw_inv : P.tw_fwd, inverse ? -1.0f : 1.0f);
		HIP_CHECK(hipGetLastError());
	};
	auto transpose = [&](const float2* in, float2* out, int W, int H) {
		hipLaunchKernelGGL(transpose_kernel, dim3((W + 31) / 32, (H + 31) / 32), dim3(kThreads), 0, nullptr, in, out,
		                   W, H);
		HIP_CHECK(hipGetLastError());
	};
	hipLaunchKernelGGL(fill_frame_kernel, dim3(grid1(N2)), dim3(kThreads), 0, nullptr, d_img.as<float>(), w, h, means,
	                   A, W2, H2);
	hipLaunchKernelGGL(fill_sample_kernel, dim3(grid1(N2)), dim3(kThreads), 0, nullptr, d_smp.as<float>(), roi,
	                   means + 1, B, W2, H2);
	HIP_CHECK(hipGetLastError());
	// forward 2D transforms: rows, then columns (as rows of the transposed arrays)
	rows(A, H2, px, false);
	rows(B, H2, px, false);
	transpose(A, C, W2, H2);  // C: frame, [W2][H2]
	transpose(B, A, W2, H2);  // A: sample, [W2][H2]
	rows(C, W2, py, false);
	rows(A, W2, py, false);
	hipLaunchKernelGGL(product_kernel, dim3(grid1(N2)), dim3(kThreads), 0, nullptr, C, (const float2*)A, N2);
	HIP_CHECK(hipGetLastError());
	// inverse 2D transform: rows first, then columns
	transpose(C, B, H2, W2);  // B: [H2][W2]
	rows(B, H2, px, true);
	transpose(B, C, W2, H2);  // C: [W2][H2]
	rows(C, W2, py, true);
	HIP_CHECK(hipMemset(d_max.p, 0, 4));
	hipLaunchKernelGGL(real_map_kernel, dim3((H2 + 31) / 32, (W2 + 31) / 32), dim3(kThreads), 0, nullptr,
	                   (const float2*)C, d_map.as<float>(), W2, H2, d_max.as<unsigned>());
	HIP_CHECK(hipGetLastError());
	std::vector<float> map(N2);
	unsigned maxbits = 0;
	HIP_CHECK(hipMemcpy(map.data(), d_map.p, N2 * 4, hipMemcpyDeviceToHost));
	HIP_CHECK(hipMemcpy(&maxbits, d_max.p, 4, hipMemcpyDeviceToHost));
	float maxVal;
	memcpy(&maxVal, &maxbits, 4);

	// the raster scan with SearchArea (cputrack/BeadFinder.cpp:81-103, 203-209)
	const float thr = maxVal * similarity;
	const int dist = (int)(0.5f * img_distance * roi);  // MinPixelDistance, passed as int
	std::vector<int2> pts;
	for (int y = 0; y < h; y++)
		for (int x = 0; x < w; x++) {
			if (!(map[(size_t)y * W2 + x] > thr)) continue;
			int bx = x, by = y;
			float best = map[(size_t)y * W2 + x];
			const int y0 = std::max(0, y - dist), x0 = std::max(0, x - dist);
			const int x1 = std::min(W2 - 1, x + dist), y1 = std::min(H2 - 1, y + dist);
			const int d2 = dist * dist;
			for (int yy = y0; yy <= y1; yy++)
				for (int xx = x0; xx <= x1; xx++) {
					const int dx = xx - x, dy = yy - y;
					if (dx * dx + dy * dy <= d2) {
						float& v = map[(size_t)yy * W2 + xx];
						if (v > best) {
							bx = xx;
							by = yy;
							best = v;
						}
						v = 0.0f;
					}
				}
			pts.push_back(make_int2(bx, by));
		}

	std::vector<ROIPosition> result;
	if (pts.empty()) return result;
	std::vector<int2> moved(pts.size());
	d_pts.ensure(pts.size() * sizeof(int2));
	d_out.ensure(pts.size() * sizeof(int2));
	HIP_CHECK(hipMemcpy(d_pts.p, pts.data(), pts.size() * sizeof(int2), hipMemcpyHostToDevice));
	const size_t lds = (size_t)roi * roi * 4;
	if (lds > 64 * 1024) HIP_CHECK(hipFuncSetAttribute((const void*)recenter_kernel,
	                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	hipLaunchKernelGGL(recenter_kernel, dim3((unsigned)pts.size()), dim3(kThreads), lds, nullptr, d_img.as<float>(), w,
	                   h, roi, d_pts.as<int2>(), d_out.as<int2>());
	HIP_CHECK(hipGetLastError());
	HIP_CHECK(hipMemcpy(moved.data(), d_out.p, moved.size() * sizeof(int2), hipMemcpyDeviceToHost));

	// the list filter (cputrack/BeadFinder.cpp:147-160): a bead closer than the minimum
	// distance to one already listed removes that one and is not listed itself
	const float m = 0.5f * img_distance * roi;
	std::list<int2> kept;
	for (const int2& bp : moved) {
		if (bp.x == INT_MIN) continue;
		if (bp.x < 0 || bp.y < 0 || bp.x > w - 1 - roi || bp.y > h - 1 - roi) continue;
		bool accepted = true;
		for (auto c = kept.begin(); c != kept.end(); ++c) {
			const int dx = c->x - bp.x, dy = c->y - bp.y;
			if ((float)(dx * dx + dy * dy) < m * m) {
				kept.erase(c);
				accepted = false;
				break;
			}
		}
		if (accepted) kept.push_back(bp);
	}
	for (const int2& p : kept) result.push_back(ROIPosition{p.x, p.y});
	return result;
}

}  // namespace qtrk
