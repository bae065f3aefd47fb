// pipeline_kernels.hip -- glue kernels of the multi-tracker pipeline.
#include "localize.hpp"

namespace qtrk {

// theta0 = (x, y, z, -1, -1): the localization's x/y and the matched z, clamped
// into the ROI and the model LUT (NaN -> the middle); I = -1 asks the fit for
// its least-squares I/bg start.
__global__ void assemble_start_kernel(const DevResult* res, const float* z, float* th, int n, float wmax, float hmax,
                                      float zmax)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const float x = res[i].x, y = res[i].y, zz = z[i];
	float* t = th + (size_t)i * 5;
	t[0] = x == x ? fminf(fmaxf(x, 0.0f), wmax) : 0.5f * wmax;
	t[1] = y == y ? fminf(fmaxf(y, 0.0f), hmax) : 0.5f * hmax;
	t[2] = zz == zz ? fminf(fmaxf(zz, 0.0f), zmax) : 0.5f * zmax;
	t[3] = -1.0f;
	t[4] = -1.0f;
}

hipError_t launch_assemble_start(const DevResult* res, const float* z, float* theta0, int n, int width, int height,
                                 int planes, hipStream_t stream)
{
	if (n <= 0) return hipSuccess;
	assemble_start_kernel<<<(n + 255) / 256, 256, 0, stream>>>(res, z, theta0, n, (float)(width - 1),
	                                                            (float)(height - 1), (float)(planes - 1));
	return hipGetLastError();
}

}  // namespace qtrk
