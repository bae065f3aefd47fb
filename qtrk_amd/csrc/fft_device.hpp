// fft_device.hpp -- kissfft's complex arithmetic and power-of-two butterflies
// for device code, shared by the localization kernel and the bead finder.
// Files that include it are built with -ffp-contract=off, so every product and
// sum rounds on its own as in the reference's kissfft (cputrack/kissfft.h).
#pragma once

#include <hip/hip_runtime.h>

namespace qtrk {

// C_MUL, C_ADD, C_SUB (cputrack/kissfft.h:112-131); std::complex<float>'s
// product rounds the same way
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
	return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// One radix-4 or radix-2 butterfly of kissfft's decimation in time
// (kf_bfly4, kf_bfly2: cputrack/kissfft.h:112-152): G is the group, k the
// butterfly, m the group's sub-length, k1 = k * fstride; nii is 1 for the
// forward transform and -1 for the inverse (the sign kf_bfly4 gives the
// rotated difference).
__device__ __forceinline__ void bfly_pow2(float2* G, int r, int k, int m, int k1, const float2* tw, float nii)
{
	if (r == 4) {
		const float2 s0 = cmul(G[k + m], tw[k1]);
		const float2 s1 = cmul(G[k + 2 * m], tw[k1 << 1]);
		const float2 s2 = cmul(G[k + 3 * m], tw[k1 + (k1 << 1)]);
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
	} else {
		const float2 tt = cmul(G[m + k], tw[k1]);
		const float2 f0 = G[k];
		G[m + k] = csub(f0, tt);
		G[k] = cadd(f0, tt);
	}
}

}  // namespace qtrk
