/*
 * qmath.h -- single-precision exp and erf built from float additions,
 * multiplications, comparisons and exponent-field construction only.
 *
 * The same source is compiled for the GPU (the 2D Gaussian MLE in localize.hip,
 * with QM_FN defined as a device function qualifier) and for the host, so a
 * computation that calls them rounds identically on both whenever both are
 * compiled without contraction (-ffp-contract=off) and with IEEE float
 * arithmetic.  The C library's expf/erff and the GPU device library's differ
 * from each other in the last place; these do not differ between the two.
 *
 * Methods:
 *  - exp: Cody-Waite argument reduction (W. J. Cody, W. Waite, "Software Manual
 *    for the Elementary Functions", 1980): x = k ln2 + r with k = nearest(x log2 e),
 *    r = (x - k LN2_HI) - k LN2_LO, where LN2_HI holds 16 significant bits so
 *    k LN2_HI is exact; exp(r) = 1 + (r + r^2 E(r)) with E a minimax polynomial
 *    on |r| <= ln2/2; the result scaled by 2^k through the exponent field (in two
 *    steps below 2^-126, so a subnormal result is rounded once).
 *  - erf: a + a S(a^2) for |x| < 1 (erf(a)/a is even and smooth); for 1 <= |x| < 4
 *    1 - exp(-a^2) R(a), where R(a) = erfc(a) exp(a^2) is smooth and slowly varying
 *    (two minimax polynomials, on [1, 2) and [2, 4)); +-1 from |x| >= 4, where
 *    erfc(x) < 2^-26 and 1 - erfc rounds to 1.
 *
 * The coefficients are derived by qmath_fit.py (minimax fits in double precision,
 * rounded to float); the block below is its output.  Measured on every float with
 * |x| <= 104 (exp) and |x| <= 5 (erf): qm_expf within 0.951 ulp and qm_erff within
 * 1.501 ulp of the exact value (1 and 2 ulp from glibc's expf and erff).
 */
#ifndef QTRK_QMATH_H
#define QTRK_QMATH_H

#ifndef QM_FN
#define QM_FN static inline
#endif

/* ---- generated by qmath_fit.py ---- */
/* E: degree 5, fit error 2.86e-09 */
#define QM_E 0x1p-1f, 0x1.555556p-3f, 0x1.5554e6p-5f, 0x1.11114ep-7f, 0x1.6d48a6p-10f, 0x1.a07452p-13f
/* S: degree 6, fit error 1.19e-09 */
#define QM_S 0x1.06eba8p-3f, -0x1.81273ep-2f, 0x1.ce2d2p-4f, -0x1.b7fbb2p-6f, 0x1.5419e8p-8f, -0x1.a485bep-11f, 0x1.4ae4b2p-14f
/* R1: degree 8, fit error 1.80e-09 */
#define QM_R1 0x1.494dbp-2f, -0x1.4f1986p-3f, 0x1.37ea28p-4f, -0x1.0dc648p-5f, 0x1.b6598p-7f, -0x1.50b8c4p-8f, 0x1.ee0748p-10f, -0x1.71351ep-11f, 0x1.ec328ap-13f
/* R2: degree 9, fit error 3.55e-09 */
#define QM_R2 0x1.6e9828p-3f, -0x1.bd6ae4p-5f, 0x1.043fcp-6f, -0x1.25905cp-8f, 0x1.40adcap-10f, -0x1.53e89ep-12f, 0x1.5b7efep-14f, -0x1.5debaap-16f, 0x1.8e238ep-18f, -0x1.77fe7p-20f
#define QM_LN2_HI 0x1.62e4p-1f
#define QM_LN2_LO 0x1.7f7d1cp-20f
#define QM_LOG2E 0x1.715476p+0f
/* ---- end of generated block ---- */

/* Polynomial c[0] + c[1] t + ... + c[n-1] t^(n-1), Horner's rule from the top. */
QM_FN float qm_horner(const float* c, int n, float t)
{
	float acc = c[n - 1];
	for (int i = n - 2; i >= 0; i--) acc = acc * t + c[i];
	return acc;
}

/* 2^e for -126 <= e <= 127, from the exponent field. */
QM_FN float qm_pow2i(int e)
{
	const unsigned bits = (unsigned)(e + 127) << 23;
	float f;
	__builtin_memcpy(&f, &bits, sizeof f);
	return f;
}

QM_FN float qm_expf(float x)
{
	if (x != x) return x + x;
	if (x > 0x1.62e42ep+6f) return qm_pow2i(127) * 2.0f; /* at or above ln(FLT_MAX): +inf */
	if (x < -0x1.9fe368p+6f) return 0.0f;                /* below ln(2^-150): rounds to 0 */
	const float E[] = { QM_E };
	const float t = x * QM_LOG2E;
	const float kf = (t + 0x1.8p23f) - 0x1.8p23f; /* nearest integer (|t| < 2^22) */
	const int k = (int)kf;
	const float r = (x - kf * QM_LN2_HI) - kf * QM_LN2_LO;
	const float p = 1.0f + (r + (r * r) * qm_horner(E, 6, r));
	if (k < -126) return (p * qm_pow2i(k + 64)) * 0x1p-64f;
	if (k > 127) return (p * qm_pow2i(k - 1)) * 2.0f;
	return p * qm_pow2i(k);
}

QM_FN float qm_erff(float x)
{
	if (x != x) return x + x;
	const float a = x < 0.0f ? -x : x;
	float y;
	if (a < 1.0f) {
		const float S[] = { QM_S };
		y = a + a * qm_horner(S, 7, a * a);
	} else if (a < 4.0f) {
		const float R1[] = { QM_R1 };
		const float R2[] = { QM_R2 };
		const float R = a < 2.0f ? qm_horner(R1, 9, a - 1.5f) : qm_horner(R2, 10, a - 3.0f);
		y = 1.0f - qm_expf(-(a * a)) * R;
	} else {
		y = 1.0f;
	}
	return x < 0.0f ? -y : y;
}

#endif /* QTRK_QMATH_H */
