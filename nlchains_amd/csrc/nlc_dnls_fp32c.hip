// DNLS split-step Yoshida-6 with FP32 fields and compensated (double-single)
// accumulation, 4096-site chains: the FP32-compensated precision of the DNLS
// step.  Same step sequence as the FP64 kernel (DNLS.cpp:159-175; kernels
// DNLS.cu:11-25; CPU form DNLS/cpu-implementation/DNLS.cpp:110-128); the
// energy, norm and entropy reductions stay FP64 (the state in HBM is FP64).
//
// The state of a site is held as psi = hi + lo, two FP32 complex values with
// |lo| <= ulp(hi)/2 (48 significant bits).  Every substep is written as an
// increment of the state, computed in FP32 and added to (hi, lo) with a
// compensated sum:
//   nonlinear: d_N = hi (e^{-i theta |hi|^2} - 1),
//   linear:    d_L = IFFT(M_k FFT(w)),  w = psi + d_N rounded to FP32,
//              M_k[j] = (e^{-i dt d_k omega_j} - 1)/N,
//   psi <- psi + d_N + d_L.
// The FP32 rounding errors of an increment are relative to the increment, not
// to psi: |e^{-i theta|psi|^2} - 1| ~ theta |psi|^2 and |M_k| <= 2 dt |d_k|
// omega_max / N are a few hundredths here, so one substep adds about 1e-9 of
// |psi| of rounding noise instead of FP32's 6e-8 times the FFT's log2 N.
// Quantities that enter every step the same way are kept to about 48 bits so
// that their rounding does not act as a systematic bias: the multipliers M_k
// and the angles theta_k = beta dt c_k are double-single constants (hi + lo);
// the FFT twiddles are FP32 (their error only scales the increment).
//
// The sum psi + d is Fast2Sum(hi, lo + d): s = hi + t, lo = t - (s - hi) with
// t = lo + d.  It is exact when |hi| >= |t|; where an increment exceeds the
// value (near zeros of psi) its error is of the size of the increment's own
// FP32 rounding.  4 FP32 operations per real component.
//
// Work per point and substep: the two FP32 FFTs (as many instructions as the
// FP64 ones: without packed FP32 an FP32 instruction issues at the FP64
// rate), the double-single multiplier (8 instead of 4), the nonlinear
// increment (FP32 polynomials of lower degree than FP64's) and the
// compensated sum (8): about 10 % more VALU instructions than the FP64 step,
// against half the LDS bytes, half the table traffic and fewer VGPRs
// (more wavefronts per SIMD).
#include "nlc_internal.hpp"

#include <cmath>
#include <vector>

namespace nlc {
namespace {

#include "nlc_dnls_common.inl"
#include "nlc_fft4096.inl"

// Angles beta dt c_k as double-single (hi + lo): k < 7 the substep angles,
// [7] the closing c_7 rotation, [8] the merged 2 c_0 of later steps.
struct DnlsArgsC {
	float th_hi[9], th_lo[9];
	double theta7;
};

constexpr float kExpiSmallF = 0.19634954f; // pi/16

// (cos a - 1, sin a) for |a| <= pi/16 in FP32: Taylor polynomials in x = a^2,
// sin to a^7 (relative truncation a^8/9! < 7e-12), cos - 1 to a^6 (relative
// truncation 2 a^6/8! < 3e-9), both below FP32's rounding.
__device__ __forceinline__ float2 expim1_small(float a) {
	const float x = a * a;
	float s = -1.98412698e-4f;      // -1/7!
	s = fmaf(s, x, 8.33333333e-3f); // 1/5!
	s = fmaf(s, x, -1.66666667e-1f); // -1/3!
	float c = -1.38888889e-3f;      // -1/6!
	c = fmaf(c, x, 4.16666667e-2f); // 1/4!
	c = fmaf(c, x, -0.5f);          // -1/2!
	return make_float2(c * x, fmaf(a * x, s, a));
}

// The nonlinear increment of all 16 values of a thread, folded into lo; w gets
// the FP32 input of the forward transform, psi + d_N.  One wave-uniform branch
// as in the FP64 kernel: angles above pi/16 here (rare: |psi|^2 far above its mean)
// take the FP64 reduced form for the whole wavefront.
__device__ __forceinline__ void nonlinear_increment(const float2 (&hi)[16], float2 (&lo)[16], float2 (&w)[16],
                                                    float th_hi, float th_lo) {
	float a[16];
	bool small = true;
#pragma unroll
	for (int j = 0; j < 16; ++j) {
		const float r2 = fmaf(hi[j].x, hi[j].x, hi[j].y * hi[j].y);
		a[j] = fmaf(-th_lo, r2, -th_hi * r2);
		small = small && fabsf(a[j]) <= kExpiSmallF;
	}
	if (__all(small)) {
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			const float2 d = cmul(hi[j], expim1_small(a[j]));
			lo[j] = cadd(lo[j], d);
			w[j] = cadd(hi[j], lo[j]);
		}
	} else {
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			const double2 e = expi_reduced(double(a[j]));
			const float2 d = cmul(hi[j], make_float2(float(e.x - 1.0), float(e.y)));
			lo[j] = cadd(lo[j], d);
			w[j] = cadd(hi[j], lo[j]);
		}
	}
}

// v * (mh + ml) with the multiplier as a double-single pair {mh.x, mh.y, ml.x, ml.y}
__device__ __forceinline__ float2 cmul_ds(float2 v, float4 m) {
	return make_float2(fmaf(v.x, m.x, fmaf(-v.y, m.y, fmaf(v.x, m.z, -v.y * m.w))),
	                   fmaf(v.x, m.y, fmaf(v.y, m.x, fmaf(v.x, m.w, v.y * m.z))));
}

// psi <- hi + lo + d as Fast2Sum(hi, lo + d)
__device__ __forceinline__ void accumulate(float &hi, float &lo, float d) {
	const float t = lo + d;
	const float s = hi + t;
	lo = t - (s - hi);
	hi = s;
}

__global__ void __launch_bounds__(kT, 3)
k_dnls_4096_c(double2 *__restrict__ state, uint32_t copies, uint32_t steps, DnlsArgsC a,
              const float2 *__restrict__ tw, const float4 *__restrict__ mult) {
	__shared__ Lds4352<float2> L;
	const uint32_t chain = blockIdx.x, t = threadIdx.x;
	double2 *psi = state + size_t(chain) * 4096;
	// psi[t + 256 j] = hi[pos(j)] + lo[pos(j)]
	float2 hi[16], lo[16];
#pragma unroll
	for (int j = 0; j < 16; ++j) {
		const double2 z = psi[t + 256 * j];
		const float hx = float(z.x), hy = float(z.y);
		hi[pos(j)] = make_float2(hx, hy);
		lo[pos(j)] = make_float2(float(z.x - double(hx)), float(z.y - double(hy)));
	}
	for (uint32_t s = 0; s < steps; ++s) {
#pragma unroll 1
		for (int k = 0; k < 7; ++k) {
			const int ki = (k == 0 && s) ? 8 : k;
			const float2 *twk = tw;
			asm volatile("" : "+s"(twk));
			float2 v[16];
			nonlinear_increment(hi, lo, v, a.th_hi[ki], a.th_lo[ki]);
			fft4096_fwd(v, L, twk);
			// the increment multiplier, stored per thread in the digit-reversed order
			const float4 *m = mult + size_t(k < 4 ? k : 6 - k) * 4096 + t;
#pragma unroll
			for (int j = 0; j < 16; ++j) v[pos(j)] = cmul_ds(v[pos(j)], m[256 * j]);
			fft4096_inv(v, L, twk);
#pragma unroll
			for (int j = 0; j < 16; ++j) {
				accumulate(hi[j].x, lo[j].x, v[j].x);
				accumulate(hi[j].y, lo[j].y, v[j].y);
			}
		}
	}
	// the closing c_7 rotation in FP64 on hi + lo
#pragma unroll
	for (int j = 0; j < 16; ++j) {
		const double2 z = make_double2(double(hi[pos(j)].x) + double(lo[pos(j)].x),
		                               double(hi[pos(j)].y) + double(lo[pos(j)].y));
		psi[t + 256 * j] = nonlinear(z, a.theta7);
	}
}

} // namespace

void dnls_fp32c_tables(uint32_t n, double dt, float2 *tw, float4 *mult) {
	// twiddles e^{-2 pi i j/n}, rounded to FP32
	for (uint32_t j = 0; j < n; ++j) {
		const double ang = 2.0 * M_PI * double(j) / double(n);
		tw[j] = make_float2(float(std::cos(ang)), float(-std::sin(ang)));
	}
	// increment multipliers (e^{-i dt d_k omega_j} - 1)/n of the 4 distinct
	// substeps (d_k = d_{6-k}), cos - 1 as -2 sin^2 to keep its relative
	// accuracy for small angles, split hi + lo, in the radix-16 kernel's
	// digit-reversed order (entry [k][256 j + t] is mode (t >> 4) + 16 (t & 15) + 256 j)
	std::vector<double> omega(n);
	for (uint32_t j = 0; j < n; ++j) {
		const double s2 = 2 * std::sin(j * M_PI / n);
		omega[j] = s2 * s2; // DNLS/dispersion.cpp:10-17
	}
	for (int k = 0; k < 4; ++k)
		for (uint32_t e = 0; e < n; ++e) {
			const uint32_t jj = e >> 8, t = e & 255;
			const uint32_t mode = (t >> 4) + 16 * (t & 15) + 256 * jj;
			const double ang = ((-dt) * kYoshidaD[k]) * omega[mode];
			const double sh = std::sin(0.5 * ang);
			const double re = (-2.0 * sh * sh) / n, im = std::sin(ang) / n;
			const float rh = float(re), ih = float(im);
			mult[size_t(k) * n + e] = make_float4(rh, ih, float(re - double(rh)), float(im - double(ih)));
		}
}

hipError_t launch_dnls_fp32c(hipStream_t st, double2 *state, uint32_t copies, uint32_t steps, const double *theta9,
                             const float2 *tw, const float4 *mult) {
	DnlsArgsC a;
	for (int i = 0; i < 9; ++i) {
		a.th_hi[i] = float(theta9[i]);
		a.th_lo[i] = float(theta9[i] - double(a.th_hi[i]));
	}
	a.theta7 = theta9[7];
	hipLaunchKernelGGL(k_dnls_4096_c, dim3(copies), dim3(kT), 0, st, state, copies, steps, a, tw, mult);
	return hipGetLastError();
}

} // namespace nlc
