// Mixed-radix Stockham FFT in LDS for lengths n = 2^a 3^b 5^c 7^d that are not
// powers of two (the passes of mixed_radix_plan: radix 4 first, then 2, 3, 5,
// 7), shared by the DNLS step and the energy pass.
//
// One pass of radix P, with ns = the product of the radices of the earlier
// passes: butterfly j < n/P reads x[j + r n/P] (r < P), multiplies input r by
// e^{-2 pi i r k/(ns P)} (k = j mod ns) and writes its P-point DFT to
// x[(j/ns) ns P + k + r ns] (Stockham's autosort: natural order in and out,
// no bit reversal).  Every index stays below n because ns P divides n.  The
// P-point DFT is the butterfly for P = 2 and 4 and the direct sum over the
// P-th roots of unity tw[t n/P] for P = 3, 5, 7.
#ifndef NLC_FFT_MIXED_HPP
#define NLC_FFT_MIXED_HPP

#include "nlc_internal.hpp"

namespace nlc {
namespace mixed {

__device__ __forceinline__ double2 mul(double2 a, double2 b) {
	return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// One pass of radix P over n points in buf (tw[t] = e^{-2 pi i t/n}) by the
// NT threads of the block (WHOLE), or by threads tid < nthr when several
// chains share a block of NT threads (then n nthr-chunks stay within MAXN: at
// most kPer butterflies per thread).
template <uint32_t MAXN, uint32_t NT, uint32_t P, bool WHOLE>
__device__ void stockham_pass(double2 *buf, uint32_t n, uint32_t ns, const double2 *__restrict__ tw, uint32_t tid,
                              uint32_t nthr_rt) {
	constexpr uint32_t kPer = (MAXN / P + NT - 1) / NT; // butterflies per thread
	const uint32_t nthr = WHOLE ? NT : nthr_rt;
	const uint32_t m = n / P, span = n / (ns * P);
	double2 v[kPer][P];
#pragma unroll
	for (uint32_t q = 0; q < kPer; ++q) {
		const uint32_t j = tid + q * nthr;
		if (j < m)
#pragma unroll
			for (uint32_t r = 0; r < P; ++r) v[q][r] = buf[j + r * m];
	}
	__syncthreads();
	double2 root[P];
	if constexpr (P != 2 && P != 4) small_roots<P>(root, tw, n);
#pragma unroll
	for (uint32_t q = 0; q < kPer; ++q) {
		const uint32_t j = tid + q * nthr;
		if (j < m) {
			const uint32_t k = j % ns;
			double2 a[P], y[P];
			a[0] = v[q][0];
#pragma unroll
			for (uint32_t r = 1; r < P; ++r) a[r] = mul(v[q][r], tw[r * k * span]); // r k span < n
			if constexpr (P == 2) {
				y[0] = make_double2(a[0].x + a[1].x, a[0].y + a[1].y);
				y[1] = make_double2(a[0].x - a[1].x, a[0].y - a[1].y);
			} else if constexpr (P == 4) {
				// forward 4-point DFT: y_r = sum_s a_s (-i)^{rs}
				const double2 s02 = make_double2(a[0].x + a[2].x, a[0].y + a[2].y);
				const double2 d02 = make_double2(a[0].x - a[2].x, a[0].y - a[2].y);
				const double2 s13 = make_double2(a[1].x + a[3].x, a[1].y + a[3].y);
				const double2 d13 = make_double2(a[1].x - a[3].x, a[1].y - a[3].y);
				y[0] = make_double2(s02.x + s13.x, s02.y + s13.y);
				y[2] = make_double2(s02.x - s13.x, s02.y - s13.y);
				y[1] = make_double2(d02.x + d13.y, d02.y - d13.x); // d02 - i d13
				y[3] = make_double2(d02.x - d13.y, d02.y + d13.x); // d02 + i d13
			} else {
				small_dft<P, false>(a, y, root);
			}
			const uint32_t o = (j / ns) * ns * P + k;
#pragma unroll
			for (uint32_t r = 0; r < P; ++r) buf[o + r * ns] = y[r];
		}
	}
	__syncthreads();
}

// Forward n-point FFT in LDS through the passes of `plan`; every thread of the
// block calls it (it contains barriers).
template <uint32_t MAXN, uint32_t NT, bool WHOLE = true>
__device__ void lds_fft(double2 *buf, uint32_t n, const double2 *__restrict__ tw, const MixedPlan &plan,
                        uint32_t tid = 0, uint32_t nthr = NT) {
	if (WHOLE) tid = threadIdx.x;
	uint32_t ns = 1;
	for (uint32_t i = 0; i < plan.passes; ++i) {
		switch (plan.radix[i]) {
		case 2: stockham_pass<MAXN, NT, 2, WHOLE>(buf, n, ns, tw, tid, nthr); break;
		case 3: stockham_pass<MAXN, NT, 3, WHOLE>(buf, n, ns, tw, tid, nthr); break;
		case 4: stockham_pass<MAXN, NT, 4, WHOLE>(buf, n, ns, tw, tid, nthr); break;
		case 5: stockham_pass<MAXN, NT, 5, WHOLE>(buf, n, ns, tw, tid, nthr); break;
		default: stockham_pass<MAXN, NT, 7, WHOLE>(buf, n, ns, tw, tid, nthr); break;
		}
		ns *= plan.radix[i];
	}
}

} // namespace mixed
} // namespace nlc

#endif
