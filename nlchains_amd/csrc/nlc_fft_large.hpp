// Two-level FFT of power-of-two lengths N = R M beyond one workgroup's LDS
// (R = 4..16 blocks of M = 4096 or 8192 points), shared by the energy pass and
// the DNLS step.  With W_N = e^{-2 pi i/N}, x[j M + m] (j < R, m < M):
//   column pass  y_r[m] = W_N^{m r} sum_j x[j M + m] W_R^{j r}          (HBM)
//   row pass     Z[r + R k] = sum_m y_r[m] W_M^{m k}                 (LDS FFT)
// and the unnormalised inverse runs the same passes backwards with conjugate
// twiddles.  The spectrum is kept in the order [r][k] (Z_{r + R k} at r M + k).
// Host side: large_fft_split gives (R, M) for N; device side: the R-point DFTs
// of the column pass, written out from their definition (R <= 16 values per
// thread in registers).
#ifndef NLC_FFT_LARGE_HPP
#define NLC_FFT_LARGE_HPP

#include <cstdint>

namespace nlc {

constexpr uint32_t kLargeFftMax = 131072; // 16 x 8192

// (R, M) for a power-of-two N in (8192, kLargeFftMax]: M = 4096 up to 65536
// (64 KiB of LDS per row, two rows per CU), 8192 above.  False otherwise.
inline bool large_fft_split(uint32_t n, uint32_t &R, uint32_t &M) {
	if (n <= 8192 || n > kLargeFftMax || (n & (n - 1))) return false;
	M = n <= 65536 ? 4096 : 8192;
	R = n / M;
	return true;
}

#ifdef __HIPCC__
// The R-th roots of unity e^{-2 pi i t/R} = tw[t N/R] (tw: the N-point table
// e^{-2 pi i t/N}), read once into registers.
template <uint32_t R>
__device__ __forceinline__ void small_roots(double2 root[R], const double2 *__restrict__ tw, uint32_t n) {
#pragma unroll
	for (uint32_t t = 0; t < R; ++t) root[t] = tw[t * (n / R)];
}

// out_r = sum_j in_j root^{j r} for r < R (INVERSE: conjugate roots), direct
// sums over a compile-time R: every root index is a constant.
template <uint32_t R, bool INVERSE>
__device__ __forceinline__ void small_dft(const double2 in[R], double2 out[R], const double2 root[R]) {
#pragma unroll
	for (uint32_t r = 0; r < R; ++r) {
		double sr = 0.0, si = 0.0;
#pragma unroll
		for (uint32_t j = 0; j < R; ++j) {
			const double2 w = root[(j * r) % R];
			const double wi = INVERSE ? -w.y : w.y;
			sr = fma(in[j].x, w.x, fma(-in[j].y, wi, sr));
			si = fma(in[j].x, wi, fma(in[j].y, w.x, si));
		}
		out[r] = make_double2(sr, si);
	}
}
#endif

} // namespace nlc

#endif
