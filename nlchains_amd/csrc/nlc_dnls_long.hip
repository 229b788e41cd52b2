// DNLS split step and energy-pass transforms for chains beyond one workgroup's
// LDS: powers of two above 8192 (up to 131072) through the two-level FFT of
// nlc_fft_large.hpp, other lengths from 4097 to 65536 through Bluestein's
// convolution on the same passes.  The chain and its spectrum live in HBM.
//
// Reference behaviour followed: DNLS.cpp:159-175 (as nlc_dnls.hip), energies
// DNLS.cpp:142-144 and DNKG_FPUT_Toda.cpp:113-123.
#include "nlc_internal.hpp"
#include "nlc_fft_large.hpp"

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

namespace nlc {
namespace {

#include "nlc_dnls_common.inl"

// Chains above 8192 sites (powers of two up to 131072, n = R M): the two-level
// FFT of nlc_fft_large.hpp with the chain and its spectrum in HBM (scratch:
// copies x n values, the spectrum in the order [r][k]).  A substep is two
// launches: the row kernel (M-point FFT in LDS, Fourier multiplier, inverse
// FFT, conjugate twiddle) and the column kernel, which finishes the inverse
// transform of one substep (R-point DFT), applies the nonlinear phase of the
// next and starts its forward transform (R-point DFT and twiddle).
// MODE 0: the first substep, reads the state; 1: between substeps; 2: after
// the last one, applies theta_7 and writes the state.
template <uint32_t R, int MODE>
__global__ void __launch_bounds__(kT)
k_dnls_col(double2 *__restrict__ state, double2 *__restrict__ scratch, uint32_t n, uint32_t M, uint32_t copies,
           double th, const double2 *__restrict__ tw) {
	const uint64_t t = uint64_t(blockIdx.x) * kT + threadIdx.x;
	if (t >= uint64_t(copies) * M) return;
	const uint32_t c = uint32_t(t / M), m = uint32_t(t % M);
	double2 root[R], x[R], y[R];
	small_roots<R>(root, tw, n);
	double2 *psi = state + size_t(c) * n + m, *sc = scratch + size_t(c) * n + m;
	if constexpr (MODE == 0) {
#pragma unroll
		for (uint32_t j = 0; j < R; ++j) x[j] = psi[size_t(j) * M];
	} else {
#pragma unroll
		for (uint32_t r = 0; r < R; ++r) y[r] = sc[size_t(r) * M];
		small_dft<R, true>(y, x, root);
	}
#pragma unroll
	for (uint32_t j = 0; j < R; ++j) x[j] = nonlinear(x[j], th);
	if constexpr (MODE == 2) {
#pragma unroll
		for (uint32_t j = 0; j < R; ++j) psi[size_t(j) * M] = x[j];
	} else {
		small_dft<R, false>(x, y, root);
#pragma unroll
		for (uint32_t r = 0; r < R; ++r) sc[size_t(r) * M] = cmul(y[r], tw[m * r]);
	}
}

// One row (chain c, block r) per workgroup: Z = FFT_M(y_r), U = Z T_k (the
// multipliers in the same [r][k] order), v_r[m] = IFFT_M(U)[m] W_N^{-m r}
// (unnormalised; T carries the 1/n), the inverse as conj(FFT(conj(.))).
template <uint32_t M>
__global__ void __launch_bounds__(kT)
k_dnls_row(double2 *__restrict__ scratch, uint32_t R, const double2 *__restrict__ tw, const double2 *__restrict__ tk) {
	extern __shared__ double2 buf[];
	const uint32_t r = blockIdx.x % R;
	double2 *row = scratch + size_t(blockIdx.x) * M;
	for (uint32_t x = threadIdx.x; x < M; x += kT) buf[x] = row[x];
	__syncthreads();
	lds_fft_fwd<M>(buf, M, tw, R);
	const double2 *t = tk + size_t(r) * M;
	for (uint32_t k = threadIdx.x; k < M; k += kT) buf[k] = cconj(cmul(buf[k], t[k]));
	__syncthreads();
	lds_fft_fwd<M>(buf, M, tw, R);
	for (uint32_t m = threadIdx.x; m < M; m += kT) row[m] = cconj(cmul(buf[m], tw[m * r]));
}

template <uint32_t R, uint32_t M>
hipError_t launch_dnls_two_level(hipStream_t st, double2 *state, double2 *scratch, uint32_t n, uint32_t copies,
                                 uint32_t steps, const DnlsArgs &a, const double2 *tw, const double2 *table) {
	if (M * sizeof(double2) > 64 * 1024) {
		hipError_t e = hipFuncSetAttribute((const void *) k_dnls_row<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   int(M * sizeof(double2)));
		if (e != hipSuccess) return e;
	}
	const dim3 cols(uint32_t((uint64_t(copies) * M + kT - 1) / kT)), rows(copies * R), block(kT);
	const size_t lds = M * sizeof(double2);
	for (uint32_t s = 0; s < steps; ++s)
		for (int k = 0; k < 7; ++k) {
			const double th = (k == 0 && s) ? a.theta[8] : a.theta[k];
			if (s == 0 && k == 0)
				hipLaunchKernelGGL((k_dnls_col<R, 0>), cols, block, 0, st, state, scratch, n, M, copies, th, tw);
			else
				hipLaunchKernelGGL((k_dnls_col<R, 1>), cols, block, 0, st, state, scratch, n, M, copies, th, tw);
			if constexpr (M == 4096) { // rows on the radix-16 register FFT
				hipError_t e = launch_dnls_rows4096(st, scratch, copies * R, R, table + size_t(7) * n, tw,
				                                    table + size_t(k) * n);
				if (e != hipSuccess) return e;
			} else {
				hipLaunchKernelGGL(k_dnls_row<M>, rows, block, lds, st, scratch, R, tw, table + size_t(k) * n);
			}
		}
	hipLaunchKernelGGL((k_dnls_col<R, 2>), cols, block, 0, st, state, scratch, n, M, copies, a.theta[7], tw);
	return hipGetLastError();
}

// Lengths n that are not powers of two, 4096 < n <= 65536: Bluestein's
// convolution (as in k_dnls_bluestein) of length L = R M >= 2n - 1 through the
// two-level FFT, one L-entry buffer per chain (natural order between the
// passes, [r][k] order inside a transform; the column passes work in place on
// the R entries j M + m of their offset m).  conv(a) = IFFT_L(FFT_L(a) bhat)
// is k_blue_fwd, k_dnls_row with bhat as the multiplier, k_blue_inv.
// SRC of k_blue_fwd: 0 w psi from the state with the nonlinear phase, 1 the
// buffer with the nonlinear phase, 2 the buffer, 3 w z from the state (energy
// pass).  Entries at or beyond n are zero.
template <uint32_t R, int SRC>
__global__ void __launch_bounds__(kT)
k_blue_fwd(const double2 *__restrict__ state, double2 *__restrict__ buf, uint32_t n, uint32_t L, uint32_t M,
           uint32_t copies, double th, const double2 *__restrict__ twL, const double2 *__restrict__ chirp) {
	const uint64_t t = uint64_t(blockIdx.x) * kT + threadIdx.x;
	if (t >= uint64_t(copies) * M) return;
	const uint32_t c = uint32_t(t / M), m = uint32_t(t % M);
	double2 root[R], a[R], y[R];
	small_roots<R>(root, twL, L);
	double2 *b = buf + size_t(c) * L + m;
#pragma unroll
	for (uint32_t j = 0; j < R; ++j) {
		const uint32_t x = j * M + m;
		double2 v = make_double2(0.0, 0.0);
		if (x < n) {
			if constexpr (SRC == 0 || SRC == 3) v = cmul(state[size_t(c) * n + x], chirp[x]);
			else v = b[size_t(j) * M];
			if constexpr (SRC == 0 || SRC == 1) v = nonlinear(v, th);
		}
		a[j] = v;
	}
	small_dft<R, false>(a, y, root);
#pragma unroll
	for (uint32_t r = 0; r < R; ++r) b[size_t(r) * M] = cmul(y[r], twL[m * r]);
}

// DST of k_blue_inv, from C = conv(.) at x = j M + m: 0 C (energy pass); 1
// conj(C T_k) for x < n, else 0; 2 conj(C) for x < n, else 0; 3 the state:
// psi = conj(w C) with the nonlinear phase (after the last substep).
template <uint32_t R, int DST>
__global__ void __launch_bounds__(kT)
k_blue_inv(double2 *__restrict__ buf, double2 *__restrict__ state, uint32_t n, uint32_t L, uint32_t M,
           uint32_t copies, double th, const double2 *__restrict__ twL, const double2 *__restrict__ chirp,
           const double2 *__restrict__ tk) {
	const uint64_t t = uint64_t(blockIdx.x) * kT + threadIdx.x;
	if (t >= uint64_t(copies) * M) return;
	const uint32_t c = uint32_t(t / M), m = uint32_t(t % M);
	double2 root[R], v[R], x[R];
	small_roots<R>(root, twL, L);
	double2 *b = buf + size_t(c) * L + m;
#pragma unroll
	for (uint32_t r = 0; r < R; ++r) v[r] = b[size_t(r) * M];
	small_dft<R, true>(v, x, root);
#pragma unroll
	for (uint32_t j = 0; j < R; ++j) {
		const uint32_t at = j * M + m;
		const double2 zero = make_double2(0.0, 0.0);
		if constexpr (DST == 0) b[size_t(j) * M] = x[j];
		else if constexpr (DST == 1) b[size_t(j) * M] = at < n ? cconj(cmul(x[j], tk[at])) : zero;
		else if constexpr (DST == 2) b[size_t(j) * M] = at < n ? cconj(x[j]) : zero;
		else if (at < n) state[size_t(c) * n + at] = nonlinear(cconj(cmul(x[j], chirp[at])), th);
	}
}

__global__ void k_zero_rows(double *p, size_t count) {
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
		p[i] = 0.0;
}

// Energy pass from C = conv(w z): X_k = w_k C_k; partial[g][k] += e_k over the
// chunk's chains c = g, g + rows, ...
template <bool DNLS_MODE>
__global__ void __launch_bounds__(kT)
k_blue_modes(const double2 *__restrict__ buf, uint32_t n, uint32_t L, uint32_t copies, const double2 *__restrict__ chirp,
             const double *__restrict__ omega, double *__restrict__ partial, uint32_t rows) {
	const uint32_t k = blockIdx.x * kT + threadIdx.x, g = blockIdx.y;
	if (k >= n) return;
	const uint32_t km = k ? n - k : 0;
	const double2 wk = chirp[k], wm = chirp[km];
	double sum = 0.0;
	for (uint32_t c = g; c < copies; c += rows) {
		const double2 z = cmul(buf[size_t(c) * L + k], wk);
		if constexpr (DNLS_MODE) {
			sum += z.x * z.x + z.y * z.y;
		} else {
			const double2 zm = cmul(buf[size_t(c) * L + km], wm);
			const double w = omega[k];
			const double a = 1.0 + w, b = 1.0 - w;
			const double er = fma(a, z.x, -b * zm.x);
			const double ei = fma(a, z.y, b * zm.y);
			sum += 0.25 * fma(er, er, ei * ei);
		}
	}
	partial[size_t(g) * n + k] += sum;
}

template <uint32_t M>
hipError_t row_attribute() {
	if (M * sizeof(double2) > 64 * 1024)
		return hipFuncSetAttribute((const void *) k_dnls_row<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                           int(M * sizeof(double2)));
	return hipSuccess;
}

// blue: bhat in [r][k] order (L), the L-point roots (L), the chirp (n).
template <uint32_t R, uint32_t M>
hipError_t launch_dnls_blue(hipStream_t st, double2 *state, double2 *buf, uint32_t n, uint32_t copies, uint32_t steps,
                            const DnlsArgs &a, const double2 *table, const double2 *blue) {
	// the row pass of a convolution: 4096-point rows on the register FFT with
	// the permuted bhat of dnls_blue_row_tables (after the 7 multipliers), else LDS
	auto blue_rows = [](hipStream_t st, double2 *buf, uint32_t copies, dim3 rows, size_t lds, uint32_t n,
	                    const double2 *twL, const double2 *bhat, const double2 *table) -> hipError_t {
		if constexpr (M == 4096) {
			const double2 *perm = table + size_t(7) * n;
			return launch_dnls_rows4096(st, buf, copies * R, R, perm + R * M, twL, perm);
		} else {
			hipLaunchKernelGGL(k_dnls_row<M>, rows, dim3(kT), lds, st, buf, R, twL, bhat);
			return hipGetLastError();
		}
	};
	if (hipError_t e = row_attribute<M>()) return e;
	constexpr uint32_t L = R * M;
	const double2 *bhat = blue, *twL = blue + L, *chirp = blue + 2 * L;
	const dim3 cols(uint32_t((uint64_t(copies) * M + kT - 1) / kT)), rows(copies * R), block(kT);
	const size_t lds = M * sizeof(double2);
	for (uint32_t s = 0; s < steps; ++s)
		for (int k = 0; k < 7; ++k) {
			const double th = (k == 0 && s) ? a.theta[8] : a.theta[k];
			if (s == 0 && k == 0)
				hipLaunchKernelGGL((k_blue_fwd<R, 0>), cols, block, 0, st, state, buf, n, L, M, copies, th, twL, chirp);
			else
				hipLaunchKernelGGL((k_blue_fwd<R, 1>), cols, block, 0, st, state, buf, n, L, M, copies, th, twL, chirp);
			if (hipError_t e = blue_rows(st, buf, copies, rows, lds, n, twL, bhat, table)) return e;
			hipLaunchKernelGGL((k_blue_inv<R, 1>), cols, block, 0, st, buf, state, n, L, M, copies, 0.0, twL, chirp,
			                   table + size_t(k) * n);
			hipLaunchKernelGGL((k_blue_fwd<R, 2>), cols, block, 0, st, state, buf, n, L, M, copies, 0.0, twL, chirp);
			if (hipError_t e = blue_rows(st, buf, copies, rows, lds, n, twL, bhat, table)) return e;
			if (s + 1 == steps && k == 6)
				hipLaunchKernelGGL((k_blue_inv<R, 3>), cols, block, 0, st, buf, state, n, L, M, copies, a.theta[7], twL,
				                   chirp, nullptr);
			else
				hipLaunchKernelGGL((k_blue_inv<R, 2>), cols, block, 0, st, buf, state, n, L, M, copies, 0.0, twL, chirp,
				                   nullptr);
		}
	return hipGetLastError();
}

template <bool DNLS_MODE, uint32_t R, uint32_t M>
hipError_t launch_energy_blue_rm(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double2 *blue,
                                 const double *omega, double *partial, uint32_t grid, double2 *scratch,
                                 uint32_t chunk) {
	if (hipError_t e = row_attribute<M>()) return e;
	constexpr uint32_t L = R * M;
	const double2 *bhat = blue, *twL = blue + L, *chirp = blue + 2 * L;
	for (uint32_t c0 = 0; c0 < copies; c0 += chunk) {
		const uint32_t cnt = std::min(chunk, copies - c0);
		const dim3 cols(uint32_t((uint64_t(cnt) * M + kT - 1) / kT)), block(kT);
		hipLaunchKernelGGL((k_blue_fwd<R, 3>), cols, block, 0, st, state + size_t(c0) * n, scratch, n, L, M, cnt, 0.0,
		                   twL, chirp);
		hipLaunchKernelGGL(k_dnls_row<M>, dim3(cnt * R), block, M * sizeof(double2), st, scratch, R, twL, bhat);
		hipLaunchKernelGGL((k_blue_inv<R, 0>), cols, block, 0, st, scratch, nullptr, n, L, M, cnt, 0.0, twL, chirp,
		                   nullptr);
		hipLaunchKernelGGL(k_blue_modes<DNLS_MODE>, dim3((n + kT - 1) / kT, grid), block, 0, st, scratch, n, L, cnt,
		                   chirp, omega, partial, grid);
	}
	return hipGetLastError();
}

} // namespace

// The DNLS step's two-level split: large_fft_split, and n = 8192 as 2 rows of
// 4096 (on the register FFT those beat the 8192-point LDS kernel: one HBM
// round trip per pass is cheaper than the LDS passes).
bool dnls_split(uint32_t n, uint32_t &R, uint32_t &M) {
	if (n == 8192) {
		R = 2, M = 4096;
		return true;
	}
	return large_fft_split(n, R, M);
}

bool dnls_two_level(uint32_t n) {
	uint32_t R, M;
	return dnls_split(n, R, M);
}

// Multipliers of the two-level path in the spectrum's [r][k] order:
// out[k][r M + q] = table[k][r + R q].  With M = 4096 the rows run on the
// radix-16 register FFT, whose spectrum is digit-reversed: slot e = t + 256 j
// (thread t, register j) holds q = (t >> 4) + 16 (t & 15) + 256 j; and the
// 4096-point roots e^{-2 pi i j/4096} follow at out[7 n, 7 n + 4096).
size_t dnls_two_level_table_size(uint32_t n) {
	uint32_t R, M;
	if (!dnls_split(n, R, M)) return 0;
	return size_t(7) * n + (M == 4096 ? 4096 : 0);
}

void dnls_two_level_tables(uint32_t n, const double2 *table, double2 *out) {
	uint32_t R, M;
	if (!dnls_split(n, R, M)) return;
	for (int k = 0; k < 7; ++k)
		for (uint32_t r = 0; r < R; ++r)
			for (uint32_t e = 0; e < M; ++e) {
				const uint32_t q = M == 4096 ? (e & 255) / 16 + 16 * (e & 15) + 256 * (e >> 8) : e;
				out[size_t(k) * n + size_t(r) * M + e] = table[size_t(k) * n + r + size_t(R) * q];
			}
	if (M == 4096)
		for (uint32_t j = 0; j < 4096; ++j) {
			const double ang = -2.0 * M_PI * double(j) / 4096.0;
			out[size_t(7) * n + j] = make_double2(std::cos(ang), std::sin(ang));
		}
}

uint32_t bluestein_two_level_size(uint32_t n) {
	if (n <= 4096 || (n & (n - 1)) == 0 || n > kLargeFftMax / 2) return 0;
	uint32_t L = 1;
	while (L < 2 * n - 1) L <<= 1;
	return L;
}

// DNLS through Bluestein with rows of 4096 points: the chirp's transform bhat
// in the row kernel's digit-reversed order (row r: slot e holds the [r][k]
// entry k = (e & 255)/16 + 16 (e & 15) + 256 (e >> 8)), then the 4096-point
// roots; appended after the 7 multipliers.  0 entries for other lengths.
size_t dnls_blue_row_table_size(uint32_t n) {
	const uint32_t L = bluestein_two_level_size(n);
	uint32_t R, M;
	return (L && large_fft_split(L, R, M) && M == 4096) ? size_t(L) + 4096 : 0;
}

void dnls_blue_row_tables(uint32_t n, double2 *out) {
	const uint32_t L = bluestein_two_level_size(n);
	uint32_t R, M;
	if (!L || !large_fft_split(L, R, M) || M != 4096) return;
	std::vector<double2> blue(size_t(2) * L + n);
	bluestein_tables(n, L, R, blue.data()); // bhat in [r][k] order at [0, L)
	for (uint32_t r = 0; r < R; ++r)
		for (uint32_t e = 0; e < 4096; ++e) {
			const uint32_t k = (e & 255) / 16 + 16 * (e & 15) + 256 * (e >> 8);
			out[size_t(r) * 4096 + e] = blue[size_t(r) * 4096 + k];
		}
	for (uint32_t j = 0; j < 4096; ++j) {
		const double ang = -2.0 * M_PI * double(j) / 4096.0;
		out[size_t(L) + j] = make_double2(std::cos(ang), std::sin(ang));
	}
}

void bluestein_two_level_tables(uint32_t n, double2 *out) {
	const uint32_t L = bluestein_two_level_size(n);
	uint32_t R, M;
	if (!L || !large_fft_split(L, R, M)) return;
	bluestein_tables(n, L, R, out);
}

hipError_t launch_energy_bluestein(hipStream_t st, bool dnls, const double2 *state, uint32_t n, uint32_t copies,
                                   const double2 *blue, const double *omega, double *partial, uint32_t grid,
                                   double2 *scratch, size_t scratch_sites) {
	const uint32_t L = bluestein_two_level_size(n);
	uint32_t R = 0, M = 0;
	if (!L || !blue || !scratch || scratch_sites < L || !large_fft_split(L, R, M)) return hipErrorInvalidValue;
	const uint32_t chunk = uint32_t(std::min<size_t>(scratch_sites / L, copies));
	hipLaunchKernelGGL(k_zero_rows, dim3(1024), dim3(256), 0, st, partial, size_t(grid) * n);
#define NLC_BLUE(DM, RR, MM)                                                                                        \
	return launch_energy_blue_rm<DM, RR, MM>(st, state, n, copies, blue, omega, partial, grid, scratch, chunk)
	if (dnls) {
		if (M == 8192) NLC_BLUE(true, 16, 8192);
		if (R == 4) NLC_BLUE(true, 4, 4096);
		if (R == 8) NLC_BLUE(true, 8, 4096);
		NLC_BLUE(true, 16, 4096);
	}
	if (M == 8192) NLC_BLUE(false, 16, 8192);
	if (R == 4) NLC_BLUE(false, 4, 4096);
	if (R == 8) NLC_BLUE(false, 8, 4096);
	NLC_BLUE(false, 16, 4096);
#undef NLC_BLUE
}

hipError_t launch_dnls_long(hipStream_t st, double2 *state, uint32_t n, uint32_t copies, uint32_t steps,
                            const double *theta9, const double2 *tw, const double2 *table, double2 *scratch,
                            const double2 *blue) {
	DnlsArgs a;
	for (int i = 0; i < 9; ++i) a.theta[i] = theta9[i];
	uint32_t R = 0, M = 0;
	if (const uint32_t L = bluestein_two_level_size(n)) {
		if (!scratch || !blue || !large_fft_split(L, R, M)) return hipErrorInvalidValue;
		if (M == 8192) return launch_dnls_blue<16, 8192>(st, state, scratch, n, copies, steps, a, table, blue);
		switch (R) {
		case 4: return launch_dnls_blue<4, 4096>(st, state, scratch, n, copies, steps, a, table, blue);
		case 8: return launch_dnls_blue<8, 4096>(st, state, scratch, n, copies, steps, a, table, blue);
		case 16: return launch_dnls_blue<16, 4096>(st, state, scratch, n, copies, steps, a, table, blue);
		}
		return hipErrorInvalidValue;
	}
	if (dnls_split(n, R, M)) {
		if (!scratch) return hipErrorInvalidValue;
		if (M == 8192) return launch_dnls_two_level<16, 8192>(st, state, scratch, n, copies, steps, a, tw, table);
		switch (R) {
		case 2: return launch_dnls_two_level<2, 4096>(st, state, scratch, n, copies, steps, a, tw, table);
		case 4: return launch_dnls_two_level<4, 4096>(st, state, scratch, n, copies, steps, a, tw, table);
		case 8: return launch_dnls_two_level<8, 4096>(st, state, scratch, n, copies, steps, a, tw, table);
		case 16: return launch_dnls_two_level<16, 4096>(st, state, scratch, n, copies, steps, a, tw, table);
		}
		return hipErrorInvalidValue;
	}
	return hipErrorInvalidValue;
}

} // namespace nlc
