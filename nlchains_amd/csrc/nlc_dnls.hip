// DNLS split-step Yoshida-6 integrator, chain resident on chip for a whole batch.
//
// Reference behaviour followed: DNLS.cpp:159-175 (per substep k: nonlinear
// rotation by beta dt c_k with the merged 2 c_0, forward FFT, Fourier multiplier
// T_k = exp(-i dt d_k w)/N (DNLS/dispersion.cpp:19-26), inverse FFT; a final
// beta dt c_7 rotation after the batch), kernels DNLS.cu:11-25 and the cuFFT
// load callbacks DNLS.cu:45-67; CPU form DNLS/cpu-implementation/DNLS.cpp:110-128.
//
// Design (MI355X-first): the reference issues 14 cuFFT launches per step, each
// a full HBM pass over the ensemble.  Here one workgroup owns one chain for all
// steps of the batch, so HBM is touched once per batch (16 B/site in and out).
//  * k_dnls_4096: 256 threads x 16 complex values in VGPRs.  The 4096-point FFT
//    is three radix-16 passes computed in registers (decimation in frequency);
//    between passes the data moves through LDS: one cross-wavefront exchange
//    and one exchange within each wavefront.  The forward transform leaves the
//    spectrum in digit-reversed order, the Fourier multiplier is read from a
//    table stored in that order, and the inverse transform (decimation in
//    time, conjugate twiddles) returns natural order, so no reordering pass is
//    ever made.  LDS: one padded region per wavefront holding what that
//    wavefront reads in the cross-wave exchange plus the rows of its in-wave
//    16 x 16 transposes; every slot is a per-lane base plus a compile-time
//    offset, and neither exchange has bank conflicts.  Each exchange moves the
//    real parts, then the imaginary parts, through 34 KB (component-split
//    exchanges of nlc_fft4096.inl): three chains fit a CU, and at 168 VGPRs
//    three wavefronts per SIMD, for six workgroup barriers per substep
//    instead of two (the 68 KB form with double2 slots held two chains per
//    CU and measured 2 % slower; results are the same bit for bit).
//  * k_dnls_4096_r8: the same transform as four radix-8 passes on 512 threads
//    x 8 values (more wavefronts per SIMD, one more LDS exchange).
//  * other lengths: nlc_dnls_generic.hip (LDS kernels up to 8192 sites) and
//    nlc_dnls_long.hip (two-level FFT in HBM beyond); launch_dnls dispatches.
#include "nlc_internal.hpp"
#include "nlc_fft_large.hpp"

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

namespace nlc {
namespace {

#include "nlc_dnls_common.inl"

// ---------------------------------------------------------------- radix 16

#include "nlc_fft4096.inl"


__global__ void __launch_bounds__(kT, 3)
k_dnls_4096(double2 *__restrict__ state, uint32_t copies, uint32_t steps, DnlsArgs a, const double2 *__restrict__ tw,
            const double2 *__restrict__ mult) {
	__shared__ Lds4352d L;
	const uint32_t chain = blockIdx.x, t = threadIdx.x;
	double2 *psi = state + size_t(chain) * 4096;
	double2 v[16]; // psi[t + 256 j] at v[pos(j)]
#pragma unroll
	for (int j = 0; j < 16; ++j) v[pos(j)] = psi[t + 256 * j];
	for (uint32_t s = 0; s < steps; ++s) {
#pragma unroll 1
		for (int k = 0; k < 7; ++k) {
			const double th = (k == 0 && s) ? a.theta[8] : a.theta[k];
			// the twiddles are re-read every substep (L1-resident) rather than held
			// in registers across the loop: registers are the occupancy limit here
			const double2 *twk = tw;
			asm volatile("" : "+s"(twk));
			nonlinear_all_halves(v, th);
			fft4096_fwd_split(v, L, twk);
			// Fourier multiplier, stored per thread in the digit-reversed order;
			// loaded and applied in two halves (the register budget of three
			// wavefronts per SIMD)
			const double2 *m = mult + size_t(k < 4 ? k : 6 - k) * 4096 + t;
#pragma unroll
			for (int h = 0; h < 2; ++h) {
#pragma unroll
				for (int j = 8 * h; j < 8 * h + 8; ++j) v[pos(j)] = cmul(v[pos(j)], m[256 * j]);
				__builtin_amdgcn_sched_barrier(0);
			}
			fft4096_inv_split(v, L, twk);
		}
	}
#pragma unroll
	for (int j = 0; j < 16; ++j) psi[t + 256 * j] = nonlinear(v[pos(j)], a.theta[7]);
}

// Row pass of the two-level FFT (nlc_dnls_long.hip) for rows of M = 4096
// points: forward radix-16 FFT in registers, the multipliers T[r + R k] of
// row r in this kernel's digit-reversed order (dnls_two_level_tables), the
// inverse FFT, and the inverse column twiddle W_n^{-m r} (as k_dnls_row);
// rows[row] in place, natural order in and out.  tw4096: the 4096-point
// roots; twn: the n-point ones (n = 4096 R).
__global__ void __launch_bounds__(kT, 2)
k_dnls_rows4096(double2 *__restrict__ rows, uint32_t R, const double2 *__restrict__ tw4096,
                const double2 *__restrict__ twn, const double2 *__restrict__ tk) {
	__shared__ Lds4352<double2> L;
	const uint32_t row = blockIdx.x, t = threadIdx.x;
	double2 *y = rows + size_t(row) * 4096;
	double2 v[16];
#pragma unroll
	for (int j = 0; j < 16; ++j) v[pos(j)] = y[t + 256 * j];
	fft4096_fwd(v, L, tw4096);
	const uint32_t r = row % R;
	const double2 *m = tk + size_t(r) * 4096 + t;
#pragma unroll
	for (int j = 0; j < 16; ++j) v[pos(j)] = cmul(v[pos(j)], m[256 * j]);
	fft4096_inv(v, L, tw4096);
#pragma unroll
	for (int j = 0; j < 16; ++j) {
		const uint32_t x = t + 256 * j; // x r < 4096 R = n
		y[x] = cmul(v[pos(j)], cconj(twn[x * r]));
	}
}

// ---------------------------------------------------------------- radix 8
//
// k_dnls_4096_r8: 512 threads x 8 complex values, 4096 = 8^4 in four radix-8
// passes.  Same DFT and twiddle work per point as the radix-16 kernel, half
// the values per thread, so about half the VGPRs and twice the wavefronts per
// SIMD to hide L2 and LDS latency, for one more LDS exchange per transform.
// Digits: x = x0 + 8 x1 + 64 x2 + 512 x3, k = k0 + 8 k1 + 64 k2 + 512 k3;
// thread t = lane + 64 w.

constexpr int kT8 = 512;

// position of element k of an 8-point transform in "transposed" order
__host__ __device__ constexpr int pos8(int k) { return 4 * (k & 1) + (k >> 1); }

// z * exp(SIGN 2 pi i j / 8), j = 0..3
template <int SIGN, int J>
__device__ __forceinline__ double2 w8mul(double2 z) {
	constexpr double h = 0.70710678118654752440;
	if constexpr (J == 0) return z;
	else if constexpr (J == 2) return SIGN < 0 ? make_double2(z.y, -z.x) : make_double2(-z.y, z.x);
	else if constexpr (J == 1) {
		return SIGN < 0 ? make_double2(h * (z.x + z.y), h * (z.y - z.x)) : make_double2(h * (z.x - z.y), h * (z.x + z.y));
	} else {
		return SIGN < 0 ? make_double2(h * (z.y - z.x), -h * (z.x + z.y)) : make_double2(-h * (z.x + z.y), h * (z.x - z.y));
	}
}

// 8-point DFT, natural in -> transposed out: radix-2 butterflies (j, j + 4)
// with twiddles W8^j, then DFT4 of the sums (even outputs, registers 0..3)
// and of the differences (odd outputs, registers 4..7).
template <int SIGN>
__device__ __forceinline__ void dft8_n2t(double2 v[8]) {
	double2 d[4];
#pragma unroll
	for (int j = 0; j < 4; ++j) {
		d[j] = csub(v[j], v[j + 4]);
		v[j] = cadd(v[j], v[j + 4]);
	}
	v[4] = d[0];
	v[5] = w8mul<SIGN, 1>(d[1]);
	v[6] = w8mul<SIGN, 2>(d[2]);
	v[7] = w8mul<SIGN, 3>(d[3]);
	dft4<SIGN>(v[0], v[1], v[2], v[3]);
	dft4<SIGN>(v[4], v[5], v[6], v[7]);
}

// transposed in -> natural out (decimation in time): DFT4 of the even and of
// the odd elements, twiddle the odd ones by W8^j, butterflies (j, j + 4).
template <int SIGN>
__device__ __forceinline__ void dft8_t2n(double2 v[8]) {
	dft4<SIGN>(v[0], v[1], v[2], v[3]);
	dft4<SIGN>(v[4], v[5], v[6], v[7]);
	const double2 o[4] = {v[4], w8mul<SIGN, 1>(v[5]), w8mul<SIGN, 2>(v[6]), w8mul<SIGN, 3>(v[7])};
#pragma unroll
	for (int j = 0; j < 4; ++j) {
		v[j + 4] = csub(v[j], o[j]);
		v[j] = cadd(v[j], o[j]);
	}
}

// v[at(j)] *= W^{e0 + j de}, j = 0..7, W = exp(-2 pi i/4096) = tw[1] (conjugated
// when CONJ), at(j) = pos8(j) when TRANSPOSED; powers refreshed from the table
// at j = 0 and 4.  ONE0 asserts e0 == 0.
template <bool CONJ, bool ONE0, bool TRANSPOSED>
__device__ __forceinline__ void twiddle_run8(double2 v[8], const double2 *__restrict__ tw, uint32_t e0, uint32_t de) {
	const double2 w = tw[de & 4095u];
	const double2 wc = CONJ ? cconj(w) : w;
#pragma unroll
	for (int q = 0; q < 2; ++q) {
		double2 r = tw[(e0 + uint32_t(4 * q) * de) & 4095u];
		if (CONJ) r = cconj(r);
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const int idx = 4 * q + j;
			const int at = TRANSPOSED ? pos8(idx) : idx;
			if (idx || !ONE0) v[at] = cmul(v[at], r);
			if (j < 3) r = cmul(r, wc);
		}
	}
}

// LDS layout of the radix-8 kernel: wavefront w owns the 640 entries from
// kR8Region w (dynamic LDS, 8 x 640 x 16 B = 80 KB per chain, two chains per
// CU).  The cross-wave exchange puts element (t, k0) at kR8Region k0 + t, so
// the wavefront w = k0 that gathers it reads only its own region; the two
// in-wave exchanges are 8 x 8 transposes per group of 8 lanes through padded
// rows (stride 9) inside the same region.  Every slot is a per-lane base plus
// a compile-time offset (no per-register address arithmetic), and 16
// consecutive lanes touch distinct bank quads except for one pair in the
// read of exchange 2.
//   exchange 2: element (x0, x1, k1) at 9 (x0 + 8 x1) + k1, written by lane
//               x0 + 8 x1 (offset k1), read by lane x0 + 8 k1 (offset 72 x1)
//   exchange 3: element (x0, k1, k2) at 72 k1 + 9 x0 + k2, written by lane
//               x0 + 8 k1 (offset k2), read by lane k2 + 8 k1 (offset 9 x0)
constexpr uint32_t kR8Region = 640;
constexpr uint32_t kR8Lds = 8 * kR8Region * sizeof(double2);

// Forward FFT.  In: thread t holds psi[t + 512 x3] at v[pos8(x3)].  Out: thread
// t = k2 + 8 k1 + 64 k0 holds X[k0 + 8 k1 + 64 k2 + 512 k3] at v[pos8(k3)].
// LDS hazards as in fft4096_fwd: exchange 1 writes only slots the same thread
// read last, every other write lands in the wavefront's own region.
__device__ __forceinline__ void fft4096_fwd8(double2 v[8], double2 *L, const double2 *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6, lo = lane & 7, hi = lane >> 3;
	double2 *own = L + kR8Region * w;
	// A: over x3 -> k0 (natural), twiddle W^{t k0}
	dft8_t2n<-1>(v);
	twiddle_run8<false, true, false>(v, tw, 0, t);
	// exchange 1 (cross-wave); thread (lane, w = k0) gathers x2
	double2 *x1w = L + t;
#pragma unroll
	for (int k0 = 0; k0 < 8; ++k0) x1w[kR8Region * k0] = v[k0];
	lds_barrier();
	const double2 *x1r = own + lane;
#pragma unroll
	for (int x2 = 0; x2 < 8; ++x2) v[x2] = x1r[64 * x2];
	// B: over x2 -> k1 (transposed), twiddle W^{8 (x0 + 8 x1) k1}, x0 + 8 x1 = lane
	dft8_n2t<-1>(v);
	twiddle_run8<false, true, true>(v, tw, 0, 8 * lane);
	// exchange 2 (in-wave): lane (x0 = lo, x1 = hi) -> lane (x0 = lo, k1 = hi)
	wave_sync();
	double2 *x2w = own + 9 * lane;
#pragma unroll
	for (int k1 = 0; k1 < 8; ++k1) x2w[k1] = v[pos8(k1)];
	wave_sync();
	const double2 *x2r = own + 9 * lo + hi;
#pragma unroll
	for (int x1 = 0; x1 < 8; ++x1) v[x1] = x2r[72 * x1];
	// C: over x1 -> k2 (transposed), twiddle W^{64 x0 k2}
	dft8_n2t<-1>(v);
	twiddle_run8<false, true, true>(v, tw, 0, 64 * lo);
	// exchange 3 (in-wave): lane (x0 = lo, k1 = hi) -> lane (k2 = lo, k1 = hi)
	wave_sync();
	double2 *x3w = own + 72 * hi + 9 * lo;
#pragma unroll
	for (int k2 = 0; k2 < 8; ++k2) x3w[k2] = v[pos8(k2)];
	wave_sync();
	const double2 *x3r = own + 72 * hi + lo;
#pragma unroll
	for (int x0 = 0; x0 < 8; ++x0) v[x0] = x3r[9 * x0];
	// D: over x0 -> k3 (transposed)
	dft8_n2t<-1>(v);
}

// Inverse (unnormalised) FFT, the exact reverse of fft4096_fwd8.
__device__ __forceinline__ void fft4096_inv8(double2 v[8], double2 *L, const double2 *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6, lo = lane & 7, hi = lane >> 3;
	double2 *own = L + kR8Region * w;
	// D': over k3 -> x0 (natural), twiddle W^{-64 x0 k2}, k2 = lo
	dft8_t2n<+1>(v);
	twiddle_run8<true, true, false>(v, tw, 0, 64 * lo);
	// exchange 3': lane (k2 = lo, k1 = hi) -> lane (x0 = lo, k1 = hi)
	wave_sync();
	double2 *x3w = own + 72 * hi + lo;
#pragma unroll
	for (int x0 = 0; x0 < 8; ++x0) x3w[9 * x0] = v[x0];
	wave_sync();
	const double2 *x3r = own + 72 * hi + 9 * lo;
#pragma unroll
	for (int k2 = 0; k2 < 8; ++k2) v[pos8(k2)] = x3r[k2];
	// C': over k2 -> x1 (natural), twiddle W^{-8 (x0 + 8 x1) k1}, x0 = lo, k1 = hi
	dft8_t2n<+1>(v);
	twiddle_run8<true, false, false>(v, tw, 8 * lo * hi, 64 * hi);
	// exchange 2': lane (x0 = lo, k1 = hi) -> lane (x0 = lo, x1 = hi)
	wave_sync();
	double2 *x2w = own + 9 * lo + hi;
#pragma unroll
	for (int x1 = 0; x1 < 8; ++x1) x2w[72 * x1] = v[x1];
	wave_sync();
	const double2 *x2r = own + 9 * lane;
#pragma unroll
	for (int k1 = 0; k1 < 8; ++k1) v[pos8(k1)] = x2r[k1];
	// B': over k1 -> x2 (natural), twiddle W^{-(lane + 64 x2) k0}, k0 = w
	dft8_t2n<+1>(v);
	twiddle_run8<true, false, false>(v, tw, lane * w, 64 * w);
	// exchange 1' (cross-wave), written into the wavefront's own region
	wave_sync();
	double2 *x1w = own + lane;
#pragma unroll
	for (int x2 = 0; x2 < 8; ++x2) x1w[64 * x2] = v[x2];
	lds_barrier();
	const double2 *x1r = L + t;
#pragma unroll
	for (int k0 = 0; k0 < 8; ++k0) v[k0] = x1r[kR8Region * k0];
	// A': over k0 -> x3 (transposed)
	dft8_n2t<+1>(v);
}

__global__ void __launch_bounds__(kT8, 4)
k_dnls_4096_r8(double2 *__restrict__ state, uint32_t copies, uint32_t steps, DnlsArgs a,
               const double2 *__restrict__ tw, const double2 *__restrict__ mult) {
	extern __shared__ double2 lds8[]; // kR8Lds bytes
	const uint32_t chain = blockIdx.x, t = threadIdx.x;
	double2 *psi = state + size_t(chain) * 4096;
	double2 v[8]; // psi[t + 512 j] at v[pos8(j)]
#pragma unroll
	for (int j = 0; j < 8; ++j) v[pos8(j)] = psi[t + 512 * j];
	for (uint32_t s = 0; s < steps; ++s) {
#pragma unroll 1
		for (int k = 0; k < 7; ++k) {
			const double th = (k == 0 && s) ? a.theta[8] : a.theta[k];
			const double2 *twk = tw;
			asm volatile("" : "+s"(twk));
			nonlinear_all(v, th);
			fft4096_fwd8(v, lds8, twk);
			// Fourier multiplier, stored per thread in the digit-reversed order;
			// the fence keeps its loads out of the transform (4 wavefronts per
			// SIMD hide their latency, registers are the scarce resource here)
			asm volatile("" ::: "memory");
			const double2 *m = mult + size_t(k < 4 ? k : 6 - k) * 4096 + t;
#pragma unroll
			for (int j = 0; j < 8; ++j) v[pos8(j)] = cmul(v[pos8(j)], m[512 * j]);
			fft4096_inv8(v, lds8, twk);
		}
	}
#pragma unroll
	for (int j = 0; j < 8; ++j) psi[t + 512 * j] = nonlinear(v[pos8(j)], a.theta[7]);
}

} // namespace

bool dnls_supported(uint32_t n) {
	uint32_t R, M;
	const bool pow2 = n >= 2 && (n & (n - 1)) == 0;
	return (pow2 && n <= kDnlsLdsMax) || mixed_radix_plan(n, nullptr) || dnls_bluestein_size(n) != 0 ||
	       large_fft_split(n, R, M) || bluestein_two_level_size(n) != 0;
}

// Fourier multiplier tables (host side).  table[k][j] = T_k[j] for the generic
// kernel; for n = 4096 the 4 distinct tables (d_k = d_{6-k}) are stored in the
// order the register kernels read them:
//   radix 16: entry [k][256 j + t] = T_k[(t >> 4) + 16 (t & 15) + 256 j],
//   radix 8:  entry [k][512 j + t] = T_k[(t >> 6) + 8 ((t >> 3) & 7) + 64 (t & 7) + 512 j].
bool dnls_fast_path(uint32_t n) { return n == 4096; }

hipError_t launch_dnls_rows4096(hipStream_t st, double2 *rows, uint32_t nrows, uint32_t R, const double2 *tw4096,
                                const double2 *twn, const double2 *tk) {
	hipLaunchKernelGGL(k_dnls_rows4096, dim3(nrows), dim3(kT), 0, st, rows, R, tw4096, twn, tk);
	return hipGetLastError();
}

void dnls_permute_tables(uint32_t n, int variant, const double2 *table, double2 *out) {
	for (int k = 0; k < 4; ++k)
		for (uint32_t e = 0; e < n; ++e) {
			uint32_t src;
			if (variant == kDnlsRadix8) {
				const uint32_t j = e >> 9, t = e & 511;
				src = (t >> 6) + 8 * ((t >> 3) & 7) + 64 * (t & 7) + 512 * j;
			} else {
				const uint32_t j = e >> 8, t = e & 255;
				src = (t >> 4) + 16 * (t & 15) + 256 * j;
			}
			out[size_t(k) * n + e] = table[size_t(k) * n + src];
		}
}

hipError_t launch_dnls(hipStream_t st, double2 *state, uint32_t n, uint32_t copies, uint32_t steps,
                       const double *theta9, const double2 *tw, const double2 *table, int variant, double2 *scratch,
                       const double2 *blue) {
	if (!dnls_fast_path(n)) {
		if (!mixed_radix_plan(n, nullptr) && (bluestein_two_level_size(n) || dnls_two_level(n)))
			return launch_dnls_long(st, state, n, copies, steps, theta9, tw, table, scratch, blue);
		return launch_dnls_generic(st, state, n, copies, steps, theta9, tw, table);
	}
	DnlsArgs a;
	for (int i = 0; i < 9; ++i) a.theta[i] = theta9[i];
	if (variant == kDnlsRadix8) {
		hipError_t e = hipFuncSetAttribute((const void *) k_dnls_4096_r8, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   int(kR8Lds));
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(k_dnls_4096_r8, dim3(copies), dim3(kT8), kR8Lds, st, state, copies, steps, a, tw, table);
	} else {
		hipLaunchKernelGGL(k_dnls_4096, dim3(copies), dim3(kT), 0, st, state, copies, steps, a, tw, table);
	}
	return hipGetLastError();
}

} // namespace nlc
