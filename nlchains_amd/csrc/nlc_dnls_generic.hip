// DNLS split step for chain lengths up to 8192 other than 4096 (the register
// kernels of nlc_dnls.hip), chain resident in LDS for a whole batch.
//
// Reference behaviour followed: DNLS.cpp:159-175 (as nlc_dnls.hip).
//
//  * k_dnls_lds: any power-of-two n <= 8192, chain in LDS, radix-4 Stockham
//    passes (the inverse transform as conj(FFT(conj(.)))).
//  * k_dnls_mixed: the other n <= 8192 whose prime factors are 2, 3, 5 and 7
//    (3000, 5000, 6000, ...), the same with the mixed-radix Stockham passes of
//    nlc_fft_mixed.hpp.
//  * k_dnls_bluestein: any other n <= 4096.  Each n-point DFT is Bluestein's
//    chirp-z convolution (L. Bluestein, 1970): with w_j = e^{-i pi j^2/n},
//    X_k = w_k sum_x (w_x psi_x) conj(w_{k-x}), a cyclic convolution of length
//    M >= 2n - 1 (a power of two) done as two M-point FFTs in LDS.
#include "nlc_internal.hpp"
#include "nlc_fft_large.hpp"
#include "nlc_fft_mixed.hpp"

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

namespace nlc {
namespace {

#include "nlc_dnls_common.inl"

// One batch of steps on a chain held in LDS by threads tid < nthr; fft(buf) is
// the forward n-point FFT in place, natural order in and out.  Between substeps
// the buffer holds conj(psi): the inverse transform is conj(FFT(conj(X T))).
// A slot past the last chain (live = false) runs on zeros and writes nothing,
// so that every thread reaches every barrier.
// NT: the thread count as a compile-time constant (0: nthr at run time).
template <uint32_t NT, class Fft>
__device__ __forceinline__ void dnls_lds_batch(double2 *buf, double2 *psi, bool live, uint32_t n, uint32_t tid,
                                               uint32_t nthr_rt, uint32_t steps, const DnlsArgs &a,
                                               const double2 *__restrict__ table, Fft fft) {
	const uint32_t nthr = NT ? NT : nthr_rt;
	for (uint32_t x = tid; x < n; x += nthr) buf[x] = live ? psi[x] : make_double2(0.0, 0.0);
	bool conj_state = false;
	__syncthreads();
	for (uint32_t s = 0; s < steps; ++s) {
		for (int k = 0; k < 7; ++k) {
			const double th = (k == 0 && s) ? a.theta[8] : a.theta[k];
			for (uint32_t x = tid; x < n; x += nthr) {
				const double2 p = conj_state ? cconj(buf[x]) : buf[x];
				buf[x] = nonlinear(p, th);
			}
			__syncthreads();
			fft(buf);
			const double2 *tk = table + size_t(k) * n;
			for (uint32_t j = tid; j < n; j += nthr) buf[j] = cconj(cmul(buf[j], tk[j]));
			__syncthreads();
			fft(buf);
			conj_state = true;
		}
	}
	if (live)
		for (uint32_t x = tid; x < n; x += nthr) {
			const double2 p = conj_state ? cconj(buf[x]) : buf[x];
			psi[x] = nonlinear(p, a.theta[7]);
		}
}

// Threads per chain of k_dnls_lds: kT from 4 kT points up, else n/4 (one
// radix-4 butterfly per thread and pass) with kT/(n/4) chains per workgroup.
__host__ __device__ inline uint32_t lds_chain_threads(uint32_t n) {
	return n >= 4 * uint32_t(kT) ? uint32_t(kT) : (n >= 4 ? n / 4 : 1u);
}

// SHARED: n < 4 kT, kT/(n/4) chains per workgroup; else one chain.
template <uint32_t MAXN, bool SHARED>
__global__ void __launch_bounds__(kT)
k_dnls_lds(double2 *__restrict__ state, uint32_t n, uint32_t copies, uint32_t steps, DnlsArgs a,
           const double2 *__restrict__ tw, const double2 *__restrict__ table) {
	extern __shared__ double2 buf[];
	if constexpr (SHARED) {
		const uint32_t nthr = lds_chain_threads(n), per_group = kT / nthr;
		const uint32_t sub = threadIdx.x / nthr, tid = threadIdx.x % nthr;
		const uint32_t chain = blockIdx.x * per_group + sub;
		const bool live = chain < copies;
		dnls_lds_batch<0>(buf + size_t(sub) * n, state + size_t(live ? chain : 0) * n, live, n, tid, nthr, steps, a,
		                  table, [&](double2 *b) { lds_fft_fwd<MAXN>(b, n, tw, 1, tid, nthr); });
	} else {
		dnls_lds_batch<kT>(buf, state + size_t(blockIdx.x) * n, true, n, threadIdx.x, kT, steps, a, table,
		                   [&](double2 *b) { lds_fft_fwd<MAXN>(b, n, tw); });
	}
}

// Chains per workgroup of k_dnls_mixed below 512 sites: ceil(n/4) threads per
// chain (no pass then has more than 2 butterflies per thread), kT / that many
// chains; the kT mod nthr threads left over run a dummy slot of their own.
__host__ __device__ inline uint32_t mixed_chain_threads(uint32_t n) { return (n + 3) / 4; }

// SHARED: several chains per workgroup (n <= 512), MAXN = 1024 bounds n per chunk.
template <uint32_t MAXN, bool SHARED>
__global__ void __launch_bounds__(kT)
k_dnls_mixed(double2 *__restrict__ state, uint32_t n, uint32_t copies, uint32_t steps, DnlsArgs a, MixedPlan plan,
             const double2 *__restrict__ tw, const double2 *__restrict__ table) {
	extern __shared__ double2 buf[];
	if constexpr (SHARED) {
		const uint32_t nthr = mixed_chain_threads(n), per_group = kT / nthr;
		const uint32_t sub = threadIdx.x / nthr, tid = threadIdx.x % nthr; // sub == per_group: the dummy slot
		const uint32_t chain = blockIdx.x * per_group + sub;
		const bool live = sub < per_group && chain < copies;
		dnls_lds_batch<0>(buf + size_t(sub) * n, state + size_t(live ? chain : 0) * n, live, n, tid, nthr, steps, a,
		                  table, [&](double2 *b) { mixed::lds_fft<MAXN, kT, false>(b, n, tw, plan, tid, nthr); });
	} else {
		dnls_lds_batch<kT>(buf, state + size_t(blockIdx.x) * n, true, n, threadIdx.x, kT, steps, a, table,
		                   [&](double2 *b) { mixed::lds_fft<MAXN, kT>(b, n, tw, plan); });
	}
}

// Cyclic convolution of buf[0..M) with b_j = conj(w_j) (|j| < n, zero
// elsewhere), returned conjugated: buf <- conj(IFFT(FFT(buf) bhat)) with
// bhat = FFT(b)/M, as FFT(conj(FFT(buf) bhat)).
template <uint32_t MAXN>
__device__ void bluestein_conv(double2 *buf, uint32_t M, const double2 *__restrict__ twm,
                               const double2 *__restrict__ bhat) {
	lds_fft_fwd<MAXN>(buf, M, twm);
	for (uint32_t j = threadIdx.x; j < M; j += kT) buf[j] = cconj(cmul(buf[j], bhat[j]));
	__syncthreads();
	lds_fft_fwd<MAXN>(buf, M, twm);
}

// Chain in LDS as u_x = w_x psi_x (x < n), M entries.  One substep:
//   nonlinear: |u_x| = |psi_x|, so u_x <- u_x e^{-i theta |u_x|^2};
//   forward DFT: conv(u) holds conj(C_k), X_k = w_k C_k;
//   backward DFT of Y = X T as conj(DFT(conj Y)): its Bluestein input
//   conj(Y_k) w_k = conj(C_k) conj(T_k), and the second conv(.) leaves
//   conj(C'_x) = w_x psi'_x, the same representation: the chirps cancel
//   between substeps and are applied only at load and store.
// table: 7 multipliers of n entries, then bhat (M), the M-point twiddles (M),
// the chirp w (n).
template <uint32_t MAXN>
__global__ void __launch_bounds__(kT)
k_dnls_bluestein(double2 *__restrict__ state, uint32_t n, uint32_t M, uint32_t copies, uint32_t steps, DnlsArgs a,
                 const double2 *__restrict__ table) {
	extern __shared__ double2 buf[];
	const double2 *bhat = table + size_t(7) * n, *twm = bhat + M, *chirp = twm + M;
	const uint32_t chain = blockIdx.x, tid = threadIdx.x;
	const double2 zero = make_double2(0.0, 0.0);
	double2 *psi = state + size_t(chain) * n;
	for (uint32_t x = tid; x < n; x += kT) buf[x] = cmul(psi[x], chirp[x]);
	for (uint32_t s = 0; s < steps; ++s) {
		for (int k = 0; k < 7; ++k) {
			const double th = (k == 0 && s) ? a.theta[8] : a.theta[k];
			for (uint32_t x = tid; x < M; x += kT) buf[x] = x < n ? nonlinear(buf[x], th) : zero;
			__syncthreads();
			bluestein_conv<MAXN>(buf, M, twm, bhat);
			const double2 *tk = table + size_t(k) * n;
			for (uint32_t j = tid; j < M; j += kT) buf[j] = j < n ? cmul(buf[j], cconj(tk[j])) : zero;
			__syncthreads();
			bluestein_conv<MAXN>(buf, M, twm, bhat);
		}
	}
	for (uint32_t x = tid; x < n; x += kT) psi[x] = nonlinear(cmul(buf[x], cconj(chirp[x])), a.theta[7]);
}

template <typename K>
hipError_t launch_lds_chain_kernel(K kernel, size_t lds) {
	if (lds > 64 * 1024)
		return hipFuncSetAttribute((const void *) kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
	return hipSuccess;
}

bool pow2(uint32_t n) { return n >= 2 && (n & (n - 1)) == 0; }

// In-place radix-2 FFT on the host (forward sign, exact twiddles), for bhat.
void host_fft(std::vector<double2> &v) {
	const size_t n = v.size();
	for (size_t i = 1, j = 0; i < n; ++i) {
		size_t bit = n >> 1;
		for (; j & bit; bit >>= 1) j ^= bit;
		j ^= bit;
		if (i < j) std::swap(v[i], v[j]);
	}
	for (size_t len = 2; len <= n; len <<= 1)
		for (size_t j = 0; j < len / 2; ++j) {
			const double ang = -2.0 * M_PI * double(j) / double(len);
			const double wr = std::cos(ang), wi = std::sin(ang);
			for (size_t b = 0; b < n; b += len) {
				double2 &u = v[b + j], &t = v[b + j + len / 2];
				const double tr = t.x * wr - t.y * wi, ti = t.x * wi + t.y * wr;
				t = make_double2(u.x - tr, u.y - ti);
				u = make_double2(u.x + tr, u.y + ti);
			}
		}
}

} // namespace

uint32_t dnls_bluestein_size(uint32_t n) {
	if (n < 2 || pow2(n) || mixed_radix_plan(n, nullptr)) return 0;
	uint32_t M = 1;
	while (M < 2 * n - 1) M <<= 1;
	return M <= kDnlsBluesteinMax ? M : 0;
}

// bhat = FFT_M(b)/M of b_j = conj(w_j) (|j| < n, placed cyclically), the M-point
// roots e^{-2 pi i j/M} and the chirp w_j = e^{-i pi j^2/n}: out[0, M), out[M, 2M),
// out[2M, 2M + n).  With R > 1 bhat is stored in the two-level order [r][k]
// (bhat[r + R k] at r M/R + k).
void bluestein_tables(uint32_t n, uint32_t M, uint32_t R, double2 *out) {
	std::vector<double2> b(M, make_double2(0.0, 0.0));
	// w_j = e^{-i pi j^2/n}, with j^2 reduced mod 2n so the angle stays in [0, 2 pi)
	auto chirp_angle = [n](uint32_t j) { return M_PI * double((uint64_t(j) * j) % (2 * uint64_t(n))) / double(n); };
	for (uint32_t j = 0; j < n; ++j) {
		const double ang = chirp_angle(j);
		b[j] = make_double2(std::cos(ang), std::sin(ang)); // conj(w_j)
		if (j) b[M - j] = b[j];
		out[size_t(2) * M + j] = make_double2(std::cos(ang), -std::sin(ang)); // w_j
	}
	host_fft(b);
	const uint32_t rows = M / R;
	for (uint32_t j = 0; j < M; ++j) {
		const uint32_t at = R > 1 ? (j % R) * rows + j / R : j;
		out[at] = make_double2(b[j].x / M, b[j].y / M);
		const double ang = 2.0 * M_PI * double(j) / double(M);
		out[M + j] = make_double2(std::cos(ang), -std::sin(ang));
	}
}

void dnls_bluestein_tables(uint32_t n, double2 *out) { bluestein_tables(n, dnls_bluestein_size(n), 1, out); }

hipError_t launch_dnls_generic(hipStream_t st, double2 *state, uint32_t n, uint32_t copies, uint32_t steps,
                               const double *theta9, const double2 *tw, const double2 *table) {
	DnlsArgs a;
	for (int i = 0; i < 9; ++i) a.theta[i] = theta9[i];
	if (pow2(n) && n <= kDnlsLdsMax) {
		const uint32_t per_group = kT / lds_chain_threads(n);
		const size_t lds = sizeof(double2) * n * per_group;
		if (per_group > 1) {
			hipLaunchKernelGGL((k_dnls_lds<4 * kT, true>), dim3((copies + per_group - 1) / per_group), dim3(kT), lds, st,
			                   state, n, copies, steps, a, tw, table);
		} else if (n <= 4096) {
			hipLaunchKernelGGL((k_dnls_lds<4096, false>), dim3(copies), dim3(kT), lds, st, state, n, copies, steps, a, tw,
			                   table);
		} else {
			hipError_t e = launch_lds_chain_kernel(k_dnls_lds<kDnlsLdsMax, false>, lds);
			if (e != hipSuccess) return e;
			hipLaunchKernelGGL((k_dnls_lds<kDnlsLdsMax, false>), dim3(copies), dim3(kT), lds, st, state, n, copies, steps,
			                   a, tw, table);
		}
	} else if (MixedPlan plan; mixed_radix_plan(n, &plan)) {
		const uint32_t nthr = mixed_chain_threads(n), per_group = kT / nthr;
		if (per_group > 1) {
			const uint32_t slots = (kT + nthr - 1) / nthr; // with the dummy slot of the left-over threads
			hipLaunchKernelGGL((k_dnls_mixed<4 * kT, true>), dim3((copies + per_group - 1) / per_group), dim3(kT),
			                   sizeof(double2) * n * slots, st, state, n, copies, steps, a, plan, tw, table);
		} else if (n <= 4096) {
			hipLaunchKernelGGL((k_dnls_mixed<4096, false>), dim3(copies), dim3(kT), sizeof(double2) * n, st, state, n,
			                   copies, steps, a, plan, tw, table);
		} else {
			const size_t lds = sizeof(double2) * n;
			hipError_t e = launch_lds_chain_kernel(k_dnls_mixed<kDnlsLdsMax, false>, lds);
			if (e != hipSuccess) return e;
			hipLaunchKernelGGL((k_dnls_mixed<kDnlsLdsMax, false>), dim3(copies), dim3(kT), lds, st, state, n, copies, steps,
			                   a, plan, tw, table);
		}
	} else if (const uint32_t M = dnls_bluestein_size(n)) {
		const size_t lds = sizeof(double2) * M;
		if (M <= 4096) {
			hipLaunchKernelGGL(k_dnls_bluestein<4096>, dim3(copies), dim3(kT), lds, st, state, n, M, copies, steps, a,
			                   table);
		} else {
			hipError_t e = launch_lds_chain_kernel(k_dnls_bluestein<kDnlsBluesteinMax>, lds);
			if (e != hipSuccess) return e;
			hipLaunchKernelGGL(k_dnls_bluestein<kDnlsBluesteinMax>, dim3(copies), dim3(kT), lds, st, state, n, M,
			                   copies, steps, a, table);
		}
	} else {
		return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

} // namespace nlc
