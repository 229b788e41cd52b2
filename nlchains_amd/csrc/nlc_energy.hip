// Linear-mode energy pass, per-mode reduction over copies, normalisation and
// entropies, and per-chain total energies (all FP64).
//
// Reference behaviour followed:
//   plane-wave energies: DNKG_FPUT_Toda.cpp:113-123 (memset + 2x memcpy2D +
//     cufftExecZ2Z of q and p separately) and make_linenergies_kernel
//     DNKG_FPUT_Toda.cu:296-323 (thread per mode, serial loop over copies);
//   DNLS energies: DNLS.cpp:142-144, DNLS.cu:70-90;
//   normalisation and entropies: common/results_entropy.cpp:11-46 (host code
//     in the reference, FP64 on the GPU here);
//   chain energies: supporting-material/post-processing/nlchains.m:40-48.
//
// Design: one workgroup owns a block of chains; each chain is loaded once
// (coalesced, 16 B/site), transformed in LDS with an in-place radix-4 Stockham
// FFT, and its per-mode energies are accumulated in registers.  Because the
// planar pair {q, p} already is the complex number z = q + i p, a single
// complex FFT per chain yields both spectra:
//   Q_k = (Z_k + conj Z_{n-k})/2,  P_k = (Z_k - conj Z_{n-k})/(2i)
// and the reference's e_k = w(w|Q|^2 + 2(Q_i P_r - Q_r P_i)) + |P|^2 = |P - i w Q|^2
// collapses to e_k = |(1+w) Z_k - (1-w) conj Z_{n-k}|^2 / 4 (half the FFT work
// of transforming q and p separately).  Per-block partial sums are reduced in
// a fixed order, so results are bit-reproducible run to run.
#include "nlc_internal.hpp"
#include "nlc_fft_large.hpp"
#include "nlc_fft_mixed.hpp"

#include <algorithm>

namespace nlc {
namespace {

constexpr int kThreads = 256;
// The LDS FFT is instantiated for n <= MAXN with MAXN = 4096 (the common
// sizes, 64 KB of LDS) and MAXN = 8192 (128 KB, one workgroup per CU).
constexpr uint32_t kEnergyFftLarge = 8192;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
	return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// Radix-4 Stockham butterfly j (< n/4) of the pass with stride ns: inputs
// v[r] = buf[j + r n/4] (read before the pass's barrier), outputs at
// (j/ns) 4 ns + j%ns + r ns.  span = n / (4 ns) is the twiddle stride.
__device__ __forceinline__ void radix4_write(double2 *buf, const double2 v[4], uint32_t j, uint32_t ns, uint32_t span,
                                             const double2 *__restrict__ tw) {
	const uint32_t k = j % ns;
	const double2 a0 = v[0];
	const double2 a1 = cmul(v[1], tw[k * span]);
	const double2 a2 = cmul(v[2], tw[2 * k * span]);
	const double2 a3 = cmul(v[3], tw[3 * k * span]);
	const double2 t0 = make_double2(a0.x + a2.x, a0.y + a2.y);
	const double2 t1 = make_double2(a0.x - a2.x, a0.y - a2.y);
	const double2 t2 = make_double2(a1.x + a3.x, a1.y + a3.y);
	const double2 t3 = make_double2(a1.x - a3.x, a1.y - a3.y);
	const uint32_t o = (j / ns) * ns * 4 + k;
	buf[o] = make_double2(t0.x + t2.x, t0.y + t2.y);
	buf[o + ns] = make_double2(t1.x + t3.y, t1.y - t3.x);     // t1 - i t3
	buf[o + 2 * ns] = make_double2(t0.x - t2.x, t0.y - t2.y);
	buf[o + 3 * ns] = make_double2(t1.x - t3.y, t1.y + t3.x); // t1 + i t3
}

// Final radix-2 butterfly j (< n/2): inputs buf[j], buf[j + n/2]; n here is
// the length of the twiddle table (the FFT length times its stride).
__device__ __forceinline__ void radix2_write(double2 *buf, double2 v0, double2 v1, uint32_t j, uint32_t ns, uint32_t n,
                                             const double2 *__restrict__ tw) {
	const uint32_t k = j % ns;
	const double2 b = cmul(v1, tw[k * (n / (2 * ns))]);
	const uint32_t o = (j / ns) * ns * 2 + k;
	buf[o] = make_double2(v0.x + b.x, v0.y + b.y);
	buf[o + ns] = make_double2(v0.x - b.x, v0.y - b.y);
}

// In-place radix-4 (plus one final radix-2 pass when log2 n is odd) Stockham
// FFT of buf[0..n), forward sign, natural-order output.  tw[j] = exp(-2 pi i j/n).
// Each pass reads all inputs into registers, synchronises, then writes.
// tws: stride of tw (an M-point FFT can use the N-point table with tws = N/M).
template <uint32_t MAXN>
__device__ void lds_fft(double2 *buf, uint32_t n, const double2 *__restrict__ tw, uint32_t tws = 1) {
	constexpr int kMaxPerThread = MAXN / 4 / kThreads; // radix-4 butterflies per thread per pass
	const uint32_t tid = threadIdx.x;
	uint32_t ns = 1;
	const uint32_t quarter = n >> 2;
	for (; ns * 4 <= n; ns *= 4) {
		double2 v[kMaxPerThread][4];
		const uint32_t span = n / (ns * 4); // twiddle stride
#pragma unroll
		for (int t = 0; t < kMaxPerThread; ++t) {
			const uint32_t j = tid + t * kThreads;
			if (j < quarter) {
#pragma unroll
				for (int r = 0; r < 4; ++r) v[t][r] = buf[j + r * quarter];
			}
		}
		__syncthreads();
#pragma unroll
		for (int t = 0; t < kMaxPerThread; ++t) {
			const uint32_t j = tid + t * kThreads;
			if (j < quarter) radix4_write(buf, v[t], j, ns, span * tws, tw);
		}
		__syncthreads();
	}
	if (ns < n) { // final radix-2 pass
		const uint32_t half = n >> 1;
		double2 v[2 * kMaxPerThread][2];
#pragma unroll
		for (int t = 0; t < 2 * kMaxPerThread; ++t) {
			const uint32_t j = tid + t * kThreads;
			if (j < half) v[t][0] = buf[j], v[t][1] = buf[j + half];
		}
		__syncthreads();
#pragma unroll
		for (int t = 0; t < 2 * kMaxPerThread; ++t) {
			const uint32_t j = tid + t * kThreads;
			if (j < half) radix2_write(buf, v[t][0], v[t][1], j, ns, n * tws, tw);
		}
		__syncthreads();
	}
}

// Direct DFT (chains that are not a power of two and whose 2n entries fit in
// LDS): Z_k = sum_x z_x W^{kx mod n}.
__device__ void lds_dft(const double2 *buf, double2 *out_k, uint32_t k, uint32_t n, const double2 *__restrict__ tw) {
	double sr = 0, si = 0;
	uint32_t idx = 0;
	for (uint32_t x = 0; x < n; ++x) {
		const double2 w = tw[idx], z = buf[x];
		sr = fma(z.x, w.x, fma(-z.y, w.y, sr));
		si = fma(z.x, w.y, fma(z.y, w.x, si));
		idx += k;
		if (idx >= n) idx -= n;
	}
	*out_k = make_double2(sr, si);
}

// use_fft: 1 = power-of-two FFT, 2 = mixed-radix FFT through `plan`, 0 = direct DFT.
template <bool DNLS_MODE, uint32_t MAXN>
__global__ void __launch_bounds__(kThreads)
k_energy(const double2 *__restrict__ state, uint32_t n, uint32_t copies, const double2 *__restrict__ tw,
         const double *__restrict__ omega, double *__restrict__ partial, int use_fft, MixedPlan plan) {
	extern __shared__ double2 lds[]; // n (FFT) or 2n (DFT) entries
	constexpr int kMaxModes = MAXN / kThreads;
	double acc[kMaxModes];
#pragma unroll
	for (int t = 0; t < kMaxModes; ++t) acc[t] = 0.0;
	const uint32_t tid = threadIdx.x;
	// modes owned by this thread: k = tid + t*kThreads; the DFT path
	// accumulates directly into partial instead.
	for (uint32_t c = blockIdx.x; c < copies; c += gridDim.x) {
		const double2 *chain = state + size_t(c) * n;
		for (uint32_t x = tid; x < n; x += kThreads) lds[x] = chain[x];
		__syncthreads();
		if (use_fft) {
			if (use_fft == 2) mixed::lds_fft<MAXN, kThreads>(lds, n, tw, plan);
			else lds_fft<MAXN>(lds, n, tw);
#pragma unroll
			for (int t = 0; t < kMaxModes; ++t) {
				const uint32_t k = tid + t * kThreads;
				if (k < n) {
					const double2 z = lds[k];
					if constexpr (DNLS_MODE) {
						acc[t] += z.x * z.x + z.y * z.y;
					} else {
						const double2 zm = lds[k ? n - k : 0];
						const double w = omega[k];
						const double a = 1.0 + w, b = 1.0 - w;
						const double er = fma(a, z.x, -b * zm.x); // Re[(1+w)Z - (1-w)conj(Zm)]
						const double ei = fma(a, z.y, b * zm.y);  // Im[...]
						acc[t] += 0.25 * fma(er, er, ei * ei);
					}
				}
			}
			__syncthreads();
		} else {
			double2 *spec = lds + n;
			for (uint32_t k = tid; k < n; k += kThreads) lds_dft(lds, &spec[k], k, n, tw);
			__syncthreads();
			for (uint32_t k = tid; k < n; k += kThreads) {
				const double2 z = spec[k];
				double e;
				if constexpr (DNLS_MODE) {
					e = z.x * z.x + z.y * z.y;
				} else {
					const double2 zm = spec[k ? n - k : 0];
					const double w = omega[k];
					const double a = 1.0 + w, b = 1.0 - w;
					const double er = fma(a, z.x, -b * zm.x);
					const double ei = fma(a, z.y, b * zm.y);
					e = 0.25 * fma(er, er, ei * ei);
				}
				partial[size_t(blockIdx.x) * n + k] += e;
			}
			__syncthreads();
		}
	}
	if (use_fft) {
#pragma unroll
		for (int t = 0; t < kMaxModes; ++t) {
			const uint32_t k = tid + t * kThreads;
			if (k < n) partial[size_t(blockIdx.x) * n + k] = acc[t];
		}
	}
}

// Chains too long for LDS (not a power of two above 5120 sites, or a power of
// two above 8192): direct DFT with the chain read from L1/L2.  Thread k
// accumulates Z_k = sum_x z_x W^{kx} and Z_{n-k} = sum_x z_x conj(W^{kx}) in the
// same pass; all threads read the same z_x at a time.  O(n^2) per chain.
template <bool DNLS_MODE>
__global__ void __launch_bounds__(kThreads)
k_energy_gdft(const double2 *__restrict__ state, uint32_t n, uint32_t copies, const double2 *__restrict__ tw,
              const double *__restrict__ omega, double *__restrict__ partial) {
	for (uint32_t c = blockIdx.x; c < copies; c += gridDim.x) {
		const double2 *chain = state + size_t(c) * n;
		for (uint32_t k = threadIdx.x; k < n; k += kThreads) {
			double zr = 0, zi = 0, mr = 0, mi = 0;
			uint32_t idx = 0;
			for (uint32_t x = 0; x < n; ++x) {
				const double2 w = tw[idx], z = chain[x];
				zr = fma(z.x, w.x, fma(-z.y, w.y, zr)); // z W
				zi = fma(z.x, w.y, fma(z.y, w.x, zi));
				if constexpr (!DNLS_MODE) {
					mr = fma(z.x, w.x, fma(z.y, w.y, mr)); // z conj(W)
					mi = fma(z.y, w.x, fma(-z.x, w.y, mi));
				}
				idx += k;
				if (idx >= n) idx -= n;
			}
			double e;
			if constexpr (DNLS_MODE) {
				e = zr * zr + zi * zi;
			} else {
				const double w = omega[k], a = 1.0 + w, b = 1.0 - w;
				const double er = fma(a, zr, -b * mr);
				const double ei = fma(a, zi, b * mi);
				e = 0.25 * fma(er, er, ei * ei);
			}
			partial[size_t(blockIdx.x) * n + k] += e;
		}
	}
}

// Short chains (N = 8..256 sites, a power of two): N/4 threads per chain, so a
// workgroup transforms 1024/N chains at once, one radix-4 butterfly per thread
// and pass (the butterflies of lds_fft).  The chains of a workgroup are summed
// in LDS in a fixed order into one partial row per workgroup.
template <bool DNLS_MODE, uint32_t N>
__global__ void __launch_bounds__(kThreads)
k_energy_small(const double2 *__restrict__ state, uint32_t copies, const double2 *__restrict__ tw,
               const double *__restrict__ omega, double *__restrict__ partial) {
	constexpr uint32_t T = N / 4, G = kThreads / T;
	__shared__ double2 lds[G * N];
	const uint32_t tid = threadIdx.x, g = tid / T, j = tid % T;
	double2 *buf = lds + g * N;
	double acc[4] = {0.0, 0.0, 0.0, 0.0}; // modes j + t T
	const uint32_t per_round = G * gridDim.x;
	const uint32_t rounds = (copies + per_round - 1) / per_round; // the same for every thread
	for (uint32_t r = 0; r < rounds; ++r) {
		const uint32_t c = r * per_round + blockIdx.x * G + g;
		const bool live = c < copies;
#pragma unroll
		for (uint32_t t = 0; t < 4; ++t) buf[j + t * T] = live ? state[size_t(c) * N + j + t * T] : make_double2(0, 0);
		__syncthreads();
		uint32_t ns = 1;
		for (; ns * 4 <= N; ns *= 4) {
			double2 v[4];
#pragma unroll
			for (uint32_t q = 0; q < 4; ++q) v[q] = buf[j + q * T];
			__syncthreads();
			radix4_write(buf, v, j, ns, N / (ns * 4), tw);
			__syncthreads();
		}
		if (ns < N) {
			double2 v[2][2];
#pragma unroll
			for (uint32_t q = 0; q < 2; ++q) v[q][0] = buf[j + q * T], v[q][1] = buf[j + q * T + N / 2];
			__syncthreads();
#pragma unroll
			for (uint32_t q = 0; q < 2; ++q) radix2_write(buf, v[q][0], v[q][1], j + q * T, ns, N, tw);
			__syncthreads();
		}
		if (live) {
#pragma unroll
			for (uint32_t t = 0; t < 4; ++t) {
				const uint32_t k = j + t * T;
				const double2 z = buf[k];
				if constexpr (DNLS_MODE) {
					acc[t] += z.x * z.x + z.y * z.y;
				} else {
					const double2 zm = buf[(N - k) & (N - 1)];
					const double w = omega[k];
					const double a = 1.0 + w, b = 1.0 - w;
					const double er = fma(a, z.x, -b * zm.x);
					const double ei = fma(a, z.y, b * zm.y);
					acc[t] += 0.25 * fma(er, er, ei * ei);
				}
			}
		}
		__syncthreads();
	}
	double *red = reinterpret_cast<double *>(lds); // G x N doubles
#pragma unroll
	for (uint32_t t = 0; t < 4; ++t) red[g * N + j + t * T] = acc[t];
	__syncthreads();
	for (uint32_t k = tid; k < N; k += kThreads) {
		double sum = 0.0;
		for (uint32_t q = 0; q < G; ++q) sum += red[q * N + k];
		partial[size_t(blockIdx.x) * N + k] = sum;
	}
}

// Two-level FFT of chains of N = R M sites (powers of two above 8192, see
// nlc_fft_large.hpp) for a chunk of copies: column pass over z = q + i p (or
// psi) into scratch[c][r][m], forward M-point FFT of every row in LDS, then the
// mode energies from Z_k and Z_{N-k} (stored at (k mod R) M + k / R).
template <uint32_t R>
__global__ void __launch_bounds__(kThreads)
k_large_col(const double2 *__restrict__ state, uint32_t n, uint32_t M, uint32_t copies, const double2 *__restrict__ tw,
            double2 *__restrict__ scratch) {
	const uint64_t t = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
	if (t >= uint64_t(copies) * M) return;
	const uint32_t c = uint32_t(t / M), m = uint32_t(t % M);
	const double2 *x = state + size_t(c) * n;
	double2 root[R], in[R], out[R];
	small_roots<R>(root, tw, n);
#pragma unroll
	for (uint32_t j = 0; j < R; ++j) in[j] = x[size_t(j) * M + m];
	small_dft<R, false>(in, out, root);
	double2 *y = scratch + size_t(c) * n + m;
#pragma unroll
	for (uint32_t r = 0; r < R; ++r) y[size_t(r) * M] = cmul(out[r], tw[m * r]);
}

template <uint32_t M>
__global__ void __launch_bounds__(kThreads)
k_large_row(double2 *__restrict__ scratch, const double2 *__restrict__ tw, uint32_t R) {
	extern __shared__ double2 buf[];
	double2 *row = scratch + size_t(blockIdx.x) * M;
	for (uint32_t x = threadIdx.x; x < M; x += kThreads) buf[x] = row[x];
	__syncthreads();
	lds_fft<M>(buf, M, tw, R);
	for (uint32_t x = threadIdx.x; x < M; x += kThreads) row[x] = buf[x];
}

// partial[g][k] += e_k summed over the chunk's chains c = g, g + rows, ...
// (one thread per stored position p = r M + k2, k = r + R k2).
template <bool DNLS_MODE>
__global__ void __launch_bounds__(kThreads)
k_large_modes(const double2 *__restrict__ scratch, uint32_t n, uint32_t R, uint32_t M, uint32_t copies,
              const double *__restrict__ omega, double *__restrict__ partial, uint32_t rows) {
	const uint32_t p = blockIdx.x * kThreads + threadIdx.x, g = blockIdx.y;
	if (p >= n) return;
	const uint32_t k = p / M + R * (p % M);
	const uint32_t km = (n - k) & (n - 1);
	const uint32_t pm = (km % R) * M + km / R;
	double sum = 0.0;
	for (uint32_t c = g; c < copies; c += rows) {
		const double2 z = scratch[size_t(c) * n + p];
		if constexpr (DNLS_MODE) {
			sum += z.x * z.x + z.y * z.y;
		} else {
			const double2 zm = scratch[size_t(c) * n + pm];
			const double w = omega[k];
			const double a = 1.0 + w, b = 1.0 - w;
			const double er = fma(a, z.x, -b * zm.x);
			const double ei = fma(a, z.y, b * zm.y);
			sum += 0.25 * fma(er, er, ei * ei);
		}
	}
	partial[size_t(g) * n + k] += sum;
}

__global__ void k_zero(double *p, size_t count) {
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
		p[i] = 0.0;
}

// First pass of the per-mode reduction: block (x, y) sums rows
// [64y, 64y+64) of partial for modes [64x, 64x+64); 4 row slices per mode,
// combined in LDS in a fixed order.
__global__ void __launch_bounds__(256)
k_reduce_rows(const double *__restrict__ partial, uint32_t rows, uint32_t n, double *__restrict__ chunk_sums) {
	__shared__ double red[4][64];
	const uint32_t j = threadIdx.x & 63, slice = threadIdx.x >> 6;
	const uint32_t k = blockIdx.x * 64 + j;
	const uint32_t r0 = blockIdx.y * 64;
	double s = 0.0;
	if (k < n) {
#pragma unroll 4
		for (uint32_t r = r0 + slice; r < min(r0 + 64, rows); r += 4) s += partial[size_t(r) * n + k];
	}
	red[slice][j] = s;
	__syncthreads();
	if (slice == 0 && k < n)
		chunk_sums[size_t(blockIdx.y) * n + k] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
}

__global__ void k_reduce_partial(const double *__restrict__ partial, uint32_t grid, uint32_t n,
                                 const double *__restrict__ scale, double *__restrict__ e_raw) {
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	double s = 0.0;
#pragma unroll 8
	for (uint32_t g = 0; g < grid; ++g) s += partial[size_t(g) * n + k];
	e_raw[k] = scale ? s * scale[k] : s;
}

// Single workgroup: normalisation (results_entropy.cpp:11-19) and entropies
// (results_entropy.cpp:21-46) with a fixed-order tree reduction.
__device__ void results_body(const double *e_raw, uint32_t n, double norm, int a0is0, const uint32_t *__restrict__ idx,
                             uint32_t nidx, double *__restrict__ out) {
	__shared__ double red[3][kThreads];
	const uint32_t tid = threadIdx.x;
	for (uint32_t k = tid; k < n; k += kThreads) out[k] = (a0is0 && k == 0) ? 0.0 : e_raw[k] * norm;
	double s_e = 0, s_log = 0, s_elog = 0;
	const uint32_t first = a0is0 ? 1 : 0;
	const uint32_t cnt = nidx ? nidx : n - first;
	for (uint32_t j = tid; j < cnt; j += kThreads) {
		const uint32_t k = nidx ? idx[j] : first + j;
		const double e = (a0is0 && k == 0) ? 0.0 : e_raw[k] * norm;
		const double lg = log(e);
		s_e += e;
		s_log += lg;
		s_elog += e * lg;
	}
	red[0][tid] = s_e;
	red[1][tid] = s_log;
	red[2][tid] = s_elog;
	__syncthreads();
	for (uint32_t w = kThreads / 2; w > 0; w >>= 1) {
		if (tid < w) {
			red[0][tid] += red[0][tid + w];
			red[1][tid] += red[1][tid + w];
			red[2][tid] += red[2][tid + w];
		}
		__syncthreads();
	}
	if (tid == 0) {
		const double tot = red[0][0], totlog = red[1][0], totelog = red[2][0];
		const double normalization = double(cnt) / tot, lognorm = log(normalization);
		out[n] = -(totlog + double(cnt) * lognorm);
		out[n + 1] = normalization * (totelog + lognorm * tot);
	}
}

__global__ void __launch_bounds__(kThreads)
k_results(const double *__restrict__ e_raw, uint32_t n, double norm, int a0is0, const uint32_t *__restrict__ idx,
          uint32_t nidx, double *__restrict__ out) {
	results_body(e_raw, n, norm, a0is0, idx, nidx, out);
}

constexpr uint32_t kFusedMaxN = 2048;

// k_reduce_partial and k_results in one workgroup (n <= kFusedMaxN): the
// per-mode sums over the rows (same order as k_reduce_partial) go to e_raw and
// to LDS, where the entropy reduction reads them.
__global__ void __launch_bounds__(kThreads)
k_reduce_results(const double *__restrict__ partial, uint32_t rows, uint32_t n, const double *__restrict__ scale,
                 double *__restrict__ e_raw, double norm, int a0is0, const uint32_t *__restrict__ idx, uint32_t nidx,
                 double *__restrict__ out) {
	__shared__ double e[kFusedMaxN];
	__shared__ double part[kThreads];
	const uint32_t tid = threadIdx.x;
	if (2 * n <= kThreads) {
		// several threads per mode, each summing every slices-th row; the row
		// loads are independent, so they are unrolled to keep several in flight
		const uint32_t slices = kThreads / n, k = tid % n, sl = tid / n;
		double sum = 0.0;
		if (sl < slices) {
#pragma unroll 8
			for (uint32_t g = sl; g < rows; g += slices) sum += partial[size_t(g) * n + k];
		}
		part[tid] = sum;
		__syncthreads();
		if (tid < n) {
			double t = 0.0;
			for (uint32_t q = 0; q < slices; ++q) t += part[q * n + tid];
			t = scale ? t * scale[tid] : t;
			e_raw[tid] = t;
			e[tid] = t;
		}
	} else {
		for (uint32_t k = tid; k < n; k += kThreads) {
			double sum = 0.0;
#pragma unroll 8
			for (uint32_t g = 0; g < rows; ++g) sum += partial[size_t(g) * n + k];
			sum = scale ? sum * scale[k] : sum;
			e_raw[k] = sum;
			e[k] = sum;
		}
	}
	__syncthreads();
	results_body(e, n, norm, a0is0, idx, nidx, out);
}

// k_reduce_results for n <= 64 modes: the four wavefronts sum every fourth
// row (the row split of k_reduce_results at n = 64), wavefront 0 adds the four
// parts in order and does the normalisation and the entropy sums with one mode
// per lane, reduced by lane exchanges (pairs i, i + 32, then + 16, ...: the
// additions of k_reduce_results' tree, which holds at most 64 nonzero values)
// instead of eight workgroup barriers.
__global__ void __launch_bounds__(kThreads)
k_reduce_results64(const double *__restrict__ partial, uint32_t rows, uint32_t n, const double *__restrict__ scale,
                   double *__restrict__ e_raw, double norm, int a0is0, const uint32_t *__restrict__ idx,
                   uint32_t nidx, double *__restrict__ out) {
	__shared__ double part[kThreads / 64][64];
	const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	double sum = 0.0;
	if (lane < n) {
#pragma unroll 8
		for (uint32_t g = wave; g < rows; g += kThreads / 64) sum += partial[size_t(g) * n + lane];
	}
	part[wave][lane] = sum;
	__syncthreads();
	if (wave != 0) return;
	double t = 0.0;
#pragma unroll
	for (uint32_t q = 0; q < kThreads / 64; ++q) t += part[q][lane];
	if (lane < n) {
		t = scale ? t * scale[lane] : t;
		e_raw[lane] = t;
		out[lane] = (a0is0 && lane == 0) ? 0.0 : t * norm;
	}
	part[1][lane] = t; // this wavefront's LDS writes are visible to its own later reads
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
	const uint32_t first = a0is0 ? 1 : 0;
	const uint32_t cnt = nidx ? nidx : n - first;
	double s_e = 0, s_log = 0, s_elog = 0;
	if (lane < cnt) {
		const uint32_t k = nidx ? idx[lane] : first + lane;
		const double e = (a0is0 && k == 0) ? 0.0 : part[1][k] * norm;
		const double lg = log(e);
		s_e = e;
		s_log = lg;
		s_elog = e * lg;
	}
#pragma unroll
	for (int w = 32; w > 0; w >>= 1) {
		s_e += __shfl_xor(s_e, w, 64);
		s_log += __shfl_xor(s_log, w, 64);
		s_elog += __shfl_xor(s_elog, w, 64);
	}
	if (lane == 0) {
		const double normalization = double(cnt) / s_e, lognorm = log(normalization);
		out[n] = -(s_log + double(cnt) * lognorm);
		out[n + 1] = normalization * (s_elog + lognorm * s_e);
	}
}

// One workgroup per chain: total energy (and DNLS norm), fixed-order reduction.
__global__ void __launch_bounds__(kThreads)
k_chain_energy(int model, const double2 *__restrict__ state, uint32_t n, uint32_t copies, StepArgs a,
               const double *__restrict__ masses, double *__restrict__ energy, double *__restrict__ norm) {
	__shared__ double red[2][kThreads];
	const uint32_t tid = threadIdx.x;
	// grid-stride over the chains (the grid is capped so that the launch stays
	// within 2^32 work-items for any number of copies)
	for (uint32_t c = blockIdx.x; c < copies; c += gridDim.x) {
		const double2 *s = state + size_t(c) * n;
		double e = 0, nr = 0;
		for (uint32_t x = tid; x < n; x += kThreads) {
			const double2 v = s[x], w = s[x + 1 == n ? 0 : x + 1];
			if (model == DNLS) {
				const double dr = w.x - v.x, di = w.y - v.y, a2 = v.x * v.x + v.y * v.y;
				e += dr * dr + di * di + a.beta * 0.5 * a2 * a2;
				nr += a2;
			} else {
				const double q = v.x, p = v.y, d = w.x - q;
				double h = 0.5 * p * p;
				switch (model) {
				case DNKG: h += 0.5 * d * d + 0.5 * a.m * q * q + 0.25 * a.beta * q * q * q * q; break;
				case DDNKG: h += 0.5 * d * d + 0.5 * masses[x] * q * q + 0.25 * a.beta * q * q * q * q; break;
				case FPUT: h += 0.5 * d * d + a.alpha * d * d * d / 3.0 + 0.25 * a.beta * d * d * d * d; break;
				case TODA: {
					const double al = 0.5 * a.a2, inv4a2 = 1.0 / (4.0 * al * al);
					h += exp(a.a2 * d) * inv4a2 - d * a.a2i - inv4a2;
					break;
				}
				}
				e += h;
			}
		}
		red[0][tid] = e;
		red[1][tid] = nr;
		__syncthreads();
		for (uint32_t w = kThreads / 2; w > 0; w >>= 1) {
			if (tid < w) red[0][tid] += red[0][tid + w], red[1][tid] += red[1][tid + w];
			__syncthreads();
		}
		if (tid == 0) {
			energy[c] = red[0][0];
			if (norm) norm[c] = red[1][0];
		}
		__syncthreads(); // red is rewritten by the next chain
	}
}

bool is_pow2(uint32_t n) { return n >= 2 && (n & (n - 1)) == 0; }

template <bool DNLS_MODE, uint32_t MAXN>
hipError_t launch_lds_energy(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double2 *tw,
                             const double *omega, double *partial, uint32_t grid, int use_fft,
                             const MixedPlan &plan = MixedPlan{}) {
	const size_t lds = sizeof(double2) * n * (use_fft ? 1 : 2);
	if (lds > 64 * 1024) {
		hipError_t e = hipFuncSetAttribute((const void *) k_energy<DNLS_MODE, MAXN>,
		                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
		if (e != hipSuccess) return e;
	}
	hipLaunchKernelGGL((k_energy<DNLS_MODE, MAXN>), dim3(grid), dim3(kThreads), lds, st, state, n, copies, tw, omega,
	                   partial, use_fft, plan);
	return hipGetLastError();
}

bool small_path(uint32_t n) { return is_pow2(n) && n >= 8 && n <= 256; }

template <bool DNLS_MODE, uint32_t N>
hipError_t launch_small(hipStream_t st, const double2 *state, uint32_t copies, const double2 *tw, const double *omega,
                        double *partial, uint32_t grid) {
	hipLaunchKernelGGL((k_energy_small<DNLS_MODE, N>), dim3(grid), dim3(kThreads), 0, st, state, copies, tw, omega,
	                   partial);
	return hipGetLastError();
}

// Chunks of copies through the two-level FFT; scratch holds chunk * n values.
template <bool DNLS_MODE, uint32_t R, uint32_t M>
hipError_t launch_large(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double2 *tw,
                        const double *omega, double *partial, uint32_t grid, double2 *scratch, uint32_t chunk) {
	if (M * sizeof(double2) > 64 * 1024) {
		hipError_t e = hipFuncSetAttribute((const void *) k_large_row<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   int(M * sizeof(double2)));
		if (e != hipSuccess) return e;
	}
	hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, st, partial, size_t(grid) * n);
	for (uint32_t c0 = 0; c0 < copies; c0 += chunk) {
		const uint32_t cnt = std::min(chunk, copies - c0);
		const uint64_t cols = uint64_t(cnt) * M;
		hipLaunchKernelGGL(k_large_col<R>, dim3(uint32_t((cols + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
		                   state + size_t(c0) * n, n, M, cnt, tw, scratch);
		hipLaunchKernelGGL(k_large_row<M>, dim3(cnt * R), dim3(kThreads), M * sizeof(double2), st, scratch, tw, R);
		hipLaunchKernelGGL(k_large_modes<DNLS_MODE>, dim3((n + kThreads - 1) / kThreads, grid), dim3(kThreads), 0, st,
		                   scratch, n, R, M, cnt, omega, partial, grid);
	}
	return hipGetLastError();
}

template <bool DNLS_MODE>
hipError_t launch_energy_mode(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double2 *tw,
                              const double *omega, double *partial, uint32_t grid, const LongChainFft &lf) {
	double2 *scratch = lf.scratch;
	const size_t scratch_sites = lf.scratch_sites;
	if (MixedPlan plan; mixed_radix_plan(n, &plan)) // n <= 8192 with prime factors 2, 3, 5, 7
		return n <= kEnergyFftMax
		           ? launch_lds_energy<DNLS_MODE, kEnergyFftMax>(st, state, n, copies, tw, omega, partial, grid, 2, plan)
		           : launch_lds_energy<DNLS_MODE, kEnergyFftLarge>(st, state, n, copies, tw, omega, partial, grid, 2,
		                                                          plan);
	if (lf.blue && scratch && bluestein_two_level_size(n))
		return launch_energy_bluestein(st, DNLS_MODE, state, n, copies, lf.blue, omega, partial, grid, scratch,
		                               scratch_sites);
	uint32_t R = 0, M = 0;
	if (scratch && large_fft_split(n, R, M) && scratch_sites >= n) {
		const uint32_t chunk = uint32_t(std::min<size_t>(scratch_sites / n, copies));
		if (M == 8192) return launch_large<DNLS_MODE, 16, 8192>(st, state, n, copies, tw, omega, partial, grid, scratch, chunk);
		switch (R) {
		case 4: return launch_large<DNLS_MODE, 4, 4096>(st, state, n, copies, tw, omega, partial, grid, scratch, chunk);
		case 8: return launch_large<DNLS_MODE, 8, 4096>(st, state, n, copies, tw, omega, partial, grid, scratch, chunk);
		case 16: return launch_large<DNLS_MODE, 16, 4096>(st, state, n, copies, tw, omega, partial, grid, scratch, chunk);
		}
	}
	switch (small_path(n) ? n : 0) {
	case 8: return launch_small<DNLS_MODE, 8>(st, state, copies, tw, omega, partial, grid);
	case 16: return launch_small<DNLS_MODE, 16>(st, state, copies, tw, omega, partial, grid);
	case 32: return launch_small<DNLS_MODE, 32>(st, state, copies, tw, omega, partial, grid);
	case 64: return launch_small<DNLS_MODE, 64>(st, state, copies, tw, omega, partial, grid);
	case 128: return launch_small<DNLS_MODE, 128>(st, state, copies, tw, omega, partial, grid);
	case 256: return launch_small<DNLS_MODE, 256>(st, state, copies, tw, omega, partial, grid);
	default: break;
	}
	if (is_pow2(n) && n <= kEnergyFftMax)
		return launch_lds_energy<DNLS_MODE, kEnergyFftMax>(st, state, n, copies, tw, omega, partial, grid, 1);
	if (is_pow2(n) && n <= kEnergyFftLarge)
		return launch_lds_energy<DNLS_MODE, kEnergyFftLarge>(st, state, n, copies, tw, omega, partial, grid, 1);
	// the DFT paths accumulate into partial over the copies of a workgroup
	hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, st, partial, size_t(grid) * n);
	if (2 * sizeof(double2) * n <= 160 * 1024)
		return launch_lds_energy<DNLS_MODE, kEnergyFftMax>(st, state, n, copies, tw, omega, partial, grid, 0);
	hipLaunchKernelGGL(k_energy_gdft<DNLS_MODE>, dim3(grid), dim3(kThreads), 0, st, state, n, copies, tw, omega,
	                   partial);
	return hipGetLastError();
}

} // namespace

uint32_t energy_rows(uint32_t n, uint32_t copies) {
	// one partial row per workgroup: short chains 1024/n per workgroup, others
	// one at a time; long chains get fewer rows (at most 256 MiB of partials)
	if (small_path(n)) return std::min<uint32_t>((copies + 1024 / n - 1) / (1024 / n), 1024u);
	return std::min<uint32_t>({copies, 1024u, std::max<uint32_t>(1u, (1u << 25) / n)});
}

size_t energy_scratch_sites(uint32_t n, uint32_t copies) {
	uint32_t R, M, len = 0;
	if (large_fft_split(n, R, M)) len = n;
	else if (!mixed_radix_plan(n, nullptr)) len = bluestein_two_level_size(n);
	if (!len) return 0;
	// chunks of copies of at most 2^25 values (512 MiB), whole chains
	return size_t(std::min<uint64_t>(copies, std::max<uint64_t>(1, (uint64_t(1) << 25) / len))) * len;
}

hipError_t launch_energy_real(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double2 *tw,
                              const double *omega, double *partial, uint32_t grid, const LongChainFft &lf) {
	return launch_energy_mode<false>(st, state, n, copies, tw, omega, partial, grid, lf);
}

hipError_t launch_energy_dnls(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double2 *tw,
                              double *partial, uint32_t grid, const LongChainFft &lf) {
	return launch_energy_mode<true>(st, state, n, copies, tw, nullptr, partial, grid, lf);
}

uint32_t reduce_scratch_rows(uint32_t grid) { return (grid + 63) / 64; }

hipError_t launch_reduce_partial(hipStream_t st, const double *partial, uint32_t grid, uint32_t n, const double *scale,
                                 double *e_raw, double *scratch) {
	if (grid > 64) {
		const uint32_t chunks = reduce_scratch_rows(grid);
		hipLaunchKernelGGL(k_reduce_rows, dim3((n + 63) / 64, chunks), dim3(256), 0, st, partial, grid, n, scratch);
		partial = scratch;
		grid = chunks;
	}
	hipLaunchKernelGGL(k_reduce_partial, dim3((n + 255) / 256), dim3(256), 0, st, partial, grid, n, scale, e_raw);
	return hipGetLastError();
}

hipError_t launch_reduce_results(hipStream_t st, const double *partial, uint32_t grid, uint32_t n, const double *scale,
                                 double *e_raw, double *scratch, double norm, int a0is0, const uint32_t *idx,
                                 uint32_t nidx, double *out) {
	if (n > kFusedMaxN) {
		hipError_t e = launch_reduce_partial(st, partial, grid, n, scale, e_raw, scratch);
		return e != hipSuccess ? e : launch_results(st, e_raw, n, norm, a0is0, idx, nidx, out);
	}
	if (grid > 64) {
		const uint32_t chunks = reduce_scratch_rows(grid);
		hipLaunchKernelGGL(k_reduce_rows, dim3((n + 63) / 64, chunks), dim3(256), 0, st, partial, grid, n, scratch);
		partial = scratch;
		grid = chunks;
	}
	if (n <= 64)
		hipLaunchKernelGGL(k_reduce_results64, dim3(1), dim3(kThreads), 0, st, partial, grid, n, scale, e_raw, norm,
		                   a0is0, idx, nidx, out);
	else
		hipLaunchKernelGGL(k_reduce_results, dim3(1), dim3(kThreads), 0, st, partial, grid, n, scale, e_raw, norm,
		                   a0is0, idx, nidx, out);
	return hipGetLastError();
}

hipError_t launch_results(hipStream_t st, const double *e_raw, uint32_t n, double norm, int a0is0, const uint32_t *idx,
                          uint32_t nidx, double *out) {
	hipLaunchKernelGGL(k_results, dim3(1), dim3(kThreads), 0, st, e_raw, n, norm, a0is0, idx, nidx, out);
	return hipGetLastError();
}

hipError_t launch_chain_energy(hipStream_t st, int model, const double2 *state, uint32_t n, uint32_t copies,
                               const StepArgs &args, const double *masses, double *energy, double *norm) {
	hipLaunchKernelGGL(k_chain_energy, dim3(std::min<uint32_t>(copies, 1u << 20)), dim3(kThreads), 0, st, model, state,
	                   n, copies, args, masses, energy, norm);
	return hipGetLastError();
}

} // namespace nlc
