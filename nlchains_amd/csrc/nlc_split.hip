// HBM-streaming step over the ensemble-major ("split") layout, for chains
// longer than the register- and LDS-resident kernels hold and for the
// reference's --split_kernel.
//
// Reference behaviour followed: move_split, DNKG_FPUT_Toda.cu:210-251 and
// dDNKG.cu:154-196 (layout q[site][copy], p[site][copy], one pass per substep
// that drifts q and kicks p); the planar<->split conversion of plane2split,
// common/utilities_cuda.cu:27-62.
//
// Design: the reference gives each copy one thread that walks the chain, which
// leaves the GPU nearly idle unless there are hundreds of thousands of copies.
// Here one launch per substep covers every (copy, 16-site segment) pair:
// consecutive threads take consecutive copies, so each access is a coalesced
// wavefront load or store; a thread drifts its segment and the two sites
// around it, kicks its segment and writes it to the other buffer of a
// ping-pong pair (neighbouring segments read the old values in the same
// launch).  34 B of HBM traffic per site and substep (16 B written, 16 B plus
// the 2/16 halo read).  The layout conversion is an LDS-tiled transpose.
#include "nlc_internal.hpp"

namespace nlc {
namespace {

template <int MODEL>
__device__ __forceinline__ double bond_term(double x, const StepArgs &a) {
	if constexpr (MODEL == FPUT) return x * fma(x, fma(a.beta, x, a.alpha), 1.0);
	else return a.a2i * exp(a.a2 * x);
}

// Force on the centre site, same expressions as the wave-resident kernels.
template <int MODEL>
__device__ __forceinline__ double force(double l, double c, double r, double mass, const StepArgs &a) {
	if constexpr (MODEL == DNKG || MODEL == DDNKG) {
		const double t = fma(a.beta, c * c, mass);
		return fma(-c, t, l + r);
	} else {
		return bond_term<MODEL>(r - c, a) - bond_term<MODEL>(c - l, a);
	}
}

constexpr uint32_t kSeg = 16; // sites per thread and substep

// One substep: q += cc p over sites [i0 - 1, i0 + kSeg] (periodic), then, with
// KICK, p += d F(q) on [i0, i0 + kSeg); results to (qo, po).
template <int MODEL, bool KICK>
__global__ void __launch_bounds__(256)
k_step_split(const double *__restrict__ qi, const double *__restrict__ pi, double *__restrict__ qo,
             double *__restrict__ po, uint32_t copies, uint32_t n, double cc, double d, StepArgs a,
             const double *__restrict__ mp2) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint32_t nseg = (n + kSeg - 1) / kSeg;
	if (t >= uint64_t(copies) * nseg) return;
	const size_t C = copies;
	const uint32_t c = uint32_t(t % copies), i0 = uint32_t(t / copies) * kSeg;
	const uint32_t cnt = min(kSeg, n - i0);
	const uint32_t il = i0 ? i0 - 1 : n - 1;
	double ql = fma(cc, pi[il * C + c], qi[il * C + c]);
	double pc = pi[i0 * C + c];
	double qc = fma(cc, pc, qi[i0 * C + c]);
#pragma unroll
	for (uint32_t j = 0; j < kSeg; ++j) {
		if (j < cnt) {
			const uint32_t i = i0 + j, inext = (i + 1 == n) ? 0 : i + 1;
			const double pn = pi[inext * C + c];
			const double qn = fma(cc, pn, qi[inext * C + c]);
			qo[i * C + c] = qc;
			if constexpr (KICK) {
				const double mass = (MODEL == DDNKG) ? mp2[i] : a.m2;
				po[i * C + c] = fma(d, force<MODEL>(ql, qc, qn, mass, a), pc);
			} else {
				po[i * C + c] = pc;
			}
			ql = qc;
			qc = qn;
			pc = pn;
		}
	}
}

template <int MODEL>
hipError_t launch_split_model(hipStream_t st, double *q[2], double *p[2], int &cur, uint32_t copies, uint32_t n,
                              uint32_t steps, const StepArgs &a, const double *mp2) {
	const uint64_t threads = uint64_t(copies) * ((n + kSeg - 1) / kSeg);
	const dim3 grid(uint32_t((threads + 255) / 256)), block(256);
	for (uint32_t s = 0; s < steps; ++s)
		for (int k = 0; k < 7; ++k) {
			// the c7 drift of a step merged with the c0 drift of the next
			const double cc = (k == 0 && s) ? 2.0 * a.dc[0] : a.dc[k];
			hipLaunchKernelGGL((k_step_split<MODEL, true>), grid, block, 0, st, q[cur], p[cur], q[cur ^ 1],
			                   p[cur ^ 1], copies, n, cc, a.dd[k], a, mp2);
			cur ^= 1;
		}
	hipLaunchKernelGGL((k_step_split<MODEL, false>), grid, block, 0, st, q[cur], p[cur], q[cur ^ 1], p[cur ^ 1],
	                   copies, n, a.dc[7], 0.0, a, mp2);
	cur ^= 1;
	return hipGetLastError();
}

constexpr int kTile = 32;

__global__ void __launch_bounds__(256)
k_planar_to_split(const double2 *__restrict__ planar, double *__restrict__ q, double *__restrict__ p, uint32_t copies,
                  uint32_t n) {
	__shared__ double2 tile[kTile][kTile + 1];
	const uint32_t c0 = blockIdx.x * kTile, x0 = blockIdx.y * kTile;
	for (uint32_t j = threadIdx.y; j < kTile; j += blockDim.y) {
		const uint32_t c = c0 + j, x = x0 + threadIdx.x;
		if (c < copies && x < n) tile[j][threadIdx.x] = planar[size_t(c) * n + x];
	}
	__syncthreads();
	for (uint32_t j = threadIdx.y; j < kTile; j += blockDim.y) {
		const uint32_t x = x0 + j, c = c0 + threadIdx.x;
		if (c < copies && x < n) {
			const double2 v = tile[threadIdx.x][j];
			q[size_t(x) * copies + c] = v.x;
			p[size_t(x) * copies + c] = v.y;
		}
	}
}

__global__ void __launch_bounds__(256)
k_split_to_planar(const double *__restrict__ q, const double *__restrict__ p, double2 *__restrict__ planar,
                  uint32_t copies, uint32_t n) {
	__shared__ double2 tile[kTile][kTile + 1];
	const uint32_t c0 = blockIdx.x * kTile, x0 = blockIdx.y * kTile;
	for (uint32_t j = threadIdx.y; j < kTile; j += blockDim.y) {
		const uint32_t x = x0 + j, c = c0 + threadIdx.x;
		if (c < copies && x < n) tile[threadIdx.x][j] = make_double2(q[size_t(x) * copies + c], p[size_t(x) * copies + c]);
	}
	__syncthreads();
	for (uint32_t j = threadIdx.y; j < kTile; j += blockDim.y) {
		const uint32_t c = c0 + j, x = x0 + threadIdx.x;
		if (c < copies && x < n) planar[size_t(c) * n + x] = tile[j][threadIdx.x];
	}
}

} // namespace

hipError_t launch_step_split(hipStream_t st, int model, double *q[2], double *p[2], int &cur, uint32_t copies,
                             uint32_t n, uint32_t steps, const StepArgs &args, const double *mp2) {
	switch (model) {
	case DNKG: return launch_split_model<DNKG>(st, q, p, cur, copies, n, steps, args, mp2);
	case FPUT: return launch_split_model<FPUT>(st, q, p, cur, copies, n, steps, args, mp2);
	case TODA: return launch_split_model<TODA>(st, q, p, cur, copies, n, steps, args, mp2);
	case DDNKG: return launch_split_model<DDNKG>(st, q, p, cur, copies, n, steps, args, mp2);
	default: return hipErrorInvalidValue;
	}
}

hipError_t launch_planar_to_split(hipStream_t st, const double2 *planar, double *q, double *p, uint32_t copies,
                                  uint32_t n) {
	const dim3 grid((copies + kTile - 1) / kTile, (n + kTile - 1) / kTile), block(kTile, 8);
	hipLaunchKernelGGL(k_planar_to_split, grid, block, 0, st, planar, q, p, copies, n);
	return hipGetLastError();
}

hipError_t launch_split_to_planar(hipStream_t st, const double *q, const double *p, double2 *planar, uint32_t copies,
                                  uint32_t n) {
	const dim3 grid((copies + kTile - 1) / kTile, (n + kTile - 1) / kTile), block(kTile, 8);
	hipLaunchKernelGGL(k_split_to_planar, grid, block, 0, st, q, p, planar, copies, n);
	return hipGetLastError();
}

} // namespace nlc
