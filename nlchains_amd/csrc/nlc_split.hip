// HBM-streaming step over the ensemble-major ("split") layout, for chain
// lengths without a wave shape and for the reference's --split_kernel.
//
// Reference behaviour followed: move_split, DNKG_FPUT_Toda.cu:210-259 and
// dDNKG.cu:154-196 (one thread per copy, layout q[site][copy], p[site][copy],
// a single pass per substep that drifts q_i and kicks p_{i-1}); the
// planar<->split conversion of plane2split, common/utilities_cuda.cu:27-62.
//
// Design: consecutive threads own consecutive copies, so every access of the
// pass is a fully coalesced 512 B wavefront load/store; sites 0 and 1 and p_0,
// which the wrap-around needs at the end of each pass, stay in registers for
// the whole batch.  32 B of HBM traffic per site and substep.  The layout
// conversion is an LDS-tiled transpose (coalesced on both sides).
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

template <int MODEL>
__global__ void __launch_bounds__(256)
k_step_split(double *__restrict__ q, double *__restrict__ p, uint32_t copies, uint32_t n, uint32_t steps, StepArgs a,
             const double *__restrict__ mp2) {
	const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= copies) return;
	const size_t C = copies;
	auto mass = [&](uint32_t i) { return MODEL == DDNKG ? mp2[i] : a.m2; };
	double q0 = q[c], q1 = q[C + c], p0 = p[c];
	for (uint32_t s = 0; s < steps; ++s) {
		for (int k = 0; k < 7; ++k) {
			const double cc = (k == 0 && s) ? 2.0 * a.dc[0] : a.dc[k];
			const double d = a.dd[k];
			const double p1 = p[C + c];
			q0 = fma(cc, p0, q0);
			q1 = fma(cc, p1, q1);
			double qm2 = q0, qm1 = q1, pm1 = p1; // q_{i-2}, q_{i-1} (drifted), p_{i-1} (not yet kicked)
			for (uint32_t i = 2; i < n; ++i) {
				const double pi = p[i * C + c];
				const double qi = fma(cc, pi, q[i * C + c]);
				q[i * C + c] = qi;
				p[(i - 1) * C + c] = fma(d, force<MODEL>(qm2, qm1, qi, mass(i - 1), a), pm1);
				pm1 = pi;
				qm2 = qm1;
				qm1 = qi;
			}
			p[(n - 1) * C + c] = fma(d, force<MODEL>(qm2, qm1, q0, mass(n - 1), a), pm1);
			p0 = fma(d, force<MODEL>(qm1, q0, q1, mass(0), a), p0);
		}
	}
	const double c7 = a.dc[7];
	q0 = fma(c7, p0, q0);
	q1 = fma(c7, p[C + c], q1);
	for (uint32_t i = 2; i < n; ++i) q[i * C + c] = fma(c7, p[i * C + c], q[i * C + c]);
	q[c] = q0;
	q[C + c] = q1;
	p[c] = p0;
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

hipError_t launch_step_split(hipStream_t st, int model, double *q, double *p, uint32_t copies, uint32_t n,
                             uint32_t steps, const StepArgs &args, const double *mp2) {
	const dim3 grid((copies + 255) / 256), block(256);
	switch (model) {
	case DNKG: hipLaunchKernelGGL(k_step_split<DNKG>, grid, block, 0, st, q, p, copies, n, steps, args, mp2); break;
	case FPUT: hipLaunchKernelGGL(k_step_split<FPUT>, grid, block, 0, st, q, p, copies, n, steps, args, mp2); break;
	case TODA: hipLaunchKernelGGL(k_step_split<TODA>, grid, block, 0, st, q, p, copies, n, steps, args, mp2); break;
	case DDNKG: hipLaunchKernelGGL(k_step_split<DDNKG>, grid, block, 0, st, q, p, copies, n, steps, args, mp2); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
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
