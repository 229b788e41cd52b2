// dDNKG linear-mode energies: projection of q and p on the eigenvectors of the
// mass-weighted periodic Laplacian, squared and summed over copies.
//
// Reference behaviour followed: dDNKG.cpp:118-135 (plane2split transpose, two
// cublasDgemm with V*diag(omega) and V, make_linenergies) and dDNKG.cu:237-265
// (warp per mode: sum over copies of proj_phi^2 + proj_pi^2); CPU form
// dDNKG/cpu-implementation/dDNKG.cpp:58-72.
//
// Design: one fused kernel reads the planar shard directly (no transpose, no
// projection matrices in HBM): a workgroup computes a 64-copy x 64-mode tile of
// both projections (q.V and p.V share the V tile), staged through LDS, and its
// epilogue squares, applies omega and reduces the 64 copies into one partial
// row; the per-tile rows are reduced in a fixed order afterwards.
#include "nlc_internal.hpp"

namespace nlc {
namespace {

constexpr int kTc = 64, kTj = 64, kTi = 16;

__global__ void __launch_bounds__(256)
k_proj_energy(const double2 *__restrict__ state, uint32_t n, uint32_t copies, const double *__restrict__ V,
              const double *__restrict__ omega, double *__restrict__ partial) {
	__shared__ double2 As[kTi][kTc + 1];
	__shared__ double Bs[kTi][kTj + 1];
	__shared__ double red[16][kTj + 1];
	const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
	const uint32_t j0 = blockIdx.x * kTj, c0 = blockIdx.y * kTc;
	double aq[4][4], ap[4][4];
#pragma unroll
	for (int r = 0; r < 4; ++r)
#pragma unroll
		for (int s = 0; s < 4; ++s) aq[r][s] = ap[r][s] = 0.0;
	for (uint32_t i0 = 0; i0 < n; i0 += kTi) {
		for (int e = threadIdx.x; e < kTi * kTc; e += 256) {
			const int ii = e % kTi, cc = e / kTi;
			const uint32_t c = c0 + cc, i = i0 + ii;
			As[ii][cc] = (c < copies && i < n) ? state[size_t(c) * n + i] : make_double2(0.0, 0.0);
		}
		for (int e = threadIdx.x; e < kTi * kTj; e += 256) {
			const int ii = e % kTi, jj = e / kTi;
			const uint32_t j = j0 + jj, i = i0 + ii;
			Bs[ii][jj] = (j < n && i < n) ? V[size_t(j) * n + i] : 0.0;
		}
		__syncthreads();
#pragma unroll 4
		for (int ii = 0; ii < kTi; ++ii) {
			double2 a[4];
			double b[4];
#pragma unroll
			for (int r = 0; r < 4; ++r) a[r] = As[ii][ty * 4 + r];
#pragma unroll
			for (int s = 0; s < 4; ++s) b[s] = Bs[ii][tx * 4 + s];
#pragma unroll
			for (int r = 0; r < 4; ++r)
#pragma unroll
				for (int s = 0; s < 4; ++s) {
					aq[r][s] = fma(a[r].x, b[s], aq[r][s]);
					ap[r][s] = fma(a[r].y, b[s], ap[r][s]);
				}
		}
		__syncthreads();
	}
	// epilogue: e(c, j) = (omega_j q.V_j)^2 + (p.V_j)^2, summed over this tile's copies
#pragma unroll
	for (int s = 0; s < 4; ++s) {
		const uint32_t j = j0 + tx * 4 + s;
		const double w = j < n ? omega[j] : 0.0;
		double sum = 0.0;
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			const double wq = w * aq[r][s];
			sum += fma(wq, wq, ap[r][s] * ap[r][s]);
		}
		red[ty][tx * 4 + s] = sum;
	}
	__syncthreads();
	if (threadIdx.x < kTj) {
		double sum = 0.0;
		for (int t = 0; t < 16; ++t) sum += red[t][threadIdx.x];
		const uint32_t j = j0 + threadIdx.x;
		if (j < n) partial[size_t(blockIdx.y) * n + j] = sum;
	}
}

} // namespace

uint32_t ddnkg_partial_rows(uint32_t copies) { return (copies + kTc - 1) / kTc; }

hipError_t launch_energy_ddnkg(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double *V,
                               const double *omega, double *partial) {
	const dim3 grid((n + kTj - 1) / kTj, ddnkg_partial_rows(copies));
	hipLaunchKernelGGL(k_proj_energy, grid, dim3(256), 0, st, state, n, copies, V, omega, partial);
	return hipGetLastError();
}

} // namespace nlc
