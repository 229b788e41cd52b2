// Workgroup-resident Yoshida-6 step kernels for long chains (gfx950, wave64).
//
// Included once per model by nlc_group_<model>.hip with NLC_STEP_MODEL and
// NLC_GROUP_LOOKUP defined.
//
// Reference behaviour followed: the same step as move_chain_in_warp
// (DNKG_FPUT_Toda.cu:74-139, dDNKG.cu:24-131); the reference runs chains that
// do not fit its warp kernel through the HBM split kernel (move_split,
// DNKG_FPUT_Toda.cu:210-251), which this kernel replaces for n <= 16384.
//
// Design:
//  * one workgroup of W wavefronts (W = 2..8, runtime) per chain; lane g of the
//    workgroup (g = 64 w + lane) holds S consecutive sites in VGPRs for the
//    whole batch, or S - 1 and a dummy last slot when n < 64 W S (the lanes
//    with S sites come first);
//  * halos: DPP wave_ror / wave_rol inside a wavefront; at the wavefront ends
//    lanes 63 and 0 exchange through a double-buffered LDS ring on every
//    other kick (one workgroup barrier per two kicks; the buffer alternates,
//    so a buffer is rewritten only after every wave has passed the barrier
//    that follows its last read).  An exchange carries the edge q, the q next
//    to it and the edge p; the receiving lane kicks and drifts the
//    neighbour's edge site itself with the same expressions, which gives the
//    halo of the following kick bit for bit;
//  * FPUT / Toda evaluate the lane's right boundary bond themselves (one bond
//    term more per lane and kick) instead of exchanging bond terms;
//  * HBM is touched once per batch: 16 B/site in, 16 B/site out.
#include <algorithm>

#include "nlc_internal.hpp"
// after the HIP runtime header: nlc_exp.hpp uses its __device__ qualifiers
#include "nlc_exp.hpp"

namespace nlc {
namespace {

#include "nlc_step_common.inl"

constexpr int kGroupMaxWaves = 8;

// The kick loop and the predicted halos evaluate the force with the same
// site_force / bond_force expressions, so a prediction is bit-identical.
template <int MODEL, int S, bool EXACT, int FORM>
struct GroupChain {
	double q[S], p[S];
	double mp[MODEL == DDNKG ? S : 1];
	double m_l, m_r; // dDNKG: m_x + 2 of the neighbouring sites across the lane edges
	bool full;
	uint32_t lane, wave, waves;
	const double *etab;
	// [buffer][0: sent to the next wave, 1: sent to the previous wave][wave][3]
	double (*ring)[2][kGroupMaxWaves][3];
	uint32_t buf;
	bool exchange;        // this kick exchanges through LDS; the next one uses the predictions
	double pred_l, pred_r; // lane 0 / lane 63: the neighbour's edge q after this kick and drift

	__device__ __forceinline__ double last_q() const { return (EXACT || full) ? q[S - 1] : q[S - 2]; }
	__device__ __forceinline__ double before_last_q() const { return (EXACT || full) ? q[S - 2] : q[S - 3]; }
	__device__ __forceinline__ double last_p() const { return (EXACT || full) ? p[S - 1] : p[S - 2]; }

	// Left and right halo of the lane: the last real q of lane g-1 and the first
	// q of lane g+1 (ring over the 64 W lanes of the chain).  Inside a
	// wavefront by DPP; across wavefronts through LDS on exchange kicks (edge
	// q, the q next to it and the edge p, from which the receiving lane
	// predicts the neighbour's edge q for the following kick), else from the
	// prediction.
	__device__ __forceinline__ void halos(double &hl, double &hr, double (&xl)[3], double (&xr)[3]) {
		hl = dpp_f64<kWaveRor1>(last_q());
		hr = dpp_f64<kWaveRol1>(q[0]);
		if (exchange) {
			if (lane == 63) {
				ring[buf][0][wave][0] = before_last_q();
				ring[buf][0][wave][1] = last_q();
				ring[buf][0][wave][2] = last_p();
			}
			if (lane == 0) {
				ring[buf][1][wave][0] = q[0];
				ring[buf][1][wave][1] = q[1];
				ring[buf][1][wave][2] = p[0];
			}
			__syncthreads();
			const uint32_t wl = wave == 0 ? waves - 1 : wave - 1, wr = wave + 1 == waves ? 0 : wave + 1;
#pragma unroll
			for (int i = 0; i < 3; ++i) xl[i] = ring[buf][0][wl][i], xr[i] = ring[buf][1][wr][i];
			if (lane == 0) hl = xl[1];
			if (lane == 63) hr = xr[0];
			buf ^= 1;
		} else {
			if (lane == 0) hl = pred_l;
			if (lane == 63) hr = pred_r;
		}
	}

	__device__ __forceinline__ void drift(double c) {
#pragma unroll
		for (int j = 0; j < S; ++j) q[j] = fma(c, p[j], q[j]);
	}

	// Kick with weight d; c_next is the drift that follows (used by the
	// predictions of an exchange kick).
	__device__ __forceinline__ void kick(double d, double c_next, const StepArgs &a) {
		double hl, hr, xl[3], xr[3];
		halos(hl, hr, xl, xr);
		if constexpr (MODEL == DNKG || MODEL == DDNKG) {
			if (exchange) {
				// the neighbours' edge sites, kicked and drifted as their own lanes do
				const double ml = (MODEL == DDNKG) ? m_l : a.m2, mr = (MODEL == DDNKG) ? m_r : a.m2;
				pred_l = fma(c_next, fma(d, site_force<FORM>(xl[0], xl[1], q[0], ml, a), xl[2]), xl[1]);
				pred_r = fma(c_next, fma(d, site_force<FORM>(last_q(), xr[0], xr[1], mr, a), xr[2]), xr[0]);
			}
			if (!EXACT && !full) q[S - 1] = hr;
#pragma unroll
			for (int j = 0; j < S; ++j) {
				const double lft = (j == 0) ? hl : q[j - 1];
				const double rgt = (j == S - 1) ? hr : q[j + 1];
				const double mass = (MODEL == DDNKG) ? mp[j] : a.m2;
				p[j] = fma(d, site_force<FORM>(lft, q[j], rgt, mass, a), p[j]);
			}
			if (!EXACT && !full) p[S - 1] = 0.0;
		} else {
			const double w = (MODEL == TODA) ? d * a.a2i : d;
			const double last = last_q();
			const double bs = bond_force<MODEL, S, FORM>(hr - last, a, etab); // right bond of the last real site
			double bl = bond_force<MODEL, S, FORM>(q[0] - hl, a, etab);
			if (exchange) {
				const double bll = bond_force<MODEL, S, FORM>(xl[1] - xl[0], a, etab); // left bond of the left neighbour
				const double brr = bond_force<MODEL, S, FORM>(xr[1] - xr[0], a, etab); // right bond of the right neighbour
				pred_l = fma(c_next, fma(w, bl - bll, xl[2]), xl[1]);
				pred_r = fma(c_next, fma(w, brr - bs, xr[2]), xr[0]);
			}
#pragma unroll
			for (int j = 0; j < S; ++j) {
				double br;
				if (j == S - 1) br = bs;
				else if (!EXACT && j == S - 2) br = full ? bond_force<MODEL, S, FORM>(q[j + 1] - q[j], a, etab) : bs;
				else br = bond_force<MODEL, S, FORM>(q[j + 1] - q[j], a, etab);
				p[j] = fma(w, br - bl, p[j]);
				bl = br;
			}
			if (!EXACT && !full) p[S - 1] = 0.0;
		}
		exchange = !exchange;
	}
};

// One chain per workgroup of 64 W threads, W = blockDim.x / 64.  Every thread
// of the workgroup runs the same number of kicks, so every barrier is reached
// by all of them.
template <int MODEL, int S, bool EXACT, int FORM>
__global__ void __launch_bounds__(64 * kGroupMaxWaves)
k_step_group(double2 *__restrict__ state, uint32_t n, uint32_t steps, StepArgs a, const double *__restrict__ mp2g) {
	__shared__ double etab[MODEL == TODA ? kExpTabSize : 1];
	__shared__ double ring[2][2][kGroupMaxWaves][3];
	load_exp_table<MODEL>(etab);
	const uint32_t g = threadIdx.x, G = blockDim.x;

	GroupChain<MODEL, S, EXACT, FORM> ch;
	ch.lane = g & 63;
	ch.wave = g >> 6;
	ch.waves = G >> 6;
	ch.etab = etab;
	ch.ring = ring;
	ch.buf = 0;
	ch.exchange = true;
	ch.pred_l = ch.pred_r = 0.0;
	const uint32_t full_lanes = EXACT ? G : n - uint32_t(S - 1) * G;
	ch.full = EXACT || g < full_lanes;
	const uint32_t off = ch.full ? g * S : full_lanes * S + (g - full_lanes) * (S - 1);
	double2 *base = state + size_t(blockIdx.x) * n + off;

#pragma unroll
	for (int j = 0; j < S; ++j) {
		const bool real_site = EXACT || j < S - 1 || ch.full;
		double2 v = make_double2(0.0, 0.0);
		if (real_site) v = to_registers<MODEL, FORM, S>(base[j], a);
		ch.q[j] = v.x;
		ch.p[j] = v.y;
		if constexpr (MODEL == DDNKG) ch.mp[j] = real_site ? mp2g[off + j] : 0.0;
	}
	if constexpr (MODEL == DDNKG) {
		ch.m_l = mp2g[off == 0 ? n - 1 : off - 1];
		const uint32_t end = off + (ch.full ? S : S - 1);
		ch.m_r = mp2g[end == n ? 0 : end];
	} else {
		ch.m_l = ch.m_r = 0.0;
	}

	const double cs = (MODEL == TODA) ? a.scale * toda_unit<MODEL, S>() : 1.0;
	const double dc0 = cs * a.dc[0];
	ch.drift(dc0);
	for (uint32_t s = 0; s < steps; ++s) {
		const double c_last = (s + 1 < steps) ? 2.0 * dc0 : cs * a.dc[7];
		if constexpr (MODEL == TODA) {
#pragma unroll 1
			for (int k = 0; k < 7; ++k) {
				const double c = k < 6 ? cs * a.dc[k + 1] : c_last;
				ch.kick(a.dd[k], c, a);
				ch.drift(c);
			}
		} else {
#pragma unroll
			for (int k = 0; k < 7; ++k) {
				const double c = k < 6 ? a.dc[k + 1] : c_last;
				ch.kick(a.dd[k], c, a);
				ch.drift(c);
			}
		}
	}

#pragma unroll
	for (int j = 0; j < S; ++j)
		if (EXACT || j < S - 1 || ch.full) base[j] = from_registers<MODEL, FORM, S>(ch.q[j], ch.p[j], a);
}

// W = ceil(n / (64 S)) wavefronts; the sites must fill every lane with S or
// S - 1 of them: 64 W (S - 1) < n <= 64 W S (n = 64 W S for EXACT).
template <int MODEL, int S, bool EXACT>
hipError_t launch_group(hipStream_t st, double2 *state, uint32_t copies, uint32_t n, uint32_t steps,
                        const StepArgs &a, const double *mp2) {
	const uint32_t G = 64u * ((n + 64u * S - 1) / (64u * S));
	if (G < 128 || G > 64u * kGroupMaxWaves || uint64_t(G) * (S - 1) >= n || uint64_t(G) * S < n ||
	    (EXACT && uint64_t(G) * S != n))
		return hipErrorInvalidValue;
	if constexpr (MODEL != TODA)
		if (a.form == kFormCubic) {
			hipLaunchKernelGGL((k_step_group<MODEL, S, EXACT, kFormCubic>), dim3(copies), dim3(G), 0, st, state, n,
			                   steps, a, mp2);
			return hipGetLastError();
		}
	if constexpr (MODEL == FPUT)
		if (a.form == kFormQuad) {
			hipLaunchKernelGGL((k_step_group<MODEL, S, EXACT, kFormQuad>), dim3(copies), dim3(G), 0, st, state, n,
			                   steps, a, mp2);
			return hipGetLastError();
		}
	hipLaunchKernelGGL((k_step_group<MODEL, S, EXACT, kFormPlain>), dim3(copies), dim3(G), 0, st, state, n, steps, a,
	                   mp2);
	return hipGetLastError();
}

} // namespace

// Shapes instantiated: S = 8, 16, 32 sites per lane with n = 64 W S exactly,
// and S = 9..32 for any n with 64 W (S - 1) < n <= 64 W S.
StepLauncher NLC_GROUP_LOOKUP(int S, bool exact) {
#define NLC_G(SS, EX)                                                                                              \
	if (S == SS && exact == EX) return &launch_group<NLC_STEP_MODEL, SS, EX>;
	NLC_G(8, true) NLC_G(16, true) NLC_G(32, true)
	NLC_G(9, false) NLC_G(10, false) NLC_G(11, false) NLC_G(12, false) NLC_G(13, false) NLC_G(14, false)
	NLC_G(15, false) NLC_G(16, false) NLC_G(17, false) NLC_G(18, false) NLC_G(19, false) NLC_G(20, false)
	NLC_G(21, false) NLC_G(22, false) NLC_G(23, false) NLC_G(24, false) NLC_G(25, false) NLC_G(26, false)
	NLC_G(27, false) NLC_G(28, false) NLC_G(29, false) NLC_G(30, false) NLC_G(31, false) NLC_G(32, false)
#undef NLC_G
	return nullptr;
}

} // namespace nlc
