// Wave-resident Yoshida-6 step kernels for the real models (gfx950, wave64).
//
// Included once per model by nlc_step_<model>.hip with NLC_STEP_MODEL and
// NLC_STEP_LOOKUP defined, and by nlc_step16_<model>.hip with
// NLC_STEP_LOOKUP16 (the 16-lane ragged shapes), so they compile in parallel.
//
// Reference behaviour followed: DNKG_FPUT_Toda.cu:74-139 (move_chain_in_warp),
// :148-182 (move_chain_in_thread), dDNKG.cu:24-131; forces DNKG_FPUT_Toda.cu:20-66
// and dDNKG.cu:17-19.
//
// Design (MI355X-first, not a translation of the warp kernel):
//  * a chain occupies L lanes of a 64-wide wavefront, S consecutive sites per
//    lane held in VGPRs for the whole batch of steps; 64/L chains share a
//    wavefront, so short chains do not waste lanes (L = 1 is thread-per-chain);
//  * nearest-neighbour halos move between lanes with DPP (wave_ror/wave_rol for
//    L = 64, row_ror for L = 16, quad_perm for L = 4 and 2) or ds_bpermute
//    (L = 32, 8), never through LDS memory;
//  * the kick of DNKG/dDNKG is 5 FP64 ops per site (4 when beta > 0: the
//    state is held as sqrt(beta) q, sqrt(beta) p, which makes the cubic
//    coefficient 1), the drift 1 FMA;
//    FPUT and Toda exchange the bond force of the lane's first bond instead of
//    recomputing it, so each bond term (a cubic or an FP64 exp) is evaluated
//    once per substep; Toda works on v = (256 alpha/ln2) q with the 2^(x/128) of
//    nlc_exp.hpp (with 25+ sites per lane on v/32 and its 2^x);
//  * the 7 substeps are unrolled with their weights in SGPRs; the c7 drift of a
//    step is fused with the c0 drift of the next (the reference's 2*c0 merge).
//  * HBM is touched once per batch: 16 B/site in, 16 B/site out.
#include <algorithm>
#include <utility>

#include "nlc_internal.hpp"
// after the HIP runtime header: nlc_exp.hpp uses its __device__ qualifiers
#include "nlc_exp.hpp"

namespace nlc {
namespace {

#include "nlc_step_common.inl"

template <int MODEL, int L, int S, bool EXACT, int FORM>
struct Chain {
	static constexpr int kSites = S;
	double q[S], p[S];
	double mp[MODEL == DDNKG ? S : 1];
	bool full; // lane holds S real sites (else S-1 and a dummy last slot)
	int lane;
	const double *etab; // Toda: exp table in LDS

	__device__ __forceinline__ void drift(double c) {
#pragma unroll
		for (int j = 0; j < S; ++j) q[j] = fma(c, p[j], q[j]);
	}

	__device__ __forceinline__ void kick(double d, const StepArgs &a) {
		// value of this lane's last real site, which the right neighbour needs
		const double send = (EXACT || full) ? q[S - 1] : q[S > 1 ? S - 2 : 0];
		const double hl = from_left<L>(send, lane);
		if constexpr (MODEL == DNKG || MODEL == DDNKG) {
			const double hr = from_right<L>(q[0], lane);
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
			// bond terms rolled through the chain: only two are live at a time
			const double w = (MODEL == TODA) ? d * a.a2i : d;
			const double b0 = bond_force<MODEL, S, FORM>(q[0] - hl, a, etab);
			const double bs = from_right<L>(b0, lane); // right bond of the lane's last site
			double bl = b0;
#pragma unroll
			for (int j = 0; j < S; ++j) {
				double br;
				if (j == S - 1) br = bs;
				else if (!EXACT && j == S - 2) br = full ? bond_force<MODEL, S, FORM>(q[j + 1] - q[j], a, etab) : bs;
				else br = bond_force<MODEL, S, FORM>(q[j + 1] - q[j], a, etab);
				p[j] = fma(w, br - bl, p[j]);
				bl = br;
			}
			if (!EXACT && !full) p[S - 1] = 0.0; // the dummy slot of a short lane
		}
	}
};

// `steps` Yoshida steps of the chain part held by ch: the c7 drift of a step
// fused with the c0 drift of the next (the reference's 2 c0 merge).  Toda
// drifts v = (256 alpha/ln2) q (times toda_unit): its drift weights carry that
// factor.
template <int MODEL, class C>
__device__ __forceinline__ void step_loop(C &ch, uint32_t steps, const StepArgs &a) {
	const double cs = (MODEL == TODA) ? a.scale * toda_unit<MODEL, C::kSites>() : 1.0;
	const double dc0 = cs * a.dc[0];
	ch.drift(dc0);
	for (uint32_t s = 0; s < steps; ++s) {
		const double c_last = (s + 1 < steps) ? 2.0 * dc0 : cs * a.dc[7];
		if constexpr (MODEL == TODA) {
			// exp is long: keep the substep loop rolled to bound code size
#pragma unroll 1
			for (int k = 0; k < 7; ++k) {
				ch.kick(a.dd[k], a);
				ch.drift(k < 6 ? cs * a.dc[k + 1] : c_last);
			}
		} else {
#pragma unroll
			for (int k = 0; k < 7; ++k) {
				ch.kick(a.dd[k], a);
				ch.drift(k < 6 ? a.dc[k + 1] : c_last);
			}
		}
	}
}

// One wavefront integrates `live_chains` (<= 64/L) consecutive chains of
// length n starting at chain0 for `steps` steps.  All 64 lanes take part in
// the halo exchange even when the wave is only partly filled.  Toda keeps
// v = (256 alpha/ln2) q in registers (the exponent of its bond term is then a
// plain difference in 1/128 octaves; v/128 and octaves from 25 sites per lane
// on) and drifts v with the same factor times c; q = v / (256 alpha/ln2) is
// stored back.
// The scaled force forms hold sigma q and sigma p.
template <int MODEL, int L, int S, bool EXACT, int FORM>
__device__ __forceinline__ void run_wave(double2 *__restrict__ chain0, uint32_t n, uint32_t live_chains, uint32_t steps,
                                         const StepArgs &a, const double *__restrict__ mp2g, const double *etab) {
	const int lane = threadIdx.x & 63;
	const int l = lane & (L - 1);
	const uint32_t sub = uint32_t(lane / L);
	const bool live = sub < live_chains;

	Chain<MODEL, L, S, EXACT, FORM> ch;
	ch.lane = lane;
	ch.etab = etab;
	const int full_lanes = EXACT ? L : int(n) - (S - 1) * L;
	ch.full = EXACT || l < full_lanes;
	const uint32_t off = ch.full ? uint32_t(l * S) : uint32_t(full_lanes * S + (l - full_lanes) * (S - 1));
	double2 *base = chain0 + size_t(live ? sub : 0) * n + off;

#pragma unroll
	for (int j = 0; j < S; ++j) {
		const bool real_site = EXACT || j < S - 1 || ch.full;
		double2 v = make_double2(0.0, 0.0);
		if (live && real_site) v = to_registers<MODEL, FORM, S>(base[j], a);
		ch.q[j] = v.x;
		ch.p[j] = v.y;
		if constexpr (MODEL == DDNKG) ch.mp[j] = real_site ? mp2g[off + j] : 0.0;
	}

	step_loop<MODEL>(ch, steps, a);

	if (live) {
#pragma unroll
		for (int j = 0; j < S; ++j)
			if (EXACT || j < S - 1 || ch.full)
				base[j] = from_registers<MODEL, FORM, S>(ch.q[j], ch.p[j], a);
	}
}

// FP32-compensated chain state (FPUT and Toda, n = L S): q and p as FP32 pairs
// (h, l) whose sum is the value.  Every increment of a drift or kick is formed
// in FP32 and added with Kahan's compensated summation, l carrying what h
// could not hold, so the state keeps about 2^-48 of its value while the
// increments are rounded to FP32.  FPUT evaluates its bond force in FP32 on
// x = (h_i - h_{i-1}) + (l_i - l_{i-1}); Toda forms the bond exponent (from
// the pairs, in FP64) and the exponential in FP64 (the exponential amplifies
// an exponent error), and only
// its kick increment is rounded to FP32.  Drift and kick weights are rounded
// to FP32.  HBM holds the FP64 state; loads split it, stores sum the pair.
template <int MODEL, int L, int S, int FORM>
struct Chain32 {
	static_assert(MODEL == FPUT || MODEL == TODA, "FP32-compensated kernels: FPUT and Toda");
	static constexpr int kSites = S;
	float qh[S], ql[S], ph[S], pl[S];
	int lane;
	const double *etab; // Toda: exp table in LDS

	// (h, l) += d
	__device__ __forceinline__ static void add(float &h, float &l, float d) {
		const float y = d + l;
		const float s = h + y;
		l = y - (s - h);
		h = s;
	}

	__device__ __forceinline__ void drift(double c) {
		const float cf = float(c);
#pragma unroll
		for (int j = 0; j < S; ++j) add(qh[j], ql[j], cf * ph[j]);
	}

	__device__ __forceinline__ static float fput_bond(float x, const StepArgs &a) {
		if constexpr (FORM == kFormCubic) return fmaf(x * x, x, x);
		else if constexpr (FORM == kFormQuad) return fmaf(x, x, x);
		else return x * fmaf(x, fmaf(float(a.beta), x, float(a.alpha)), 1.0f);
	}

	__device__ __forceinline__ double toda_bond(float h1, float l1, float h0, float l0) const {
		// each difference of two FP32 values is exact in FP64 (unless their
		// magnitudes differ by more than 2^29); one rounding in the sum
		const double x = (double(h1) - double(h0)) + (double(l1) - double(l0));
		if constexpr (S >= 25) return exp2_poly(x);
		else return exp2_128_tab(x, etab);
	}

	__device__ __forceinline__ void kick(double d, const StepArgs &a) {
		const float lh = from_left<L>(qh[S - 1], lane), ll = from_left<L>(ql[S - 1], lane);
		if constexpr (MODEL == FPUT) {
			const float w = float(d);
			const float b0 = fput_bond((qh[0] - lh) + (ql[0] - ll), a);
			const float bs = from_right<L>(b0, lane); // right bond of the lane's last site
			float bl = b0;
#pragma unroll
			for (int j = 0; j < S; ++j) {
				const float br = (j == S - 1) ? bs : fput_bond((qh[j + 1] - qh[j]) + (ql[j + 1] - ql[j]), a);
				add(ph[j], pl[j], w * (br - bl));
				bl = br;
			}
		} else {
			const double w = d * a.a2i;
			const double b0 = toda_bond(qh[0], ql[0], lh, ll);
			const double bs = from_right<L>(b0, lane);
			double bl = b0;
#pragma unroll
			for (int j = 0; j < S; ++j) {
				const double br = (j == S - 1) ? bs : toda_bond(qh[j + 1], ql[j + 1], qh[j], ql[j]);
				add(ph[j], pl[j], float(w * (br - bl)));
				bl = br;
			}
		}
	}
};

__device__ __forceinline__ void split_f32(double v, float &h, float &l) {
	h = float(v);
	l = float(v - double(h));
}

// run_wave for the FP32-compensated state (exact shapes only).
template <int MODEL, int L, int S, int FORM>
__device__ __forceinline__ void run_wave32(double2 *__restrict__ chain0, uint32_t n, uint32_t live_chains,
                                           uint32_t steps, const StepArgs &a, const double *etab) {
	const int lane = threadIdx.x & 63;
	const uint32_t sub = uint32_t(lane / L);
	const bool live = sub < live_chains;
	Chain32<MODEL, L, S, FORM> ch;
	ch.lane = lane;
	ch.etab = etab;
	double2 *base = chain0 + size_t(live ? sub : 0) * n + uint32_t((lane & (L - 1)) * S);
#pragma unroll
	for (int j = 0; j < S; ++j) {
		const double2 v = live ? to_registers<MODEL, FORM, S>(base[j], a) : make_double2(0.0, 0.0);
		split_f32(v.x, ch.qh[j], ch.ql[j]);
		split_f32(v.y, ch.ph[j], ch.pl[j]);
	}
	step_loop<MODEL>(ch, steps, a);
	if (live) {
#pragma unroll
		for (int j = 0; j < S; ++j)
			base[j] = from_registers<MODEL, FORM, S>(double(ch.qh[j]) + double(ch.ql[j]),
			                                         double(ch.ph[j]) + double(ch.pl[j]), a);
	}
}

template <int MODEL, int L, int S, int FORM>
__global__ void __launch_bounds__(256)
k_step_wave32(double2 *__restrict__ state, uint32_t copies, uint32_t n, uint32_t steps, StepArgs a) {
	constexpr uint32_t CPW = 64 / L;
	__shared__ double etab[MODEL == TODA ? kExpTabSize : 1];
	load_exp_table<MODEL>(etab);
	const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const uint32_t first_chain = wave * CPW;
	if (first_chain >= copies) return; // whole wavefront idle: uniform exit
	run_wave32<MODEL, L, S, FORM>(state + size_t(first_chain) * n, n, min(CPW, copies - first_chain), steps, a, etab);
}

template <int MODEL, int L, int S, int FORM>
hipError_t launch_wave32_form(hipStream_t st, double2 *state, uint32_t copies, uint32_t n, uint32_t steps,
                              const StepArgs &a) {
	if (n != uint32_t(L * S)) return hipErrorInvalidValue;
	constexpr uint32_t CPW = 64 / L;
	const uint64_t waves = (uint64_t(copies) + CPW - 1) / CPW;
	const uint64_t blocks = (waves + 3) / 4;
	hipLaunchKernelGGL((k_step_wave32<MODEL, L, S, FORM>), dim3(uint32_t(blocks)), dim3(256), 0, st, state, copies,
	                   n, steps, a);
	return hipGetLastError();
}

template <int MODEL, int L, int S>
hipError_t launch_wave32(hipStream_t st, double2 *state, uint32_t copies, uint32_t n, uint32_t steps,
                         const StepArgs &a, const double *) {
	if constexpr (MODEL == FPUT) {
		if (a.form == kFormCubic) return launch_wave32_form<MODEL, L, S, kFormCubic>(st, state, copies, n, steps, a);
		if (a.form == kFormQuad) return launch_wave32_form<MODEL, L, S, kFormQuad>(st, state, copies, n, steps, a);
	}
	return launch_wave32_form<MODEL, L, S, kFormPlain>(st, state, copies, n, steps, a);
}

template <int MODEL, int L, int S, bool EXACT, int FORM>
__global__ void __launch_bounds__(256)
k_step_wave(double2 *__restrict__ state, uint32_t copies, uint32_t n, uint32_t steps, StepArgs a,
            const double *__restrict__ mp2g) {
	constexpr uint32_t CPW = 64 / L; // chains per wavefront
	__shared__ double etab[MODEL == TODA ? kExpTabSize : 1];
	load_exp_table<MODEL>(etab);
	const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const uint32_t first_chain = wave * CPW;
	if (first_chain >= copies) return; // whole wavefront idle: uniform exit
	run_wave<MODEL, L, S, EXACT, FORM>(state + size_t(first_chain) * n, n, min(CPW, copies - first_chain), steps, a, mp2g,
	                             etab);
}

template <int MODEL, int L, int S, bool EXACT, int FORM>
hipError_t launch_wave_form(hipStream_t st, double2 *state, uint32_t copies, uint32_t n, uint32_t steps,
                            const StepArgs &a, const double *mp2) {
	constexpr uint32_t CPW = 64 / L;
	const uint64_t waves = (uint64_t(copies) + CPW - 1) / CPW;
	const uint64_t blocks = (waves + 3) / 4;
	hipLaunchKernelGGL((k_step_wave<MODEL, L, S, EXACT, FORM>), dim3(uint32_t(blocks)), dim3(256), 0, st, state,
	                   copies, n, steps, a, mp2);
	return hipGetLastError();
}

// The instance for the force form of `a` (forms a model does not have are not
// instantiated; the host never selects them).
template <int MODEL, int L, int S, bool EXACT>
hipError_t launch_wave(hipStream_t st, double2 *state, uint32_t copies, uint32_t n, uint32_t steps,
                       const StepArgs &a, const double *mp2) {
	if constexpr (MODEL != TODA)
		if (a.form == kFormCubic) return launch_wave_form<MODEL, L, S, EXACT, kFormCubic>(st, state, copies, n, steps, a, mp2);
	if constexpr (MODEL == FPUT)
		if (a.form == kFormQuad) return launch_wave_form<MODEL, L, S, EXACT, kFormQuad>(st, state, copies, n, steps, a, mp2);
	return launch_wave_form<MODEL, L, S, EXACT, kFormPlain>(st, state, copies, n, steps, a, mp2);
}

#ifdef NLC_STEP_TILED
// Temporal blocking for chains longer than a workgroup holds.  A wavefront
// takes an extended segment of E = 64 S consecutive sites (64 lanes x S, as a
// 64 x S wave shape) around the C = E - 2 H sites it owns, H on each side, and
// runs T <= H/7 steps on it as if it were a ring.  The two ends see wrong
// neighbours, and a wrong value moves one site inward per kick (every kick
// reads only nearest neighbours, drifts are local), so after 7 T kicks the C
// owned sites are exact; only they are written, to the other buffer.  HBM
// traffic: 16 E bytes in and 16 C out per segment and launch, about 2 B per
// site and step at T = 18.
template <int S>
__host__ __device__ constexpr uint32_t tiled_halo() { return S >= 32 ? 126 : 63; } // 7 T: T = 18 or 9

template <int MODEL, int S, int FORM>
__global__ void __launch_bounds__(256)
k_step_tiled(const double2 *__restrict__ src, double2 *__restrict__ dst, uint32_t n, uint32_t copies, uint32_t segs,
             uint32_t steps, StepArgs a, const double *__restrict__ mp2g) {
	constexpr uint32_t E = 64 * S, H = tiled_halo<S>(), C = E - 2 * H;
	__shared__ double etab[MODEL == TODA ? kExpTabSize : 1];
	load_exp_table<MODEL>(etab);
	const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	if (wave >= copies * segs) return; // whole wavefront idle: uniform exit, no barrier follows
	const uint32_t copy = wave / segs, seg = wave % segs;
	const int lane = threadIdx.x & 63;
	Chain<MODEL, 64, S, true, FORM> ch;
	ch.lane = lane;
	ch.etab = etab;
	ch.full = true;
	const double2 *in = src + size_t(copy) * n;
	// extended position e = lane S + j is global site (seg C - H + e) mod n; one
	// 64-bit reduction per lane (n > H >= S: the sites of a lane wrap at most once)
	uint32_t g = uint32_t((uint64_t(seg) * C + n - H + uint32_t(lane * S)) % n);
#pragma unroll
	for (int j = 0; j < S; ++j, g = (g + 1 == n) ? 0 : g + 1) {
		const double2 v = to_registers<MODEL, FORM, S>(in[g], a);
		ch.q[j] = v.x;
		ch.p[j] = v.y;
		if constexpr (MODEL == DDNKG) ch.mp[j] = mp2g[g];
	}
	step_loop<MODEL>(ch, steps, a);
	double2 *out = dst + size_t(copy) * n;
#pragma unroll
	for (int j = 0; j < S; ++j) {
		const uint32_t e = uint32_t(lane * S + j);
		const uint64_t g = uint64_t(seg) * C + e - H;
		if (e >= H && e < H + C && g < n) out[g] = from_registers<MODEL, FORM, S>(ch.q[j], ch.p[j], a);
	}
}

// `steps` steps of every chain in launches of at most T steps, alternating
// between state and scratch (the result ends in state: an odd launch count
// copies it back).  Needs n > H.
template <int MODEL, int S, int FORM>
hipError_t launch_tiled_form(hipStream_t st, double2 *state, double2 *scratch, uint32_t copies, uint32_t n,
                             uint32_t steps, const StepArgs &a, const double *mp2) {
	constexpr uint32_t E = 64 * S, H = tiled_halo<S>(), C = E - 2 * H, T = H / 7;
	if (n <= H) return hipErrorInvalidValue;
	const uint32_t segs = (n + C - 1) / C;
	const uint64_t blocks = (uint64_t(copies) * segs + 3) / 4;
	if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
	double2 *from = state, *to = scratch;
	for (uint32_t done = 0; done < steps;) {
		const uint32_t k = std::min(T, steps - done);
		hipLaunchKernelGGL((k_step_tiled<MODEL, S, FORM>), dim3(uint32_t(blocks)), dim3(256), 0, st, from, to, n, copies,
		                   segs, k, a, mp2);
		hipError_t e = hipGetLastError();
		if (e != hipSuccess) return e;
		std::swap(from, to);
		done += k;
	}
	if (from != state)
		return hipMemcpyAsync(state, from, sizeof(double2) * size_t(n) * copies, hipMemcpyDeviceToDevice, st);
	return hipSuccess;
}

template <int MODEL, int S>
hipError_t launch_tiled(hipStream_t st, double2 *state, double2 *scratch, uint32_t copies, uint32_t n, uint32_t steps,
                        const StepArgs &a, const double *mp2) {
	if constexpr (MODEL != TODA)
		if (a.form == kFormCubic) return launch_tiled_form<MODEL, S, kFormCubic>(st, state, scratch, copies, n, steps, a, mp2);
	if constexpr (MODEL == FPUT)
		if (a.form == kFormQuad) return launch_tiled_form<MODEL, S, kFormQuad>(st, state, scratch, copies, n, steps, a, mp2);
	return launch_tiled_form<MODEL, S, kFormPlain>(st, state, scratch, copies, n, steps, a, mp2);
}
#endif

#ifdef NLC_PERSIST_S
// Persistent step over a ragged (several chain lengths) or mixed (several
// models) ensemble.  An item is one wavefront's worth of chains of one class:
// 64/L chains of n = L S' sites, S' = S or S/2 sites per lane (flag bit 2), of
// the model and parameters of its parameter set (flag bits 0-1).  Waves take
// items from a queue (an atomic counter) in the order the host sorted them,
// most expensive first, so the cheap items fill the end of the launch (the
// longest-processing-time rule); which wave runs an item does not change its
// result.  The queue counters are zero at every launch: the last wave to finish
// resets them.
//
// KMASK: the kinds compiled in (bit k = kind k of PersistentArgs): one kind
// for a single-model ensemble, so the kernel has that model's register
// footprint; all six for a mixed one.
//
// C32: the FP32-compensated chain state (FPUT and Toda kinds).
template <int KIND, int S, bool C32>
__device__ __forceinline__ void run_item(double2 *__restrict__ state, const RaggedItem &it, uint32_t steps,
                                         const StepArgs &a, const double *etab) {
	constexpr int MODEL = kind_model(KIND), FORM = kind_form(KIND);
	double2 *c0 = state + it.site0;
	if constexpr (C32) {
		switch (it.log2_lanes) {
		case 0: run_wave32<MODEL, 1, S, FORM>(c0, it.n, it.nchains, steps, a, etab); break;
		case 1: run_wave32<MODEL, 2, S, FORM>(c0, it.n, it.nchains, steps, a, etab); break;
		case 2: run_wave32<MODEL, 4, S, FORM>(c0, it.n, it.nchains, steps, a, etab); break;
		case 3: run_wave32<MODEL, 8, S, FORM>(c0, it.n, it.nchains, steps, a, etab); break;
		case 4: run_wave32<MODEL, 16, S, FORM>(c0, it.n, it.nchains, steps, a, etab); break;
		case 5: run_wave32<MODEL, 32, S, FORM>(c0, it.n, it.nchains, steps, a, etab); break;
		default: run_wave32<MODEL, 64, S, FORM>(c0, it.n, it.nchains, steps, a, etab); break;
		}
	} else {
		switch (it.log2_lanes) {
		case 0: run_wave<MODEL, 1, S, true, FORM>(c0, it.n, it.nchains, steps, a, nullptr, etab); break;
		case 1: run_wave<MODEL, 2, S, true, FORM>(c0, it.n, it.nchains, steps, a, nullptr, etab); break;
		case 2: run_wave<MODEL, 4, S, true, FORM>(c0, it.n, it.nchains, steps, a, nullptr, etab); break;
		case 3: run_wave<MODEL, 8, S, true, FORM>(c0, it.n, it.nchains, steps, a, nullptr, etab); break;
		case 4: run_wave<MODEL, 16, S, true, FORM>(c0, it.n, it.nchains, steps, a, nullptr, etab); break;
		case 5: run_wave<MODEL, 32, S, true, FORM>(c0, it.n, it.nchains, steps, a, nullptr, etab); break;
		default: run_wave<MODEL, 64, S, true, FORM>(c0, it.n, it.nchains, steps, a, nullptr, etab); break;
		}
	}
}

template <int KIND, int S, uint32_t KMASK, bool C32>
__device__ __forceinline__ void run_kind(double2 *__restrict__ state, const RaggedItem &it, uint32_t steps,
                                         const StepArgs &a, const double *etab) {
	if constexpr ((KMASK >> KIND) & 1u) {
		if constexpr (S >= 2) {
			if (it.flags & kItemHalfS) {
				run_item<KIND, S / 2, C32>(state, it, steps, a, etab);
				return;
			}
		}
		run_item<KIND, S, C32>(state, it, steps, a, etab);
	}
}

template <int S, uint32_t KMASK, bool C32>
__global__ void __launch_bounds__(256)
k_step_persistent(double2 *__restrict__ state, const RaggedItem *__restrict__ items, uint32_t nitems, uint32_t steps,
                  PersistentArgs pa, uint32_t *__restrict__ queue) {
	constexpr bool kHasToda = (KMASK >> kKindToda) & 1u;
	__shared__ double etab[kHasToda ? kExpTabSize : 1];
	if constexpr (kHasToda) load_exp_table<TODA>(etab);
	const int lane = threadIdx.x & 63;
	for (;;) {
		uint32_t i = 0;
		if (lane == 0) i = atomicAdd(&queue[0], 1u);
		i = __builtin_amdgcn_readfirstlane(i);
		if (i >= nitems) break; // wave-uniform
		const RaggedItem it = items[i];
		const int set = it.flags & kItemSetMask;
		const StepArgs &a = pa.a[set];
		switch (pa.kind[set]) {
		case 0: run_kind<0, S, KMASK, C32>(state, it, steps, a, etab); break;
		case 1: run_kind<1, S, KMASK, C32>(state, it, steps, a, etab); break;
		case 2: run_kind<2, S, KMASK, C32>(state, it, steps, a, etab); break;
		case 3: run_kind<3, S, KMASK, C32>(state, it, steps, a, etab); break;
		case 4: run_kind<4, S, KMASK, C32>(state, it, steps, a, etab); break;
		default: run_kind<5, S, KMASK, C32>(state, it, steps, a, etab); break;
		}
	}
	// the last wave to get here resets the queue for the next launch
	if (lane == 0) {
		const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
		if (atomicAdd(&queue[1], 1u) == nwaves - 1) {
			atomicExch(&queue[0], 0u);
			atomicExch(&queue[1], 0u);
		}
	}
}

template <int S, uint32_t KMASK, bool C32>
hipError_t launch_persistent(hipStream_t st, double2 *state, const RaggedItem *items, uint32_t nitems, uint32_t steps,
                             const PersistentArgs &pa, uint32_t *queue, uint32_t max_blocks) {
	int per_cu = 0, cus = 0, dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_step_persistent<S, KMASK, C32>, 256, 0);
	if (e != hipSuccess) return e;
	// as many waves as are resident at once, no more than there are items
	uint64_t blocks = std::min<uint64_t>(uint64_t(std::max(per_cu, 1)) * std::max(cus, 1), (nitems + 3) / 4);
	if (max_blocks) blocks = std::min<uint64_t>(blocks, max_blocks);
	hipLaunchKernelGGL((k_step_persistent<S, KMASK, C32>), dim3(uint32_t(std::max<uint64_t>(blocks, 1))), dim3(256), 0, st,
	                   state, items, nitems, steps, pa, queue);
	return hipGetLastError();
}
#endif

} // namespace

#ifdef NLC_STEP_TILED
// Toda: 64 x 16 sites with the table-driven exp (T = 9; 64 x 32 with the
// table-free exp measured 9 % slower); the others 64 x 32 (T = 18: every
// launch reads and writes the whole state, about 0.14 ms at 2^24 sites, so
// 64 x 16 with T = 9 measured 15 % slower).
TiledLauncher NLC_STEP_TILED() {
	constexpr int S = NLC_STEP_MODEL == TODA ? 16 : 32;
	return &launch_tiled<NLC_STEP_MODEL, S>;
}
#endif

#ifdef NLC_PERSIST_S
#ifndef NLC_PERSIST_C32
#define NLC_PERSIST_C32 0
#endif
// The persistent kernels with S (and S/2) sites per lane: one per kind, and
// for S >= 16 one with every kind (mixed ensembles); the FP32-compensated
// build (NLC_PERSIST_C32) has the FPUT and Toda kinds.
PersistentLauncher NLC_PERSIST_LOOKUP(uint32_t kind_mask) {
#if NLC_PERSIST_C32
	constexpr uint32_t kC32Kinds = (1u << kKindFputCubic) | (1u << kKindFputQuad) | (1u << kKindFputPlain) |
	                               (1u << kKindToda);
	switch (kind_mask) {
	case 1u << kKindFputCubic: return &launch_persistent<NLC_PERSIST_S, 1u << kKindFputCubic, true>;
	case 1u << kKindFputQuad: return &launch_persistent<NLC_PERSIST_S, 1u << kKindFputQuad, true>;
	case 1u << kKindFputPlain: return &launch_persistent<NLC_PERSIST_S, 1u << kKindFputPlain, true>;
	case 1u << kKindToda: return &launch_persistent<NLC_PERSIST_S, 1u << kKindToda, true>;
	}
#if NLC_PERSIST_S >= 16
	if (kind_mask && !(kind_mask & ~kC32Kinds)) return &launch_persistent<NLC_PERSIST_S, kC32Kinds, true>;
#endif
	return nullptr;
#else
	switch (kind_mask) {
	case 1u << 0: return &launch_persistent<NLC_PERSIST_S, 1u << 0, false>;
	case 1u << 1: return &launch_persistent<NLC_PERSIST_S, 1u << 1, false>;
	case 1u << 2: return &launch_persistent<NLC_PERSIST_S, 1u << 2, false>;
	case 1u << 3: return &launch_persistent<NLC_PERSIST_S, 1u << 3, false>;
	case 1u << 4: return &launch_persistent<NLC_PERSIST_S, 1u << 4, false>;
	case 1u << 5: return &launch_persistent<NLC_PERSIST_S, 1u << 5, false>;
	}
#if NLC_PERSIST_S >= 16
	if (kind_mask && !(kind_mask & ~kAllKinds)) return &launch_persistent<NLC_PERSIST_S, kAllKinds, false>;
#endif
	return nullptr;
#endif
}
#endif

#ifdef NLC_STEP_LOOKUP32
// FP32-compensated wave kernels (FPUT, Toda): every power-of-two shape with
// n = L S, L <= 64, S <= 32.
StepLauncher NLC_STEP_LOOKUP32(int L, int S) {
#define NLC_C32(LL, SS)                                                                                            \
	if (L == LL && S == SS) return &launch_wave32<NLC_STEP_MODEL, LL, SS>;
#define NLC_C32_ROW(LL) NLC_C32(LL, 1) NLC_C32(LL, 2) NLC_C32(LL, 4) NLC_C32(LL, 8) NLC_C32(LL, 16) NLC_C32(LL, 32)
	NLC_C32_ROW(1) NLC_C32_ROW(2) NLC_C32_ROW(4) NLC_C32_ROW(8) NLC_C32_ROW(16) NLC_C32_ROW(32) NLC_C32_ROW(64)
#undef NLC_C32_ROW
#undef NLC_C32
	return nullptr;
}
#endif

#ifdef NLC_STEP_LOOKUP
// Shapes instantiated: every power-of-two (L, S) with 2 <= L*S <= 2048, S <= 32
// (exact fit), plus L = 64 with S = 2..32 for any n in (64(S-1), 64S]; for
// short chains whose length is not a power of two, L = 1 with S = n <= 32
// (thread per chain) and L = 4 with S = 9..16 for n in 33..63.  (L = 16 with
// S = 5..32, for n in 65..512, is compiled in its own translation unit.)
StepLauncher NLC_STEP_LOOKUP(int L, int S, bool exact) {
#define NLC_EXACT(LL, SS)                                                                                          \
	if (exact && L == LL && S == SS) return &launch_wave<NLC_STEP_MODEL, LL, SS, true>;
#define NLC_RAGGED64(SS)                                                                                           \
	if (!exact && L == 64 && S == SS) return &launch_wave<NLC_STEP_MODEL, 64, SS, false>;
#define NLC_RAGGED4(SS)                                                                                            \
	if (!exact && L == 4 && S == SS) return &launch_wave<NLC_STEP_MODEL, 4, SS, false>;
	NLC_EXACT(1, 3) NLC_EXACT(1, 5) NLC_EXACT(1, 6) NLC_EXACT(1, 7) NLC_EXACT(1, 9) NLC_EXACT(1, 10)
	NLC_EXACT(1, 11) NLC_EXACT(1, 12) NLC_EXACT(1, 13) NLC_EXACT(1, 14) NLC_EXACT(1, 15) NLC_EXACT(1, 17)
	NLC_EXACT(1, 18) NLC_EXACT(1, 19) NLC_EXACT(1, 20) NLC_EXACT(1, 21) NLC_EXACT(1, 22) NLC_EXACT(1, 23)
	NLC_EXACT(1, 24) NLC_EXACT(1, 25) NLC_EXACT(1, 26) NLC_EXACT(1, 27) NLC_EXACT(1, 28) NLC_EXACT(1, 29)
	NLC_EXACT(1, 30) NLC_EXACT(1, 31)
	NLC_RAGGED4(9) NLC_RAGGED4(10) NLC_RAGGED4(11) NLC_RAGGED4(12) NLC_RAGGED4(13) NLC_RAGGED4(14)
	NLC_RAGGED4(15) NLC_RAGGED4(16)
	NLC_EXACT(1, 2) NLC_EXACT(1, 4) NLC_EXACT(1, 8) NLC_EXACT(1, 16) NLC_EXACT(1, 32)
	NLC_EXACT(2, 1) NLC_EXACT(2, 2) NLC_EXACT(2, 4) NLC_EXACT(2, 8) NLC_EXACT(2, 16) NLC_EXACT(2, 32)
	NLC_EXACT(4, 1) NLC_EXACT(4, 2) NLC_EXACT(4, 4) NLC_EXACT(4, 8) NLC_EXACT(4, 16) NLC_EXACT(4, 32)
	NLC_EXACT(8, 1) NLC_EXACT(8, 2) NLC_EXACT(8, 4) NLC_EXACT(8, 8) NLC_EXACT(8, 16) NLC_EXACT(8, 32)
	NLC_EXACT(16, 1) NLC_EXACT(16, 2) NLC_EXACT(16, 4) NLC_EXACT(16, 8) NLC_EXACT(16, 16) NLC_EXACT(16, 32)
	NLC_EXACT(32, 1) NLC_EXACT(32, 2) NLC_EXACT(32, 4) NLC_EXACT(32, 8) NLC_EXACT(32, 16) NLC_EXACT(32, 32)
	NLC_EXACT(64, 1) NLC_EXACT(64, 2) NLC_EXACT(64, 4) NLC_EXACT(64, 8) NLC_EXACT(64, 16) NLC_EXACT(64, 32)
	NLC_RAGGED64(2) NLC_RAGGED64(3) NLC_RAGGED64(4) NLC_RAGGED64(5) NLC_RAGGED64(6) NLC_RAGGED64(7)
	NLC_RAGGED64(8) NLC_RAGGED64(9) NLC_RAGGED64(10) NLC_RAGGED64(11) NLC_RAGGED64(12) NLC_RAGGED64(13)
	NLC_RAGGED64(14) NLC_RAGGED64(15) NLC_RAGGED64(16) NLC_RAGGED64(17) NLC_RAGGED64(18) NLC_RAGGED64(19)
	NLC_RAGGED64(20) NLC_RAGGED64(21) NLC_RAGGED64(22) NLC_RAGGED64(23) NLC_RAGGED64(24) NLC_RAGGED64(25)
	NLC_RAGGED64(26) NLC_RAGGED64(27) NLC_RAGGED64(28) NLC_RAGGED64(29) NLC_RAGGED64(30) NLC_RAGGED64(31)
	NLC_RAGGED64(32)
#undef NLC_EXACT
#undef NLC_RAGGED64
#undef NLC_RAGGED4
	return nullptr;
}
#endif

#ifdef NLC_STEP_LOOKUP16
// L = 16 lanes (one DPP row) with S = 5..32 sites per lane for any n in
// (16(S-1), 16S]: chains of 65..512 sites whose length is not a power of two,
// four to a wavefront.
StepLauncher NLC_STEP_LOOKUP16(int S) {
	switch (S) {
#define NLC_R16(SS)                                                                                                \
	case SS: return &launch_wave<NLC_STEP_MODEL, 16, SS, false>;
		NLC_R16(5) NLC_R16(6) NLC_R16(7) NLC_R16(8) NLC_R16(9) NLC_R16(10) NLC_R16(11) NLC_R16(12) NLC_R16(13)
		NLC_R16(14) NLC_R16(15) NLC_R16(16) NLC_R16(17) NLC_R16(18) NLC_R16(19) NLC_R16(20) NLC_R16(21)
		NLC_R16(22) NLC_R16(23) NLC_R16(24) NLC_R16(25) NLC_R16(26) NLC_R16(27) NLC_R16(28) NLC_R16(29)
		NLC_R16(30) NLC_R16(31) NLC_R16(32)
#undef NLC_R16
	}
	return nullptr;
}
#endif

} // namespace nlc
