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
//    16 x 16 transposes, so a substep needs two workgroup barriers (one per
//    cross-wave exchange), every slot is a per-lane base plus a compile-time
//    offset, and neither exchange has bank conflicts.
//  * k_dnls_4096_r8: the same transform as four radix-8 passes on 512 threads
//    x 8 values (more wavefronts per SIMD, one more LDS exchange).
//  * k_dnls_lds: any power-of-two n <= 8192, chain in LDS, radix-4 Stockham
//    passes (the inverse transform as conj(FFT(conj(.)))).
//  * k_dnls_bluestein: any other n <= 4096.  Each n-point DFT is Bluestein's
//    chirp-z convolution (L. Bluestein, 1970): with w_j = e^{-i pi j^2/n},
//    X_k = w_k sum_x (w_x psi_x) conj(w_{k-x}), a cyclic convolution of length
//    M >= 2n - 1 (a power of two) done as two M-point FFTs in LDS.
//  * the nonlinear phase e^{-i theta}: Taylor polynomials of sin and cos on
//    |r| <= pi/16 (degree 11 and 12), after a Cody & Waite reduction by pi/8
//    for larger angles.
#include "nlc_internal.hpp"
#include "nlc_fft_large.hpp"

#include <cmath>
#include <utility>
#include <vector>

namespace nlc {
namespace {

constexpr int kT = 256; // threads per chain

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
	return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// (cos r, sin r) for |r| <= pi/16: Taylor series about 0 in x = r^2 (degree
// 12 and 11); the first omitted terms are below 1e-18 relative.
__device__ __forceinline__ double2 expi_small(double r) {
	const double x = r * r;
	double s = -2.505210838544172e-08;         // -1/11!
	s = fma(s, x, 2.7557319223985893e-06);     // 1/9!
	s = fma(s, x, -0.0001984126984126984);     // -1/7!
	s = fma(s, x, 0.008333333333333333);       // 1/5!
	s = fma(s, x, -0.16666666666666666);       // -1/3!
	double c = 2.08767569878681e-09;           // 1/12!
	c = fma(c, x, -2.755731922398589e-07);     // -1/10!
	c = fma(c, x, 2.48015873015873e-05);       // 1/8!
	c = fma(c, x, -0.001388888888888889);      // -1/6!
	c = fma(c, x, 0.041666666666666664);       // 1/4!
	c = fma(c, x, -0.5);                       // -1/2!
	return make_double2(fma(c, x, 1.0), fma(r * x, s, r));
}

// e^{i k pi/8}, k = 0..15
__constant__ double2 kOctant[16] = {
    {1.0, 0.0},
    {0.9238795325112867, 0.3826834323650898},
    {0.7071067811865476, 0.7071067811865476},
    {0.3826834323650898, 0.9238795325112867},
    {0.0, 1.0},
    {-0.3826834323650898, 0.9238795325112867},
    {-0.7071067811865476, 0.7071067811865476},
    {-0.9238795325112867, 0.3826834323650898},
    {-1.0, 0.0},
    {-0.9238795325112867, -0.3826834323650898},
    {-0.7071067811865476, -0.7071067811865476},
    {-0.3826834323650898, -0.9238795325112867},
    {0.0, -1.0},
    {0.3826834323650898, -0.9238795325112867},
    {0.7071067811865476, -0.7071067811865476},
    {0.9238795325112867, -0.3826834323650898},
};

// (cos a, sin a).  |a| <= pi/16 directly (the nonlinear angles of a DNLS run
// are far smaller, so the other branch is normally skipped by the whole
// wavefront); otherwise Cody & Waite reduction a = k pi/8 + r, |r| <= pi/16,
// with pi/8 split in three parts (33, 33 and 53 significant bits, so k*Q1 and
// k*Q2 are exact for |k| < 2^20), and e^{ia} = e^{i k pi/8} e^{ir}.
constexpr double kExpiSmall = 0.19634954084936207; // pi/16

// The reduced form, valid for every angle; for |a| <= pi/16 it gives exactly
// expi_small(a) (k = 0, r = a, and the octant factor is 1 + 0i).
__device__ __forceinline__ double2 expi_reduced(double a) {
	constexpr double kEightOverPi = 2.5464790894703255, kQ1 = 0x1.921fb54400000p-2, kQ2 = 0x1.0b4611a600000p-36,
	                 kQ3 = 0x1.3198a2e037073p-71;
	const double k = rint(a * kEightOverPi);
	double r = fma(-k, kQ1, a);
	r = fma(-k, kQ2, r);
	r = fma(-k, kQ3, r);
	return cmul(expi_small(r), kOctant[int(k) & 15]);
}

__device__ __forceinline__ double2 expi(double a) {
	return fabs(a) <= kExpiSmall ? expi_small(a) : expi_reduced(a);
}

__device__ __forceinline__ double2 rotate(double2 p, double2 e) {
	return make_double2(fma(p.x, e.x, -p.y * e.y), fma(p.x, e.y, p.y * e.x));
}

// psi * e^{-i theta |psi|^2} (DNLS.cu:11-17)
__device__ __forceinline__ double2 nonlinear(double2 p, double theta) {
	return rotate(p, expi(-theta * fma(p.x, p.x, p.y * p.y)));
}

// The nonlinear step on all N values of a thread.  The branch between the
// short and the reduced form is taken once for the whole wavefront (the
// reduced form only when some angle of some lane exceeds pi/16), so each path
// is straight-line code over the N values: their polynomial chains interleave
// instead of running one after another behind a per-value branch.
template <int N>
__device__ __forceinline__ void nonlinear_all(double2 (&v)[N], double theta) {
	double a[N];
	bool small = true;
#pragma unroll
	for (int j = 0; j < N; ++j) {
		a[j] = -theta * fma(v[j].x, v[j].x, v[j].y * v[j].y);
		small = small && fabs(a[j]) <= kExpiSmall;
	}
	if (__all(small)) {
#pragma unroll
		for (int j = 0; j < N; ++j) v[j] = rotate(v[j], expi_small(a[j]));
	} else {
#pragma unroll
		for (int j = 0; j < N; ++j) v[j] = rotate(v[j], expi_reduced(a[j]));
	}
}

// ---------------------------------------------------------------- radix 16

// 4-point DFT in place, SIGN = -1 forward, +1 inverse.
template <int SIGN>
__device__ __forceinline__ void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3) {
	const double2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
	a0 = cadd(t0, t2);
	a2 = csub(t0, t2);
	// (SIGN i) t3
	const double2 it3 = SIGN < 0 ? make_double2(t3.y, -t3.x) : make_double2(-t3.y, t3.x);
	a1 = cadd(t1, it3);
	a3 = csub(t1, it3);
}

// exp(SIGN 2 pi i e / 16) for e in {1, 2, 3, 6, 9}
template <int SIGN, int E>
__device__ __forceinline__ double2 w16() {
	constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
	if constexpr (E == 1) return make_double2(c1, SIGN * s1);
	else if constexpr (E == 2) return make_double2(h, SIGN * h);
	else if constexpr (E == 3) return make_double2(s1, SIGN * c1);
	else if constexpr (E == 6) return make_double2(-h, SIGN * h);
	else return make_double2(-c1, -SIGN * s1); // E == 9
}

// 16-point DFT in registers, in place, v[.] <- sum_j v[j] exp(SIGN 2 pi i jk/16),
// as two layers of 4-point DFTs (j = 4a + b, k = c + 4d) around the twiddles
// W16^{bc}.  Element k of a transform lives either at register k ("natural")
// or at register pos(k) = 4 (k & 3) + (k >> 2) ("transposed"); pos is an
// involution and every index is a compile-time constant, so the order is
// tracked for free instead of copying registers.
__host__ __device__ constexpr int pos(int k) { return 4 * (k & 3) + (k >> 2); }

// natural in -> transposed out
template <int SIGN>
__device__ __forceinline__ void dft16_n2t(double2 v[16]) {
#pragma unroll
	for (int b = 0; b < 4; ++b) dft4<SIGN>(v[b], v[4 + b], v[8 + b], v[12 + b]); // u[b][c] at 4c + b
	v[4 * 1 + 1] = cmul(v[4 * 1 + 1], w16<SIGN, 1>());
	v[4 * 2 + 1] = cmul(v[4 * 2 + 1], w16<SIGN, 2>());
	v[4 * 3 + 1] = cmul(v[4 * 3 + 1], w16<SIGN, 3>());
	v[4 * 1 + 2] = cmul(v[4 * 1 + 2], w16<SIGN, 2>());
	{
		const double2 t = v[4 * 2 + 2]; // W16^4 = SIGN i
		v[4 * 2 + 2] = SIGN < 0 ? make_double2(t.y, -t.x) : make_double2(-t.y, t.x);
	}
	v[4 * 3 + 2] = cmul(v[4 * 3 + 2], w16<SIGN, 6>());
	v[4 * 1 + 3] = cmul(v[4 * 1 + 3], w16<SIGN, 3>());
	v[4 * 2 + 3] = cmul(v[4 * 2 + 3], w16<SIGN, 6>());
	v[4 * 3 + 3] = cmul(v[4 * 3 + 3], w16<SIGN, 9>());
#pragma unroll
	for (int c = 0; c < 4; ++c) dft4<SIGN>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]); // X[c+4d] at 4c + d
}

// transposed in -> natural out: element k = c + 4d at register 4c + d; DFT4
// over d (contiguous), twiddle W16^{c b'}, DFT4 over c gives j = 4a' + b'.
template <int SIGN>
__device__ __forceinline__ void dft16_t2n(double2 v[16]) {
#pragma unroll
	for (int c = 0; c < 4; ++c) dft4<SIGN>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]); // Y[c][b'] at 4c + b'
	v[4 * 1 + 1] = cmul(v[4 * 1 + 1], w16<SIGN, 1>());
	v[4 * 2 + 1] = cmul(v[4 * 2 + 1], w16<SIGN, 2>());
	v[4 * 3 + 1] = cmul(v[4 * 3 + 1], w16<SIGN, 3>());
	v[4 * 1 + 2] = cmul(v[4 * 1 + 2], w16<SIGN, 2>());
	{
		const double2 t = v[4 * 2 + 2];
		v[4 * 2 + 2] = SIGN < 0 ? make_double2(t.y, -t.x) : make_double2(-t.y, t.x);
	}
	v[4 * 3 + 2] = cmul(v[4 * 3 + 2], w16<SIGN, 6>());
	v[4 * 1 + 3] = cmul(v[4 * 1 + 3], w16<SIGN, 3>());
	v[4 * 2 + 3] = cmul(v[4 * 2 + 3], w16<SIGN, 6>());
	v[4 * 3 + 3] = cmul(v[4 * 3 + 3], w16<SIGN, 9>());
#pragma unroll
	for (int b = 0; b < 4; ++b) dft4<SIGN>(v[b], v[4 + b], v[8 + b], v[12 + b]); // x[4a' + b'] at 4a' + b'
}

// v[j] *= W^{e0 + j de} for j = 0..15 with W = exp(-2 pi i/4096) = tw[1]
// (conjugated when CONJ); ONE0 asserts e0 == 0, so v[0] is left alone.  The
// running power is refreshed from the exact table every 4 steps to bound
// rounding.
template <bool CONJ, bool ONE0, bool TRANSPOSED>
__device__ __forceinline__ void twiddle_run(double2 v[16], const double2 *__restrict__ tw, uint32_t e0, uint32_t de) {
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		double2 r = tw[(e0 + uint32_t(4 * q) * de) & 4095u];
		const double2 w = tw[de & 4095u];
		const double2 wc = CONJ ? cconj(w) : w;
		if (CONJ) r = cconj(r);
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const int idx = 4 * q + j;
			const int at = TRANSPOSED ? pos(idx) : idx;
			if (idx || !ONE0) v[at] = cmul(v[at], r);
			if (j < 3) r = cmul(r, wc);
		}
	}
}

// LDS of the radix-16 kernel: 4 regions of 1088 entries, one per wavefront.
// Cross-wave exchange: element a in [0, 4096) of the natural layout at
// a + 64 (a >> 10), so wavefront w reads exactly [1088 w, 1088 w + 1024) of
// its own region.  In-wave exchange: 16 x 16 transposes with rows of 17
// entries (bank-conflict free) at 1088 w + 272 g, g = 16-lane group in the
// wave, inside the same region.  Every slot is a per-lane base plus a
// compile-time offset.
struct Lds4352 {
	double2 a[4352];
};

// Barriers of the LDS exchanges.  The fences are restricted to LDS: a plain
// __syncthreads() also drains every global load in flight (s_waitcnt vmcnt(0)),
// which would stall on the Fourier multipliers and twiddles read ahead of
// their use; the exchanges only need the LDS writes of the other lanes.
__device__ __forceinline__ void wave_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ void lds_barrier() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Forward FFT.  In: thread t = x0 + 16 x1 holds psi[t + 256 x2] at v[pos(x2)].
// Out: thread t with k1 = t & 15, k0 = t >> 4 holds X[k0 + 16 k1 + 256 k2] at
// v[pos(k2)].
//
// LDS hazards: exchange 1 writes only the slots t + 256 k0 that the same
// thread read last (in exchange 1' of the previous inverse transform), and
// every other write goes to the wavefront's own region, last read by that
// wavefront; LDS operations of one wavefront complete in order.  So the only
// workgroup barriers are the ones between the writes and the reads of the
// cross-wave exchanges.
__device__ __forceinline__ void fft4096_fwd(double2 v[16], Lds4352 &L, const double2 *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lo = t & 15, hi = t >> 4;
	const uint32_t w = t >> 6, g = hi & 3;
	// stage 1: DFT over x2 (k0 natural), twiddle W4096^{(x0+16x1) k0}
	dft16_t2n<-1>(v);
	twiddle_run<false, true, false>(v, tw, 0, t);
	// exchange 1 (cross-wave): C'[x0,x1,k0] at x0 + 16 x1 + 256 k0 (+ region padding)
	double2 *e1w = L.a + t;
#pragma unroll
	for (int k0 = 0; k0 < 16; ++k0) e1w[256 * k0 + 64 * (k0 >> 2)] = v[k0];
	lds_barrier();
	// thread (x0 = lo, k0 = hi) gathers x1
	const double2 *e1r = L.a + lo + 256 * hi + 64 * w;
#pragma unroll
	for (int i = 0; i < 16; ++i) { // in the order the first DFT-4 layer consumes them
		const int x1 = 4 * (i & 3) + (i >> 2);
		v[x1] = e1r[16 * x1];
	}
	// stage 2: DFT over x1 (k1 transposed), twiddle W256^{x0 k1} = W4096^{16 x0 k1}
	dft16_n2t<-1>(v);
	twiddle_run<false, true, true>(v, tw, 0, 16 * lo);
	// exchange 2 (within the 16-lane group): element (x0 = lo, k1)
	double2 *grp = L.a + 1088 * w + 272 * g;
	wave_sync();
#pragma unroll
	for (int k1 = 0; k1 < 16; ++k1) grp[lo + 17 * k1] = v[pos(k1)];
	wave_sync();
	// thread (k1 = lo, k0 = hi) gathers x0
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int x0 = 4 * (i & 3) + (i >> 2);
		v[x0] = grp[x0 + 17 * lo];
	}
	// stage 3: DFT over x0 (k2 transposed)
	dft16_n2t<-1>(v);
}

// Inverse (unnormalised) FFT, the exact reverse of fft4096_fwd: spectrum at
// v[pos(k2)] in, psi[t + 256 x2] at v[pos(x2)] out.
__device__ __forceinline__ void fft4096_inv(double2 v[16], Lds4352 &L, const double2 *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lo = t & 15, hi = t >> 4;
	const uint32_t w = t >> 6, g = hi & 3;
	// stage 3': over k2 -> x0 natural (thread k1 = lo), twiddle W256^{-x0 k1}
	dft16_t2n<+1>(v);
	twiddle_run<true, true, false>(v, tw, 0, 16 * lo);
	// exchange 2' (within the 16-lane group): element (x0, k1 = lo)
	double2 *grp = L.a + 1088 * w + 272 * g;
	wave_sync();
#pragma unroll
	for (int x0 = 0; x0 < 16; ++x0) grp[x0 + 17 * lo] = v[x0];
	wave_sync();
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int k1 = 4 * (i & 3) + (i >> 2);
		v[k1] = grp[lo + 17 * k1];
	}
	// stage 2': over k1 -> x1 transposed (thread x0 = lo, k0 = hi), twiddle W4096^{-(x0 + 16 x1) k0}
	dft16_n2t<+1>(v);
	twiddle_run<true, false, true>(v, tw, lo * hi, 16 * hi);
	// exchange 1' (cross-wave), written into the wavefront's own region
	double2 *e1w = L.a + lo + 256 * hi + 64 * w;
	wave_sync();
#pragma unroll
	for (int x1 = 0; x1 < 16; ++x1) e1w[16 * x1] = v[pos(x1)];
	lds_barrier();
	const double2 *e1r = L.a + t;
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int k0 = 4 * (i & 3) + (i >> 2);
		v[k0] = e1r[256 * k0 + 64 * (k0 >> 2)];
	}
	// stage 1': over k0 -> x2 transposed
	dft16_n2t<+1>(v);
}

// theta[k] = beta dt c_k (k < 7), theta[7] = beta dt c_7, theta[8] = beta dt 2 c_0.
struct DnlsArgs {
	double theta[9];
};

__global__ void __launch_bounds__(kT, 2)
k_dnls_4096(double2 *__restrict__ state, uint32_t copies, uint32_t steps, DnlsArgs a, const double2 *__restrict__ tw,
            const double2 *__restrict__ mult) {
	__shared__ Lds4352 L;
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
			nonlinear_all(v, th);
			fft4096_fwd(v, L, twk);
			// Fourier multiplier, stored per thread in the digit-reversed order
			const double2 *m = mult + size_t(k < 4 ? k : 6 - k) * 4096 + t;
#pragma unroll
			for (int j = 0; j < 16; ++j) v[pos(j)] = cmul(v[pos(j)], m[256 * j]);
			fft4096_inv(v, L, twk);
		}
	}
#pragma unroll
	for (int j = 0; j < 16; ++j) psi[t + 256 * j] = nonlinear(v[pos(j)], a.theta[7]);
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

// ------------------------------------------------------------ generic n

// In-place radix-4 (+ final radix-2) Stockham FFT of n <= MAXN points, forward
// sign, natural order; tw[j tws] = e^{-2 pi i j/n} (tws > 1: a longer table).
template <uint32_t MAXN>
__device__ void lds_fft_fwd(double2 *buf, uint32_t n, const double2 *__restrict__ tw, uint32_t tws = 1) {
	constexpr int kMaxPer = MAXN / 4 / kT; // radix-4 butterflies per thread per pass
	const uint32_t tid = threadIdx.x, quarter = n >> 2;
	uint32_t ns = 1;
	for (; ns * 4 <= n; ns *= 4) {
		double2 v[kMaxPer][4];
		const uint32_t span = n / (ns * 4) * tws;
#pragma unroll
		for (int q = 0; q < kMaxPer; ++q) {
			const uint32_t j = tid + q * kT;
			if (j < quarter)
#pragma unroll
				for (int r = 0; r < 4; ++r) v[q][r] = buf[j + r * quarter];
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < kMaxPer; ++q) {
			const uint32_t j = tid + q * kT;
			if (j < quarter) {
				const uint32_t k = j % ns;
				double2 a0 = v[q][0], a1 = cmul(v[q][1], tw[k * span]), a2 = cmul(v[q][2], tw[2 * k * span]),
				        a3 = cmul(v[q][3], tw[3 * k * span]);
				dft4<-1>(a0, a1, a2, a3);
				const uint32_t o = (j / ns) * ns * 4 + k;
				buf[o] = a0, buf[o + ns] = a1, buf[o + 2 * ns] = a2, buf[o + 3 * ns] = a3;
			}
		}
		__syncthreads();
	}
	if (ns < n) {
		const uint32_t half = n >> 1;
		double2 v[2 * kMaxPer][2];
#pragma unroll
		for (int q = 0; q < 2 * kMaxPer; ++q) {
			const uint32_t j = tid + q * kT;
			if (j < half) v[q][0] = buf[j], v[q][1] = buf[j + half];
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < 2 * kMaxPer; ++q) {
			const uint32_t j = tid + q * kT;
			if (j < half) {
				const uint32_t k = j % ns;
				const double2 b = cmul(v[q][1], tw[k * (n / (2 * ns)) * tws]);
				const uint32_t o = (j / ns) * ns * 2 + k;
				buf[o] = cadd(v[q][0], b);
				buf[o + ns] = csub(v[q][0], b);
			}
		}
		__syncthreads();
	}
}

// Chain in LDS.  Between substeps the buffer holds conj(psi): the inverse
// transform is conj(FFT(conj(X T))).
template <uint32_t MAXN>
__global__ void __launch_bounds__(kT)
k_dnls_lds(double2 *__restrict__ state, uint32_t n, uint32_t copies, uint32_t steps, DnlsArgs a,
           const double2 *__restrict__ tw, const double2 *__restrict__ table) {
	extern __shared__ double2 buf[];
	const uint32_t chain = blockIdx.x, tid = threadIdx.x;
	double2 *psi = state + size_t(chain) * n;
	for (uint32_t x = tid; x < n; x += kT) buf[x] = psi[x];
	bool conj_state = false;
	__syncthreads();
	for (uint32_t s = 0; s < steps; ++s) {
		for (int k = 0; k < 7; ++k) {
			const double th = (k == 0 && s) ? a.theta[8] : a.theta[k];
			for (uint32_t x = tid; x < n; x += kT) {
				const double2 p = conj_state ? cconj(buf[x]) : buf[x];
				buf[x] = nonlinear(p, th);
			}
			__syncthreads();
			lds_fft_fwd<MAXN>(buf, n, tw);
			const double2 *tk = table + size_t(k) * n;
			for (uint32_t j = tid; j < n; j += kT) buf[j] = cconj(cmul(buf[j], tk[j]));
			__syncthreads();
			lds_fft_fwd<MAXN>(buf, n, tw);
			conj_state = true;
		}
	}
	for (uint32_t x = tid; x < n; x += kT) {
		const double2 p = conj_state ? cconj(buf[x]) : buf[x];
		psi[x] = nonlinear(p, a.theta[7]);
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
			hipLaunchKernelGGL(k_dnls_row<M>, rows, block, lds, st, scratch, R, tw, table + size_t(k) * n);
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
			hipLaunchKernelGGL(k_dnls_row<M>, rows, block, lds, st, buf, R, twL, bhat);
			hipLaunchKernelGGL((k_blue_inv<R, 1>), cols, block, 0, st, buf, state, n, L, M, copies, 0.0, twL, chirp,
			                   table + size_t(k) * n);
			hipLaunchKernelGGL((k_blue_fwd<R, 2>), cols, block, 0, st, state, buf, n, L, M, copies, 0.0, twL, chirp);
			hipLaunchKernelGGL(k_dnls_row<M>, rows, block, lds, st, buf, R, twL, bhat);
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
	if (n < 2 || pow2(n)) return 0;
	uint32_t M = 1;
	while (M < 2 * n - 1) M <<= 1;
	return M <= kDnlsBluesteinMax ? M : 0;
}

bool dnls_supported(uint32_t n) {
	uint32_t R, M;
	return (pow2(n) && n <= kDnlsLdsMax) || dnls_bluestein_size(n) != 0 || large_fft_split(n, R, M) ||
	       bluestein_two_level_size(n) != 0;
}

bool dnls_two_level(uint32_t n) {
	uint32_t R, M;
	return large_fft_split(n, R, M);
}

// Multipliers of the two-level path in the spectrum's [r][k] order:
// out[k][r M + q] = table[k][r + R q].
void dnls_two_level_tables(uint32_t n, const double2 *table, double2 *out) {
	uint32_t R, M;
	if (!large_fft_split(n, R, M)) return;
	for (int k = 0; k < 7; ++k)
		for (uint32_t r = 0; r < R; ++r)
			for (uint32_t q = 0; q < M; ++q)
				out[size_t(k) * n + size_t(r) * M + q] = table[size_t(k) * n + r + size_t(R) * q];
}

namespace {
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
} // namespace

void dnls_bluestein_tables(uint32_t n, double2 *out) { bluestein_tables(n, dnls_bluestein_size(n), 1, out); }

uint32_t bluestein_two_level_size(uint32_t n) {
	if (n <= 4096 || pow2(n) || n > kLargeFftMax / 2) return 0;
	uint32_t L = 1;
	while (L < 2 * n - 1) L <<= 1;
	return L;
}

void bluestein_two_level_tables(uint32_t n, double2 *out) {
	const uint32_t L = bluestein_two_level_size(n);
	uint32_t R, M;
	if (!L || !large_fft_split(L, R, M)) return;
	bluestein_tables(n, L, R, out);
}

// Fourier multiplier tables (host side).  table[k][j] = T_k[j] for the generic
// kernel; for n = 4096 the 4 distinct tables (d_k = d_{6-k}) are stored in the
// order the register kernels read them:
//   radix 16: entry [k][256 j + t] = T_k[(t >> 4) + 16 (t & 15) + 256 j],
//   radix 8:  entry [k][512 j + t] = T_k[(t >> 6) + 8 ((t >> 3) & 7) + 64 (t & 7) + 512 j].
bool dnls_fast_path(uint32_t n) { return n == 4096; }

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

hipError_t launch_dnls(hipStream_t st, double2 *state, uint32_t n, uint32_t copies, uint32_t steps,
                       const double *theta9, const double2 *tw, const double2 *table, int variant, double2 *scratch,
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
	if (large_fft_split(n, R, M)) {
		if (!scratch) return hipErrorInvalidValue;
		if (M == 8192) return launch_dnls_two_level<16, 8192>(st, state, scratch, n, copies, steps, a, tw, table);
		switch (R) {
		case 4: return launch_dnls_two_level<4, 4096>(st, state, scratch, n, copies, steps, a, tw, table);
		case 8: return launch_dnls_two_level<8, 4096>(st, state, scratch, n, copies, steps, a, tw, table);
		case 16: return launch_dnls_two_level<16, 4096>(st, state, scratch, n, copies, steps, a, tw, table);
		}
		return hipErrorInvalidValue;
	}
	if (dnls_fast_path(n) && variant == kDnlsRadix8) {
		hipError_t e = hipFuncSetAttribute((const void *) k_dnls_4096_r8, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   int(kR8Lds));
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(k_dnls_4096_r8, dim3(copies), dim3(kT8), kR8Lds, st, state, copies, steps, a, tw, table);
	} else if (dnls_fast_path(n)) {
		hipLaunchKernelGGL(k_dnls_4096, dim3(copies), dim3(kT), 0, st, state, copies, steps, a, tw, table);
	} else if (pow2(n) && n <= kDnlsLdsMax) {
		const size_t lds = sizeof(double2) * n;
		if (n <= 4096) {
			hipLaunchKernelGGL(k_dnls_lds<4096>, dim3(copies), dim3(kT), lds, st, state, n, copies, steps, a, tw, table);
		} else {
			hipError_t e = launch_lds_chain_kernel(k_dnls_lds<kDnlsLdsMax>, lds);
			if (e != hipSuccess) return e;
			hipLaunchKernelGGL(k_dnls_lds<kDnlsLdsMax>, dim3(copies), dim3(kT), lds, st, state, n, copies, steps, a, tw,
			                   table);
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
