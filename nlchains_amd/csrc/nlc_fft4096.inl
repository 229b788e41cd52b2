// The 4096-point FFT of the DNLS register kernels, templated on the complex
// type (double2 for the FP64 step, float2 for the FP32-compensated one):
// three radix-16 passes in registers, 256 threads x 16 values per chain, data
// exchanged through LDS between passes.  Included inside
// namespace nlc { namespace { ... } } after nlc_dnls_common.inl.
//
// Decimation in frequency forward (natural order in, digit-reversed spectrum
// out), decimation in time backward (digit-reversed in, natural out), so no
// reordering pass is ever made; the Fourier multipliers are stored in the
// digit-reversed order.

// exp(SIGN 2 pi i e / 16) for e in {1, 2, 3, 6, 9}, rounded to C's precision
template <int SIGN, int E, class C>
__device__ __forceinline__ C w16() {
	using R = typename Cplx<C>::R;
	constexpr R c1 = R(0.92387953251128675613), s1 = R(0.38268343236508977173), h = R(0.70710678118654752440);
	if constexpr (E == 1) return Cplx<C>::make(c1, SIGN * s1);
	else if constexpr (E == 2) return Cplx<C>::make(h, SIGN * h);
	else if constexpr (E == 3) return Cplx<C>::make(s1, SIGN * c1);
	else if constexpr (E == 6) return Cplx<C>::make(-h, SIGN * h);
	else return Cplx<C>::make(-c1, -SIGN * s1); // E == 9
}

// 16-point DFT in registers, in place, v[.] <- sum_j v[j] exp(SIGN 2 pi i jk/16),
// as two layers of 4-point DFTs (j = 4a + b, k = c + 4d) around the twiddles
// W16^{bc}.  Element k of a transform lives either at register k ("natural")
// or at register pos(k) = 4 (k & 3) + (k >> 2) ("transposed"); pos is an
// involution and every index is a compile-time constant, so the order is
// tracked for free instead of copying registers.
__host__ __device__ constexpr int pos(int k) { return 4 * (k & 3) + (k >> 2); }

// DFT4 (signs as dft4) of (x0, w^A x1, w^2A x2, w^3A x3), w = exp(SIGN 2 pi i/16),
// A = 1, 2, 3, in FMA form (Linzer & Feig, IBM J. Res. Dev. 37 (1993)): each
// twiddle is a real factor, cos(pi/4) or cos(pi/8), times a complex number
// with components +-1 and +-tan(pi/8), so the rotation costs two adds or two
// FMAs and the factor rides on the FMAs that combine the results.  20 or 22
// instructions where the complex products and adds take 24 or 28; the same
// transform up to rounding.
template <int SIGN, int A, class C>
__device__ __forceinline__ void dft4_tw(C &x0, C &x1, C &x2, C &x3) {
	using R = typename Cplx<C>::R;
	constexpr R h = R(0.70710678118654752440), c1 = R(0.92387953251128675613), tau = R(0.41421356237309504880);
	constexpr R S = R(SIGN);
	// t0, t1 = x0 +- w^{2A} x2
	C t0, t1;
	if constexpr (A == 2) { // w^4 = S i
		const C z = Cplx<C>::make(-S * x2.y, S * x2.x);
		t0 = cadd(x0, z);
		t1 = csub(x0, z);
	} else { // w^2 = h (1 + S i), w^6 = h (-1 + S i)
		constexpr R g = A == 1 ? R(1) : R(-1);
		const C z = Cplx<C>::make(g * x2.x - S * x2.y, g * x2.y + S * x2.x);
		t0 = Cplx<C>::make(fma(h, z.x, x0.x), fma(h, z.y, x0.y));
		t1 = Cplx<C>::make(fma(-h, z.x, x0.x), fma(-h, z.y, x0.y));
	}
	// t2, t3 = w^A x1 +- w^{3A} x3 = f (up, um)
	C up, um;
	R f;
	if constexpr (A == 2) { // w^2 (x1 +- S i x3), w^2 = h (1 + S i)
		const C p = Cplx<C>::make(x1.x - S * x3.y, x1.y + S * x3.x);
		const C m = Cplx<C>::make(x1.x + S * x3.y, x1.y - S * x3.x);
		up = Cplx<C>::make(p.x - S * p.y, p.y + S * p.x);
		um = Cplx<C>::make(m.x - S * m.y, m.y + S * m.x);
		f = h;
	} else {
		// w^1 = c1 (1 + S i tau), w^3 = c1 (tau + S i), w^9 = -c1 (1 + S i tau)
		C a, b;
		if constexpr (A == 1) {
			a = Cplx<C>::make(fma(-S * tau, x1.y, x1.x), fma(S * tau, x1.x, x1.y));
			b = Cplx<C>::make(fma(tau, x3.x, -S * x3.y), fma(tau, x3.y, S * x3.x));
		} else {
			a = Cplx<C>::make(fma(tau, x1.x, -S * x1.y), fma(tau, x1.y, S * x1.x));
			b = Cplx<C>::make(fma(S * tau, x3.y, -x3.x), fma(-S * tau, x3.x, -x3.y));
		}
		up = cadd(a, b);
		um = csub(a, b);
		f = c1;
	}
	// y0, y2 = t0 +- f up; y1, y3 = t1 +- S i f um
	x0 = Cplx<C>::make(fma(f, up.x, t0.x), fma(f, up.y, t0.y));
	x2 = Cplx<C>::make(fma(-f, up.x, t0.x), fma(-f, up.y, t0.y));
	x1 = Cplx<C>::make(fma(-S * f, um.y, t1.x), fma(S * f, um.x, t1.y));
	x3 = Cplx<C>::make(fma(S * f, um.y, t1.x), fma(-S * f, um.x, t1.y));
}

// The second layer's twiddles are folded into its DFT4s (dft4_tw).
template <class C>
constexpr bool kFoldTwiddles = true;

// natural in -> transposed out
template <int SIGN, class C>
__device__ __forceinline__ void dft16_n2t(C v[16]) {
#pragma unroll
	for (int b = 0; b < 4; ++b) dft4<SIGN>(v[b], v[4 + b], v[8 + b], v[12 + b]); // u[b][c] at 4c + b
	if constexpr (kFoldTwiddles<C>) { // DFT4 over b of W16^{bc} u[b][c]
		dft4<SIGN>(v[0], v[1], v[2], v[3]);
		dft4_tw<SIGN, 1>(v[4], v[5], v[6], v[7]);
		dft4_tw<SIGN, 2>(v[8], v[9], v[10], v[11]);
		dft4_tw<SIGN, 3>(v[12], v[13], v[14], v[15]);
		return;
	}
	v[4 * 1 + 1] = cmul(v[4 * 1 + 1], w16<SIGN, 1, C>());
	v[4 * 2 + 1] = cmul(v[4 * 2 + 1], w16<SIGN, 2, C>());
	v[4 * 3 + 1] = cmul(v[4 * 3 + 1], w16<SIGN, 3, C>());
	v[4 * 1 + 2] = cmul(v[4 * 1 + 2], w16<SIGN, 2, C>());
	{
		const C t = v[4 * 2 + 2]; // W16^4 = SIGN i
		v[4 * 2 + 2] = SIGN < 0 ? Cplx<C>::make(t.y, -t.x) : Cplx<C>::make(-t.y, t.x);
	}
	v[4 * 3 + 2] = cmul(v[4 * 3 + 2], w16<SIGN, 6, C>());
	v[4 * 1 + 3] = cmul(v[4 * 1 + 3], w16<SIGN, 3, C>());
	v[4 * 2 + 3] = cmul(v[4 * 2 + 3], w16<SIGN, 6, C>());
	v[4 * 3 + 3] = cmul(v[4 * 3 + 3], w16<SIGN, 9, C>());
#pragma unroll
	for (int c = 0; c < 4; ++c) dft4<SIGN>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]); // X[c+4d] at 4c + d
}

// transposed in -> natural out: element k = c + 4d at register 4c + d; DFT4
// over d (contiguous), twiddle W16^{c b'}, DFT4 over c gives j = 4a' + b'.
template <int SIGN, class C>
__device__ __forceinline__ void dft16_t2n(C v[16]) {
#pragma unroll
	for (int c = 0; c < 4; ++c) dft4<SIGN>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]); // Y[c][b'] at 4c + b'
	if constexpr (kFoldTwiddles<C>) { // DFT4 over c of W16^{cb} Y[c][b]
		dft4<SIGN>(v[0], v[4], v[8], v[12]);
		dft4_tw<SIGN, 1>(v[1], v[5], v[9], v[13]);
		dft4_tw<SIGN, 2>(v[2], v[6], v[10], v[14]);
		dft4_tw<SIGN, 3>(v[3], v[7], v[11], v[15]);
		return;
	}
	v[4 * 1 + 1] = cmul(v[4 * 1 + 1], w16<SIGN, 1, C>());
	v[4 * 2 + 1] = cmul(v[4 * 2 + 1], w16<SIGN, 2, C>());
	v[4 * 3 + 1] = cmul(v[4 * 3 + 1], w16<SIGN, 3, C>());
	v[4 * 1 + 2] = cmul(v[4 * 1 + 2], w16<SIGN, 2, C>());
	{
		const C t = v[4 * 2 + 2];
		v[4 * 2 + 2] = SIGN < 0 ? Cplx<C>::make(t.y, -t.x) : Cplx<C>::make(-t.y, t.x);
	}
	v[4 * 3 + 2] = cmul(v[4 * 3 + 2], w16<SIGN, 6, C>());
	v[4 * 1 + 3] = cmul(v[4 * 1 + 3], w16<SIGN, 3, C>());
	v[4 * 2 + 3] = cmul(v[4 * 2 + 3], w16<SIGN, 6, C>());
	v[4 * 3 + 3] = cmul(v[4 * 3 + 3], w16<SIGN, 9, C>());
#pragma unroll
	for (int b = 0; b < 4; ++b) dft4<SIGN>(v[b], v[4 + b], v[8 + b], v[12 + b]); // x[4a' + b'] at 4a' + b'
}

// v[j] *= W^{e0 + j de} for j = 0..15 with W = exp(-2 pi i/4096) = tw[1]
// (conjugated when CONJ); ONE0 asserts e0 == 0, so v[0] is left alone.  The
// running power is refreshed from the exact table every 4 steps to bound
// rounding.  In FP64 the third and fourth power of each run of four come from
// the three-term recurrence r_{j+1} = 2 cos(theta) r_j - r_{j-1} (two FMAs
// instead of a 4-instruction complex product, a few ulp of FP64 at most);
// the FP32 kernel keeps the products, whose smaller rounding error it needs.
template <bool CONJ, bool ONE0, bool TRANSPOSED, class C>
__device__ __forceinline__ void twiddle_run(C v[16], const C *__restrict__ tw, uint32_t e0, uint32_t de) {
	using R = typename Cplx<C>::R;
	constexpr bool kRecur = sizeof(R) == 8;
	const C w = tw[de & 4095u];
	const C wc = CONJ ? cconj(w) : w;
	const R c2 = R(2) * w.x; // 2 cos(theta), the same for W and its conjugate
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		// with e0 == 0 the first run starts at the table's entry 0, (1, -0)
		// (conjugated when CONJ), and its next power is W^{de} itself: the same
		// bits as the load and the product, without either (FP64 only: in the
		// FP32 kernel the shorter run changed the register allocation and put
		// spill stores inside the step loop)
		const bool unit = kRecur && ONE0 && q == 0;
		C r;
		if (unit) {
			r = Cplx<C>::make(R(1), CONJ ? R(0) : -R(0));
		} else {
			r = tw[(e0 + uint32_t(4 * q) * de) & 4095u];
			if (CONJ) r = cconj(r);
		}
		C rp = r; // the power before r (used from j = 1 on)
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const int idx = 4 * q + j;
			const int at = TRANSPOSED ? pos(idx) : idx;
			if (idx || !ONE0) v[at] = cmul(v[at], r);
			if (j < 3) {
				C rn;
				if (kRecur && j >= 1) rn = Cplx<C>::make(fma(c2, r.x, -rp.x), fma(c2, r.y, -rp.y));
				else if (unit && j == 0) rn = wc;
				else rn = cmul(r, wc);
				rp = r;
				r = rn;
			}
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
template <class C>
struct Lds4352 {
	C a[4352];
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
template <class C>
__device__ __forceinline__ void fft4096_fwd(C v[16], Lds4352<C> &L, const C *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lo = t & 15, hi = t >> 4;
	const uint32_t w = t >> 6, g = hi & 3;
	// stage 1: DFT over x2 (k0 natural), twiddle W4096^{(x0+16x1) k0}
	dft16_t2n<-1>(v);
	twiddle_run<false, true, false>(v, tw, 0, t);
	// exchange 1 (cross-wave): C'[x0,x1,k0] at x0 + 16 x1 + 256 k0 (+ region padding)
	C *e1w = L.a + t;
#pragma unroll
	for (int k0 = 0; k0 < 16; ++k0) e1w[256 * k0 + 64 * (k0 >> 2)] = v[k0];
	lds_barrier();
	// thread (x0 = lo, k0 = hi) gathers x1
	const C *e1r = L.a + lo + 256 * hi + 64 * w;
#pragma unroll
	for (int i = 0; i < 16; ++i) { // in the order the first DFT-4 layer consumes them
		const int x1 = 4 * (i & 3) + (i >> 2);
		v[x1] = e1r[16 * x1];
	}
	// stage 2: DFT over x1 (k1 transposed), twiddle W256^{x0 k1} = W4096^{16 x0 k1}
	dft16_n2t<-1>(v);
	twiddle_run<false, true, true>(v, tw, 0, 16 * lo);
	// exchange 2 (within the 16-lane group): element (x0 = lo, k1)
	C *grp = L.a + 1088 * w + 272 * g;
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
template <class C>
__device__ __forceinline__ void fft4096_inv(C v[16], Lds4352<C> &L, const C *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lo = t & 15, hi = t >> 4;
	const uint32_t w = t >> 6, g = hi & 3;
	// stage 3': over k2 -> x0 natural (thread k1 = lo), twiddle W256^{-x0 k1}
	dft16_t2n<+1>(v);
	twiddle_run<true, true, false>(v, tw, 0, 16 * lo);
	// exchange 2' (within the 16-lane group): element (x0, k1 = lo)
	C *grp = L.a + 1088 * w + 272 * g;
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
	C *e1w = L.a + lo + 256 * hi + 64 * w;
	wave_sync();
#pragma unroll
	for (int x1 = 0; x1 < 16; ++x1) e1w[16 * x1] = v[pos(x1)];
	lds_barrier();
	const C *e1r = L.a + t;
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int k0 = 4 * (i & 3) + (i >> 2);
		v[k0] = e1r[256 * k0 + 64 * (k0 >> 2)];
	}
	// stage 1': over k0 -> x2 transposed
	dft16_n2t<+1>(v);
}

// ------------------------------------------------ component-split exchanges
//
// The same FP64 transforms with every LDS exchange made twice, real parts
// then imaginary parts, through an array of 4352 doubles (34 KB instead of
// the 68 KB of Lds4352<double2>): three chains' LDS fit a CU, so with at most
// 168 VGPRs three wavefronts per SIMD hide the LDS and L2 latency that two
// leave exposed.  Each cross-wave exchange takes three workgroup barriers
// instead of one (write re, read re, write im, read im: the imaginary parts
// reuse the slots the real parts were read from), each in-wave exchange three
// wave syncs instead of two, and the LDS operations move 8 bytes instead of 16.
//
// Slot of element a in the cross-wave exchanges: a + 64 (a >> 10) + 16 ((a >> 8)
// & 3), injective (the 256-element blocks of a 1024-element quarter start 272
// slots apart) and inside a 1088-slot region per quarter: each wavefront reads
// only its own region (as in Lds4352).  The second term moves consecutive
// 16-lane groups of a wavefront by 16 doubles, so the two groups that one
// 8-byte access instruction serves at a time (32 lanes, 256 bytes) fall in
// different banks.  In-wave exchanges use
// the same 17-entry rows inside the region, now conflict-free for 32 lanes
// (17 is odd, and the next group starts 272 = 16 mod 32 doubles later).
//
// LDS hazards, as for fft4096_fwd: every slot that a thread writes first in
// an exchange is one it read last itself; the other writes follow a
// workgroup barrier (cross-wave) or a wave sync (in-wave, own region).
struct Lds4352d {
	double a[4352];
};

template <int PART>
__device__ __forceinline__ double &part(double2 &z) {
	if constexpr (PART == 0) return z.x;
	else return z.y;
}

template <int PART>
__device__ __forceinline__ void xch1_fwd_write(const double2 v[16], double *L, uint32_t t) {
#pragma unroll
	for (int k0 = 0; k0 < 16; ++k0) L[t + 256 * k0 + 64 * (k0 >> 2) + 16 * (k0 & 3)] = PART == 0 ? v[k0].x : v[k0].y;
}

template <int PART>
__device__ __forceinline__ void xch1_fwd_read(double2 v[16], const double *L, uint32_t lo, uint32_t hi) {
	const double *r = L + lo + 256 * hi + 64 * (hi >> 2) + 16 * (hi & 3);
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int x1 = 4 * (i & 3) + (i >> 2);
		part<PART>(v[x1]) = r[16 * x1];
	}
}

__device__ __forceinline__ void fft4096_fwd_split(double2 v[16], Lds4352d &L, const double2 *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lo = t & 15, hi = t >> 4;
	const uint32_t w = t >> 6, g = hi & 3;
	dft16_t2n<-1>(v);
	twiddle_run<false, true, false>(v, tw, 0, t);
	// exchange 1 (cross-wave), real then imaginary parts
	xch1_fwd_write<0>(v, L.a, t);
	lds_barrier();
	xch1_fwd_read<0>(v, L.a, lo, hi);
	lds_barrier();
	xch1_fwd_write<1>(v, L.a, t);
	lds_barrier();
	xch1_fwd_read<1>(v, L.a, lo, hi);
	dft16_n2t<-1>(v);
	twiddle_run<false, true, true>(v, tw, 0, 16 * lo);
	// exchange 2 (within the 16-lane group)
	double *grp = L.a + 1088 * w + 272 * g;
	wave_sync();
#pragma unroll
	for (int k1 = 0; k1 < 16; ++k1) grp[lo + 17 * k1] = v[pos(k1)].x;
	wave_sync();
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int x0 = 4 * (i & 3) + (i >> 2);
		v[x0].x = grp[x0 + 17 * lo];
	}
	wave_sync();
	// the real pass touched only the .x halves: v[pos(k1)].y still hold the
	// imaginary parts in the old order
#pragma unroll
	for (int k1 = 0; k1 < 16; ++k1) grp[lo + 17 * k1] = v[pos(k1)].y;
	wave_sync();
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int x0 = 4 * (i & 3) + (i >> 2);
		v[x0].y = grp[x0 + 17 * lo];
	}
	dft16_n2t<-1>(v);
}

__device__ __forceinline__ void fft4096_inv_split(double2 v[16], Lds4352d &L, const double2 *__restrict__ tw) {
	const uint32_t t = threadIdx.x, lo = t & 15, hi = t >> 4;
	const uint32_t w = t >> 6, g = hi & 3;
	dft16_t2n<+1>(v);
	twiddle_run<true, true, false>(v, tw, 0, 16 * lo);
	// exchange 2' (within the 16-lane group)
	double *grp = L.a + 1088 * w + 272 * g;
	wave_sync();
#pragma unroll
	for (int x0 = 0; x0 < 16; ++x0) grp[x0 + 17 * lo] = v[x0].x;
	wave_sync();
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int k1 = 4 * (i & 3) + (i >> 2);
		v[k1].x = grp[lo + 17 * k1];
	}
	wave_sync();
#pragma unroll
	for (int x0 = 0; x0 < 16; ++x0) grp[x0 + 17 * lo] = v[x0].y;
	wave_sync();
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int k1 = 4 * (i & 3) + (i >> 2);
		v[k1].y = grp[lo + 17 * k1];
	}
	dft16_n2t<+1>(v);
	twiddle_run<true, false, true>(v, tw, lo * hi, 16 * hi);
	// exchange 1' (cross-wave), written into the wavefront's own region
	double *e1w = L.a + lo + 256 * hi + 64 * w + 16 * (hi & 3);
	const double *e1r = L.a + t;
	wave_sync();
#pragma unroll
	for (int x1 = 0; x1 < 16; ++x1) e1w[16 * x1] = v[pos(x1)].x;
	lds_barrier();
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int k0 = 4 * (i & 3) + (i >> 2);
		v[k0].x = e1r[256 * k0 + 64 * (k0 >> 2) + 16 * (k0 & 3)];
	}
	lds_barrier();
#pragma unroll
	for (int x1 = 0; x1 < 16; ++x1) e1w[16 * x1] = v[pos(x1)].y;
	lds_barrier();
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int k0 = 4 * (i & 3) + (i >> 2);
		v[k0].y = e1r[256 * k0 + 64 * (k0 >> 2) + 16 * (k0 & 3)];
	}
	dft16_n2t<+1>(v);
}
