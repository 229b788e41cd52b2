// Device pieces shared by the DNLS kernels: complex helpers, the nonlinear
// phase e^{-i theta |psi|^2} (Taylor polynomials of sin and cos on |r| <= pi/32,
// degree 9 and 10, after a Cody & Waite reduction by pi/16 for larger angles),
// the 4-point DFT and the LDS Stockham FFT.  Included inside
// namespace nlc { namespace { ... } } by each DNLS translation unit.
constexpr int kT = 256; // threads per chain

// Complex values as double2 (FP64 kernels) or float2 (the FP32 transforms of
// the compensated DNLS step); Cplx<C> gives the component type and a maker.
template <class C>
struct Cplx;
template <>
struct Cplx<double2> {
	using R = double;
	static __device__ __forceinline__ double2 make(double x, double y) { return make_double2(x, y); }
};
template <>
struct Cplx<float2> {
	using R = float;
	static __device__ __forceinline__ float2 make(float x, float y) { return make_float2(x, y); }
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
	return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
	return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// (cos r, sin r) for |r| <= pi/32: Taylor series about 0 in x = r^2 (degree
// 10 and 9); the first omitted terms, r^12/12! and r^11/11!, are below 3e-18
// relative.
__device__ __forceinline__ double2 expi_small(double r) {
	const double x = r * r;
	double s = 2.7557319223985893e-06;         // 1/9!
	s = fma(s, x, -0.0001984126984126984);     // -1/7!
	s = fma(s, x, 0.008333333333333333);       // 1/5!
	s = fma(s, x, -0.16666666666666666);       // -1/3!
	double c = -2.755731922398589e-07;         // -1/10!
	c = fma(c, x, 2.48015873015873e-05);       // 1/8!
	c = fma(c, x, -0.001388888888888889);      // -1/6!
	c = fma(c, x, 0.041666666666666664);       // 1/4!
	c = fma(c, x, -0.5);                       // -1/2!
	return make_double2(fma(c, x, 1.0), fma(r * x, s, r));
}

// e^{i k pi/16}, k = 0..31, correctly rounded (the first octant's values
// carried to the others by the quarter-turn symmetries)
__constant__ double2 kSixteenth[32] = {
    {1.0, 0.0},
    {0.9807852804032304, 0.19509032201612828},
    {0.9238795325112867, 0.3826834323650898},
    {0.8314696123025452, 0.5555702330196022},
    {0.7071067811865476, 0.7071067811865476},
    {0.5555702330196022, 0.8314696123025452},
    {0.3826834323650898, 0.9238795325112867},
    {0.19509032201612828, 0.9807852804032304},
    {0.0, 1.0},
    {-0.19509032201612828, 0.9807852804032304},
    {-0.3826834323650898, 0.9238795325112867},
    {-0.5555702330196022, 0.8314696123025452},
    {-0.7071067811865476, 0.7071067811865476},
    {-0.8314696123025452, 0.5555702330196022},
    {-0.9238795325112867, 0.3826834323650898},
    {-0.9807852804032304, 0.19509032201612828},
    {-1.0, 0.0},
    {-0.9807852804032304, -0.19509032201612828},
    {-0.9238795325112867, -0.3826834323650898},
    {-0.8314696123025452, -0.5555702330196022},
    {-0.7071067811865476, -0.7071067811865476},
    {-0.5555702330196022, -0.8314696123025452},
    {-0.3826834323650898, -0.9238795325112867},
    {-0.19509032201612828, -0.9807852804032304},
    {0.0, -1.0},
    {0.19509032201612828, -0.9807852804032304},
    {0.3826834323650898, -0.9238795325112867},
    {0.5555702330196022, -0.8314696123025452},
    {0.7071067811865476, -0.7071067811865476},
    {0.8314696123025452, -0.5555702330196022},
    {0.9238795325112867, -0.3826834323650898},
    {0.9807852804032304, -0.19509032201612828},
};

// (cos a, sin a).  |a| <= 0.098 directly (the nonlinear angles of a DNLS run
// are far smaller, so the other branch is normally skipped by the whole
// wavefront); otherwise Cody & Waite reduction a = k pi/16 + r, |r| <= pi/32,
// with pi/16 split in three parts (33, 33 and 53 significant bits, so k*Q1 and
// k*Q2 are exact for |k| < 2^20), and e^{ia} = e^{i k pi/16} e^{ir}.  The bound
// 0.098 sits just under pi/32 so that every angle it admits has k = 0.
constexpr double kExpiSmall = 0.098;

// The reduced form, valid for every angle; for |a| <= kExpiSmall it gives
// exactly expi_small(a) (k = 0, r = a, and the table factor is 1 + 0i).
__device__ __forceinline__ double2 expi_reduced(double a) {
	constexpr double kSixteenOverPi = 5.092958178940651, kQ1 = 0x1.921fb54400000p-3, kQ2 = 0x1.0b4611a600000p-37,
	                 kQ3 = 0x1.3198a2e037073p-72;
	const double k = rint(a * kSixteenOverPi);
	double r = fma(-k, kQ1, a);
	r = fma(-k, kQ2, r);
	r = fma(-k, kQ3, r);
	return cmul(expi_small(r), kSixteenth[int(k) & 31]);
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
// reduced form only when some angle of some lane exceeds kExpiSmall), so each path
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

// nonlinear_all in two halves of 8 values, the scheduler kept from
// interleaving them: half the polynomial temporaries live at a time, for a
// kernel that runs at 168 VGPRs (three wavefronts per SIMD hide the latency
// the interleaving hid).  Each half takes its own wave-uniform branch, so the
// angles are computed once (both forms agree bit for bit up to kExpiSmall, so the
// result does not depend on how the values are grouped).
__device__ __forceinline__ void nonlinear_all_halves(double2 (&v)[16], double theta) {
#pragma unroll
	for (int h = 0; h < 2; ++h) {
		double a[8];
		bool small = true;
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			a[j] = -theta * fma(v[8 * h + j].x, v[8 * h + j].x, v[8 * h + j].y * v[8 * h + j].y);
			small = small && fabs(a[j]) <= kExpiSmall;
		}
		if (__all(small)) {
#pragma unroll
			for (int j = 0; j < 8; ++j) v[8 * h + j] = rotate(v[8 * h + j], expi_small(a[j]));
		} else {
#pragma unroll
			for (int j = 0; j < 8; ++j) v[8 * h + j] = rotate(v[8 * h + j], expi_reduced(a[j]));
		}
		__builtin_amdgcn_sched_barrier(0);
	}
}

// 4-point DFT in place, SIGN = -1 forward, +1 inverse.
template <int SIGN, class C>
__device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3) {
	const C t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
	a0 = cadd(t0, t2);
	a2 = csub(t0, t2);
	// (SIGN i) t3
	const C it3 = SIGN < 0 ? Cplx<C>::make(t3.y, -t3.x) : Cplx<C>::make(-t3.y, t3.x);
	a1 = cadd(t1, it3);
	a3 = csub(t1, it3);
}

// theta[k] = beta dt c_k (k < 7), theta[7] = beta dt c_7, theta[8] = beta dt 2 c_0.
struct DnlsArgs {
	double theta[9];
};

// In-place radix-4 (+ final radix-2) Stockham FFT of n <= MAXN points, forward
// sign, natural order; tw[j tws] = e^{-2 pi i j/n} (tws > 1: a longer table).
// Threads tid < nthr of the chain work on it (nthr = kT, or n/4 when several
// chains share the workgroup: one radix-4 butterfly per thread and pass); every
// thread of the workgroup reaches the same barriers.
template <uint32_t MAXN, bool WHOLE>
__device__ void lds_fft_fwd_impl(double2 *buf, uint32_t n, const double2 *__restrict__ tw, uint32_t tws,
                                 uint32_t tid, uint32_t nthr_rt) {
	constexpr int kMaxPer = MAXN / 4 / kT; // radix-4 butterflies per thread per pass (nthr = kT)
	// a compile-time stride for the whole workgroup keeps the LDS offsets immediate
	const uint32_t nthr = WHOLE ? uint32_t(kT) : nthr_rt;
	const uint32_t quarter = n >> 2;
	uint32_t ns = 1;
	for (; ns * 4 <= n; ns *= 4) {
		double2 v[kMaxPer][4];
		const uint32_t span = n / (ns * 4) * tws;
#pragma unroll
		for (int q = 0; q < kMaxPer; ++q) {
			const uint32_t j = tid + q * nthr;
			if (j < quarter)
#pragma unroll
				for (int r = 0; r < 4; ++r) v[q][r] = buf[j + r * quarter];
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < kMaxPer; ++q) {
			const uint32_t j = tid + q * nthr;
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
			const uint32_t j = tid + q * nthr;
			if (j < half) v[q][0] = buf[j], v[q][1] = buf[j + half];
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < 2 * kMaxPer; ++q) {
			const uint32_t j = tid + q * nthr;
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

// The whole workgroup on one chain.
template <uint32_t MAXN>
__device__ void lds_fft_fwd(double2 *buf, uint32_t n, const double2 *__restrict__ tw, uint32_t tws = 1) {
	lds_fft_fwd_impl<MAXN, true>(buf, n, tw, tws, threadIdx.x, kT);
}

// Threads tid < nthr of the workgroup on one of its chains.
template <uint32_t MAXN>
__device__ void lds_fft_fwd(double2 *buf, uint32_t n, const double2 *__restrict__ tw, uint32_t tws, uint32_t tid,
                            uint32_t nthr) {
	lds_fft_fwd_impl<MAXN, false>(buf, n, tw, tws, tid, nthr);
}
