// Double-precision exponentials for the Toda bond force, on an argument
// already expressed in a power-of-two unit: exp2_32_tab(x) = 2^(x/32) (table
// driven) and exp2_poly(y) = 2^y (table-free).  The Toda kernels keep
// v = (64 alpha/ln2) q in registers for the first, (2 alpha/ln2) q for the
// second (the two scalings differ by 2^5, so every product and sum of the step
// scales exactly and both give the same state), and the bond term
// e^{2 alpha (q_i - q_{i-1})} is then exp2_32_tab(v_i - v_{i-1}), resp.
// exp2_poly(v_i - v_{i-1}).  Each unit lets the reduction's constants be the
// instruction's inline constants (2, -1/2) or no constant at all.
//
// Table form: P. T. P. Tang, "Table-driven implementation of the exponential
// function in IEEE floating-point arithmetic", ACM TOMS 15(2), 1989, with a
// 64-entry table and a degree-5 polynomial, on the pre-scaled argument:
//   2x = N + 2r,  N = nearest integer to 2x,  |r| <= 1/4,
//   2^(x/32) = 2^(N >> 6) * 2^((N & 63)/64) * 2^(r/32),
//   2^(r/32) - 1 = r (p0 + p1 r + ... + p4 r^4),
// the p_k interpolating (2^(r/32) - 1)/r at the five Chebyshev nodes of
// |r| <= 1/4 (computed in 60-digit decimal arithmetic; error of the
// polynomial < 2.4e-18 on the interval, against 3.5e-17 for the degree-5
// Taylor polynomial).  N is 2x + 1.5 * 2^52 - 1.5 * 2^52 (an FMA and a
// subtraction) and r = x - N/2 is exact (one FMA), so the reduction of Tang's
// method (y - N ln2/64 in two FMAs) costs nothing extra here; 2^(N >> 6) is
// applied with ldexp, which saturates to +inf / 0 beyond the double range.
// Error about 1 ulp: the table entries are 2^(j/64) rounded to nearest, one
// rounding in the last FMA, and |2^(r/32) - 1| < 0.0055 carries the
// polynomial's rounding errors down to 1/100 ulp.
//
// Range: N is converted to an integer by the saturating conversion (values
// beyond the int32 range clamp to INT32_MIN / INT32_MAX), so ldexp's exponent
// N >> 6 saturates with it and every |x| < 2^50 gives the correctly rounded
// result or the correctly saturated +inf / 0.  (Taking N from the low 32 bits
// of the shifted argument, which costs no instruction, wrapped for |x| >= 2^30:
// a bond stretched that far by a blown-up step got a finite force instead of
// +inf, and the run's non-finite-entropy stop might never come.)  Beyond 2^50,
// where x - N/2 is no longer exact, a positive x gives +inf or -inf and a
// negative one 0 or a non-finite value; non-finite x give NaN.  Such arguments
// come only from a chain that has already blown up; its state is non-finite
// or becomes so.
//
// Cost on gfx950: 13 VALU + 1 LDS read per call, against about 25 VALU for the
// device library's exp (a degree-12 polynomial on |r| <= ln2/2 plus range
// checks).  The table lives in LDS (512 bytes: entry j in banks 2j mod 64 and
// 2j + 1 mod 64, so a wavefront's reads conflict at most two ways).
//
// The header is plain C++ apart from NLC_EXP_QUAL and the saturating
// conversion, so the same code can be compiled on the host (with NLC_EXP_HOST
// defined) to measure its error.
#pragma once

#ifndef NLC_EXP_QUAL
#define NLC_EXP_QUAL __device__ __forceinline__
#endif

namespace nlc {

// double -> int32 with saturation (NaN gives 0).  On the device this is HIP's
// __double2int_rz, which compiles to v_cvt_i32_f64: the instruction clamps
// out-of-range values to INT32_MIN / INT32_MAX.  (An inline-asm form of the same
// instruction kept the compiler from interleaving the 32 exponentials of a
// 32-site lane and doubled the register use of those kernels.)  The host
// build, used to measure the error, spells the saturation out.
#ifdef NLC_EXP_HOST
inline int cvt_sat_i32(double x) {
	if (!(x == x)) return 0;
	if (x >= 2147483647.0) return 2147483647;
	if (x <= -2147483648.0) return -2147483647 - 1;
	return int(x);
}
#else
NLC_EXP_QUAL int cvt_sat_i32(double x) { return __double2int_rz(x); }
#endif

constexpr int kExpTabSize = 64;

// 2^(j/64), j = 0..63, rounded to nearest double.
#define NLC_EXP2_TABLE                                                                                              \
	{0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,                           \
	 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,                           \
	 0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,                           \
	 0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,                           \
	 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,                           \
	 0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,                           \
	 0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,                           \
	 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,                           \
	 0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,                           \
	 0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,                           \
	 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,                           \
	 0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,                           \
	 0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,                           \
	 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,                           \
	 0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,                           \
	 0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0}

constexpr double kExpShifter = 0x1.8p52; // 1.5 * 2^52: y + kExpShifter rounds y to an integer

// 2^(x/32) = e^(x ln2/32).
NLC_EXP_QUAL double exp2_32_tab(double x, const double *tab) {
	const double s = __builtin_fma(x, 2.0, kExpShifter); // rounds 2x to the nearest integer N
	const double nd = s - kExpShifter;                    // N as a double, exact for |x| < 2^50
	const int n = cvt_sat_i32(nd);
	const double r = __builtin_fma(nd, -0.5, x); // x - N/2, exact, |r| <= 1/4
	// (2^(r/32) - 1) / r = sum_k p_k r^k
	double t = __builtin_fma(r, 0x1.5d8812769ed72p-35, 0x1.3b2ad0385b409p-27); // p4, p3
	t = __builtin_fma(t, r, 0x1.c6b08d70496c1p-20);                            // p2
	t = __builtin_fma(t, r, 0x1.ebfbdff82ac52p-13);                            // p1
	t = __builtin_fma(t, r, 0x1.62e42fefa39efp-6);                             // p0
	const double q = t * r;
	const double T = tab[n & (kExpTabSize - 1)];
	return __builtin_ldexp(__builtin_fma(T, q, T), n >> 6);
}

// Table-free variant: y = N + r, N = y rounded to the nearest integer (ties
// to even, by the rounding instruction), |r| <= 1/2, 2^r - 1 = r (q0 + q1 r +
// ... + q10 r^10), the q_k interpolating (2^r - 1)/r at the eleven Chebyshev
// nodes of |r| <= 1/2 (error < 2.1e-17 with the coefficients rounded to
// double, against 1.3e-17 for the degree-13 Taylor polynomial), 2^N by ldexp
// with N converted as for the table form.  15 VALU and no memory access (on
// an argument in 1/32 octaves it takes 16: the scaling to octaves before the
// rounding; here the state carries it).  With the table, the scheduler hoists
// the table loads of a whole unrolled run of bonds to hide their latency and
// the kernel loses occupancy; this form keeps a chain of bonds in few
// registers.  r = y - N is exact for every finite y, so this form saturates
// exactly to +inf / 0 for all finite arguments.
NLC_EXP_QUAL double exp2_poly(double y) {
	const double nd = __builtin_rint(y);
	const int n = cvt_sat_i32(nd);
	const double r = y - nd; // exact, |r| <= 1/2
	double t = __builtin_fma(r, 0x1.e9d3fe3952179p-32, 0x1.e6063f7217bc6p-28); // q10, q9
	t = __builtin_fma(t, r, 0x1.b524fae627834p-24);                                // q8
	t = __builtin_fma(t, r, 0x1.62bfd47773353p-20);                                // q7
	t = __builtin_fma(t, r, 0x1.ffcbfc670dcd4p-17);                                // q6
	t = __builtin_fma(t, r, 0x1.430913096fd9fp-13);                                // q5
	t = __builtin_fma(t, r, 0x1.5d87fe78a5276p-10);                                // q4
	t = __builtin_fma(t, r, 0x1.3b2ab6fba1ddap-7);                                 // q3
	t = __builtin_fma(t, r, 0x1.c6b08d704a0c2p-5);                                 // q2
	t = __builtin_fma(t, r, 0x1.ebfbdff82c598p-3);                                 // q1
	t = __builtin_fma(t, r, 0x1.62e42fefa39efp-1);                                 // q0
	return __builtin_ldexp(__builtin_fma(t, r, 1.0), n);
}

} // namespace nlc
