// Double-precision e^y for the Toda bond force, table driven.
//
// Method: P. T. P. Tang, "Table-driven implementation of the exponential
// function in IEEE floating-point arithmetic", ACM TOMS 15(2), 1989, with a
// 32-entry table and a Taylor polynomial:
//   y = N ln2/32 + r,  N = nearest integer to 32 y / ln2,  |r| <= ln2/64,
//   e^y = 2^(N >> 5) * 2^((N & 31)/32) * e^r,
//   e^r - 1 = r (1 + r/2! + r^2/3! + ... + r^5/6!)   (truncation < 4e-18).
// N comes from the low word of fma(y, 32/ln2, 1.5 * 2^52) (no conversion
// instruction); r = y - N ln2/32 is formed with two FMAs (ln2/32 split into a
// rounded head and the rounded remainder); 2^(N >> 5) is applied with ldexp,
// which saturates to +inf / 0 beyond the double range.  Error about 1 ulp:
// the table entries are 2^(j/32) rounded to nearest, one rounding in the last
// FMA, |e^r - 1| < 0.011 carries the polynomial's few-ulp error down to 1/100 ulp.
// Valid for |y| < 4e7 (larger magnitudes are reached only by a diverged chain,
// whose state is non-finite long before); NaN propagates.
//
// Cost on gfx950: 14 VALU + 1 LDS read per call, against about 25 VALU for the
// device library's exp (a degree-12 polynomial on |r| <= ln2/2 plus range
// checks).  The table lives in LDS (256 bytes: entry j in banks 2j, 2j+1, so a
// wavefront's reads never conflict).
//
// The header is plain C++ apart from NLC_EXP_QUAL, so the same code can be
// compiled on the host to measure its error.
#pragma once

#ifndef NLC_EXP_QUAL
#define NLC_EXP_QUAL __device__ __forceinline__
#endif

namespace nlc {

constexpr int kExpTabSize = 32;

// 2^(j/32), j = 0..31, rounded to nearest double.
#define NLC_EXP2_TABLE                                                                                              \
	{0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,                        \
	 0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,                        \
	 0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,                        \
	 0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,                        \
	 0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,                        \
	 0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,                        \
	 0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,                        \
	 0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0}

constexpr double kExpInvStep = 0x1.71547652b82fep+5; // 32/ln2
constexpr double kExpStepHi = 0x1.62e42fefa39efp-6;  // ln2/32 rounded
constexpr double kExpStepLo = 0x1.abc9e3b39803fp-61; // ln2/32 - kExpStepHi, rounded
constexpr double kExpShifter = 0x1.8p52;             // 1.5 * 2^52

NLC_EXP_QUAL double exp_tab(double y, const double *tab) {
	const double s = __builtin_fma(y, kExpInvStep, kExpShifter);
	const double nd = s - kExpShifter; // N as a double, exact
	// low 32 bits of s's significand hold N in two's complement (|N| < 2^31)
	union {
		double d;
		unsigned long long u;
	} bits = {s};
	const int n = int(unsigned(bits.u));
	double r = __builtin_fma(-nd, kExpStepHi, y);
	r = __builtin_fma(-nd, kExpStepLo, r);
	double t = __builtin_fma(r, 0x1.6c16c16c16c17p-10, 0x1.1111111111111p-7); // 1/720, 1/120
	t = __builtin_fma(t, r, 0x1.5555555555555p-5);                             // 1/24
	t = __builtin_fma(t, r, 0x1.5555555555555p-3);                             // 1/6
	t = __builtin_fma(t, r, 0.5);
	t = __builtin_fma(t, r, 1.0);
	const double q = t * r; // e^r - 1
	const double T = tab[n & (kExpTabSize - 1)];
	return __builtin_ldexp(__builtin_fma(T, q, T), n >> 5);
}

// Table-free variant: y = N ln2 + r, |r| <= ln2/2, e^r by its degree-13 Taylor
// polynomial (truncation below 5e-18 relative), 2^N by ldexp.  18 VALU and no
// memory access: with the table, the scheduler hoists the table loads of a
// whole unrolled run of bonds to hide their latency and the kernel loses
// occupancy; this form keeps a chain of bonds in few registers.
constexpr double kExpInvLn2 = 0x1.71547652b82fep+0; // 1/ln2
constexpr double kExpLn2Hi = 0x1.62e42fefa39efp-1;  // ln2 rounded
constexpr double kExpLn2Lo = 0x1.abc9e3b39803fp-56; // ln2 - kExpLn2Hi, rounded

NLC_EXP_QUAL double exp_poly(double y) {
	const double s = __builtin_fma(y, kExpInvLn2, kExpShifter);
	const double nd = s - kExpShifter;
	union {
		double d;
		unsigned long long u;
	} bits = {s};
	const int n = int(unsigned(bits.u));
	double r = __builtin_fma(-nd, kExpLn2Hi, y);
	r = __builtin_fma(-nd, kExpLn2Lo, r);
	double t = __builtin_fma(r, 0x1.6124613a86d09p-33, 0x1.1eed8eff8d898p-29); // 1/13!, 1/12!
	t = __builtin_fma(t, r, 0x1.ae64567f544e4p-26);                            // 1/11!
	t = __builtin_fma(t, r, 0x1.27e4fb7789f5cp-22);                            // 1/10!
	t = __builtin_fma(t, r, 0x1.71de3a556c734p-19);                            // 1/9!
	t = __builtin_fma(t, r, 0x1.a01a01a01a01ap-16);                            // 1/8!
	t = __builtin_fma(t, r, 0x1.a01a01a01a01ap-13);                            // 1/7!
	t = __builtin_fma(t, r, 0x1.6c16c16c16c17p-10);                            // 1/6!
	t = __builtin_fma(t, r, 0x1.1111111111111p-7);                             // 1/5!
	t = __builtin_fma(t, r, 0x1.5555555555555p-5);                             // 1/4!
	t = __builtin_fma(t, r, 0x1.5555555555555p-3);                             // 1/3!
	t = __builtin_fma(t, r, 0.5);
	t = __builtin_fma(t, r, 1.0);
	return __builtin_ldexp(__builtin_fma(t, r, 1.0), n);
}

} // namespace nlc
