/*
 * oracle.c -- CPU FP64 restatement of the nlchains hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load the library built from it; the product (nlchains_amd) never does.
 *
 * Every function restates the reference's CPU validation path (the "<model>-cpu"
 * subprograms) in plain C; the reference file:line followed is given at each
 * function.  Paths are relative to the reference tree.
 *
 * The reference CPU path cannot be compiled here (FFTW3, Boost, Armadillo and
 * the CUDA headers are absent), so the third-party arithmetic it relies on is
 * restated from its textbook definition:
 *   - FFTW3 forward/backward DFT  -> orc_fft (radix-2 Cooley-Tukey for powers of
 *     two, direct O(n^2) DFT otherwise), unnormalised, FFTW sign convention;
 *   - Armadillo eig_sym           -> orc_eigensystem (cyclic Jacobi rotations),
 *     eigenvalues ascending, eigenvectors as columns (column-major);
 *   - the DNLS CPU path's sqrt(1-cos^2) sine workaround
 *     (DNLS/cpu-implementation/DNLS.cpp:79-96) is replaced by a true sin(): the
 *     workaround is a vectorisation trick, not part of the method.
 *
 * Pinning (each function is checked by tests/test_oracle.py):
 *   orc_coeffs            literal values + palindrome/sum invariants
 *   orc_move_real         harmonic closed form (2x2 shear products), force vs
 *                         -dH/dq finite differences, energy-drift slope ~ dt^6
 *   orc_fft               numpy.fft fixtures, brute-force DFT, Parseval
 *   orc_linenergies_real  running-wave one-hot, Parseval-type harmonic energy sum
 *   orc_eigensystem       numpy.linalg.eigh fixtures, A V = V L, V^T V = I,
 *                         uniform-mass closed-form spectrum
 *   orc_linenergies_ddnkg harmonic-energy identity, uniform-mass equivalence
 *                         with the plane-wave energies
 *   orc_move_dnls         norm conservation, beta=0 closed form, single-site
 *                         nonlinear phase
 *   orc_entropies         equipartition (0), one-hot and two-mode closed forms,
 *                         the formula of supporting-material/post-processing/nlchains.m:56-57
 *   orc_chain_energy      closed forms on small hand-made chains
 *
 * Build: oracle/build.sh (strict: -O2 -fno-fast-math -ffp-contract=off) and a
 * second "reference-flags" build used only as the CPU timing baseline.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { ORC_DNKG = 0, ORC_FPUT = 1, ORC_TODA = 2, ORC_DDNKG = 3, ORC_DNLS = 4 };

/* Yoshida 6th-order weights: the literal values of common/symplectic.hpp:5-8. */
static const double orc_sym_c[8] = {0.392256805238780, 0.510043411918458, -0.471053385409757, 0.068753168252518,
                                    0.068753168252518, -0.471053385409757, 0.510043411918458, 0.392256805238780};
static const double orc_sym_d[8] = {0.784513610477560, 0.235573213359357, -1.177679984178870, 1.315186320683906,
                                    -1.177679984178870, 0.235573213359357, 0.784513610477560, 0};

/* Raw weights (for the invariant tests). */
void orc_yoshida(double *c, double *d) {
	for (int i = 0; i < 8; i++) c[i] = orc_sym_c[i], d[i] = orc_sym_d[i];
}

/* dt-scaled weights: DNKG_FPUT_Toda/cpu-implementation/DNKG_FPUT_Toda.cpp:86-87. */
void orc_coeffs(double dt, double *dt_c, double *dt_d) {
	for (int i = 0; i < 8; i++) {
		dt_c[i] = orc_sym_c[i] * dt;
		dt_d[i] = orc_sym_d[i] * dt;
	}
}

/* ------------------------------------------------------------------------- */
/* Real models: force (right-hand side of dp/dt)                             */
/* ------------------------------------------------------------------------- */

typedef struct {
	int model;
	double m, alpha, beta, alpha2, alpha2_inv;
	const double *mp2; /* dDNKG: m_x + 2 per site */
} orc_rhs_params;

/*
 * DNKG:  DNKG_FPUT_Toda/cpu-implementation/DNKG_FPUT_Toda.cpp:23-26
 * FPUT:  ...:28-34   (uncached CPU form)
 * Toda:  ...:36-40   (two exponentials, CPU form)
 * dDNKG: dDNKG/cpu-implementation/dDNKG.cpp:17-19
 * The operand order of each expression is kept as written there.
 */
static inline double orc_rhs(const orc_rhs_params *P, double l, double c, double r, uint32_t site) {
	switch (P->model) {
	case ORC_DNKG: {
		double bracket = (P->m + 2) + (P->beta * c) * c;
		return -((c * bracket - l) - r);
	}
	case ORC_DDNKG: {
		double bracket = P->mp2[site] + (P->beta * c) * c;
		return -((c * bracket - l) - r);
	}
	case ORC_FPUT: {
		double lin = (l + r) - 2 * c;
		lin = lin * (1 + P->alpha * (r - l));
		double dr = r - c, dl = c - l;
		return lin + P->beta * (dr * dr * dr - dl * dl * dl);
	}
	case ORC_TODA:
		return P->alpha2_inv * (exp(P->alpha2 * (r - c)) - exp(P->alpha2 * (c - l)));
	}
	return 0;
}

/*
 * The timing build compiles the per-chain integrators for several x86 vector
 * ISAs and picks one at load time (the reference does the same with
 * make_simd_clones, common/utilities.hpp:65-69); the strict build does not.
 */
#if defined(ORC_FAST) && defined(__x86_64__)
#define ORC_HOT __attribute__((target_clones("avx512f", "avx2", "default")))
#else
#define ORC_HOT
#endif

/* One kick sweep of the real models, one loop per model so each vectorises. */
#define ORC_KICK_LOOP(EXPR)                                                                                        \
	for (size_t i = 0; i < (size_t) n; i++) {                                                                         \
		double l = phi[i], c = phi[i + 1], r = phi[i + 2];                                                            \
		(void) l; (void) r;                                                                                           \
		pi[i] += dk * (EXPR);                                                                                         \
	}

ORC_HOT static void orc_chain_real(const orc_rhs_params *P, uint32_t n, double *chain, uint32_t steps,
                                   const double *dt_c, const double *dt_d, double *restrict phi,
                                   double *restrict pi) {
	const double m2 = P->m + 2, beta = P->beta, alpha = P->alpha, a2 = P->alpha2, a2i = P->alpha2_inv;
	const double *mp2 = P->mp2;
	for (uint32_t i = 0; i < n; i++) phi[i + 1] = chain[2 * i], pi[i] = chain[2 * i + 1];
	for (uint32_t s = 0; s < steps; s++) {
		for (int k = 0; k < 7; k++) {
			double dck = dt_c[k], dk = dt_d[k];
			if (s && !k) dck *= 2;
			for (uint32_t i = 0; i < n; i++) phi[i + 1] += dck * pi[i];
			phi[0] = phi[n];
			phi[n + 1] = phi[1];
			switch (P->model) {
			case ORC_DNKG: ORC_KICK_LOOP(-((c * (m2 + (beta * c) * c) - l) - r)) break;
			case ORC_DDNKG: ORC_KICK_LOOP(-((c * (mp2[i] + (beta * c) * c) - l) - r)) break;
			case ORC_FPUT:
				ORC_KICK_LOOP(((l + r) - 2 * c) * (1 + alpha * (r - l)) +
				              beta * ((r - c) * (r - c) * (r - c) - (c - l) * (c - l) * (c - l)))
				break;
			case ORC_TODA: ORC_KICK_LOOP(a2i * (exp(a2 * (r - c)) - exp(a2 * (c - l)))) break;
			}
		}
	}
	for (uint32_t i = 0; i < n; i++) {
		phi[i + 1] += dt_c[7] * pi[i];
		chain[2 * i] = phi[i + 1];
		chain[2 * i + 1] = pi[i];
	}
}

/*
 * Advance `copies` chains of length n by `steps` Yoshida-6 steps.
 * planar: double[copies][n][2] = {q, p} (configuration.hpp:71, README.md:78).
 * Loop structure: DNKG_FPUT_Toda/cpu-implementation/DNKG_FPUT_Toda.cpp:113-147
 * (dDNKG: dDNKG/cpu-implementation/dDNKG.cpp:86-110): drift with dt*c[k]
 * (doubled for k==0 after the first step, merging the c[7] drift of the
 * previous step), periodic halo, kick with dt*d[k]; a final dt*c[7] drift.
 * The force expressions keep the operand order of orc_rhs.
 * mp2: per-site m_x+2 for dDNKG (NULL otherwise).
 * threads > 1 parallelises over copies (OpenMP builds only; each copy is
 * independent, so the result does not depend on it).
 */
int orc_move_real(int model, uint32_t n, uint32_t copies, double *planar, uint32_t steps, double dt, double m,
                  double alpha, double beta, const double *mp2, int threads) {
	if (n < 2 || (model == ORC_DDNKG && !mp2) || model < 0 || model > ORC_DDNKG) return -1;
	orc_rhs_params P = {model, m, alpha, beta, 2 * alpha, 1 / (2 * alpha), mp2};
	double dt_c[8], dt_d[8];
	orc_coeffs(dt, dt_c, dt_d);
	(void) threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
	{
		double *phi = (double *) malloc(sizeof(double) * (n + 2));
		double *pi = (double *) malloc(sizeof(double) * n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
		for (int64_t c = 0; c < (int64_t) copies; c++)
			orc_chain_real(&P, n, planar + (size_t) c * n * 2, steps, dt_c, dt_d, phi, pi);
		free(phi);
		free(pi);
	}
	return 0;
}

/* Force evaluation alone (for the -dH/dq checks). out[i] = F at site i. */
int orc_forces(int model, uint32_t n, const double *q, double m, double alpha, double beta, const double *mp2,
               double *out) {
	if (n < 2) return -1;
	orc_rhs_params P = {model, m, alpha, beta, 2 * alpha, 1 / (2 * alpha), mp2};
	for (uint32_t i = 0; i < n; i++) {
		double l = q[(i + n - 1) % n], r = q[(i + 1) % n];
		out[i] = orc_rhs(&P, l, q[i], r, i);
	}
	return 0;
}

/* ------------------------------------------------------------------------- */
/* DFT (restates what FFTW3 computes; FFTW sign convention, unnormalised)    */
/* ------------------------------------------------------------------------- */

static int orc_is_pow2(uint32_t n) { return n && !(n & (n - 1)); }

/*
 * In-place DFT of (re, im), X_k = sum_x z_x exp(sign * 2 pi i k x / n), sign = -1
 * forward (FFTW_FORWARD), +1 backward.  Twiddles are evaluated directly from
 * cos/sin of the exact angle (no recurrences).
 */
void orc_fft(uint32_t n, double *re, double *im, int sign) {
	if (n < 2) return;
	if (orc_is_pow2(n)) {
		/* bit reversal */
		for (uint32_t i = 1, j = 0; i < n; i++) {
			uint32_t bit = n >> 1;
			for (; j & bit; bit >>= 1) j ^= bit;
			j ^= bit;
			if (i < j) {
				double t = re[i]; re[i] = re[j]; re[j] = t;
				t = im[i]; im[i] = im[j]; im[j] = t;
			}
		}
		for (uint32_t len = 2; len <= n; len <<= 1) {
			uint32_t half = len >> 1, stride = n / len;
			for (uint32_t j = 0; j < half; j++) {
				/* w = exp(sign 2 pi i j/len) = exp(sign 2 pi i (j*stride)/n) */
				double ang = 2 * M_PI * (double) (j * stride) / (double) n;
				double wr = cos(ang), wi = sign * sin(ang);
				for (uint32_t b = 0; b < n; b += len) {
					uint32_t u = b + j, v = u + half;
					double tr = re[v] * wr - im[v] * wi;
					double ti = re[v] * wi + im[v] * wr;
					re[v] = re[u] - tr;
					im[v] = im[u] - ti;
					re[u] = re[u] + tr;
					im[u] = im[u] + ti;
				}
			}
		}
	} else {
		double *tr = (double *) malloc(sizeof(double) * n), *ti = (double *) malloc(sizeof(double) * n);
		double *cs = (double *) malloc(sizeof(double) * n), *sn = (double *) malloc(sizeof(double) * n);
		for (uint32_t j = 0; j < n; j++) {
			double ang = 2 * M_PI * (double) j / (double) n;
			cs[j] = cos(ang);
			sn[j] = sign * sin(ang);
		}
		for (uint32_t k = 0; k < n; k++) {
			double sr = 0, si = 0;
			for (uint32_t x = 0; x < n; x++) {
				uint32_t idx = (uint32_t) (((uint64_t) k * x) % n);
				sr += re[x] * cs[idx] - im[x] * sn[idx];
				si += re[x] * sn[idx] + im[x] * cs[idx];
			}
			tr[k] = sr;
			ti[k] = si;
		}
		memcpy(re, tr, sizeof(double) * n);
		memcpy(im, ti, sizeof(double) * n);
		free(tr); free(ti); free(cs); free(sn);
	}
}

/* ------------------------------------------------------------------------- */
/* Linear-mode energies and entropies                                        */
/* ------------------------------------------------------------------------- */

/* omega_k = sqrt(m + (2 sin(pi k/n))^2): DNKG_FPUT_Toda/dispersion.cpp:9-16. */
void orc_dispersion_real(uint32_t n, double m, double *omega) {
	for (uint32_t k = 0; k < n; k++) {
		double s2 = 2 * sin(k * M_PI / n);
		omega[k] = sqrt(m + s2 * s2);
	}
}

/* omega_k = (2 sin(pi k/n))^2: DNLS/dispersion.cpp:10-17. */
void orc_dispersion_dnls(uint32_t n, double *omega) {
	for (uint32_t k = 0; k < n; k++) {
		double s2 = 2 * sin(k * M_PI / n);
		omega[k] = s2 * s2;
	}
}

/*
 * Raw (unnormalised) per-mode energies of the real plane-wave models, summed
 * over copies: DNKG_FPUT_Toda/cpu-implementation/DNKG_FPUT_Toda.cpp:89-104.
 * e_raw[k] = sum_c w_k (w_k |Q|^2 + 2 (Q_i P_r - Q_r P_i)) + |P|^2 with Q, P the
 * forward DFTs of q and p.  e_raw is overwritten.
 */
void orc_linenergies_real(uint32_t n, uint32_t copies, const double *planar, const double *omega, double *e_raw) {
	double *qr = (double *) malloc(sizeof(double) * n), *qi = (double *) malloc(sizeof(double) * n);
	double *pr = (double *) malloc(sizeof(double) * n), *pim = (double *) malloc(sizeof(double) * n);
	for (uint32_t k = 0; k < n; k++) e_raw[k] = 0;
	for (uint32_t c = 0; c < copies; c++) {
		const double *chain = planar + (size_t) c * n * 2;
		for (uint32_t i = 0; i < n; i++) qr[i] = chain[2 * i], pr[i] = chain[2 * i + 1], qi[i] = pim[i] = 0;
		orc_fft(n, qr, qi, -1);
		orc_fft(n, pr, pim, -1);
		for (uint32_t k = 0; k < n; k++) {
			double e = (qr[k] * qr[k] + qi[k] * qi[k]) * omega[k];
			e += 2 * (qi[k] * pr[k] - qr[k] * pim[k]);
			e *= omega[k];
			e += pr[k] * pr[k] + pim[k] * pim[k];
			e_raw[k] += e;
		}
	}
	free(qr); free(qi); free(pr); free(pim);
}

/*
 * results::calc_linenergies for one process: common/results_entropy.cpp:11-19.
 * e[0] = 0 when a0is0, then e *= norm.
 */
void orc_calc_linenergies(uint32_t n, double *e, double norm, int a0is0) {
	if (a0is0) e[0] = 0;
	for (uint32_t k = 0; k < n; k++) e[k] *= norm;
}

/*
 * Entropy mode set: common/main.cpp:119-128 (mask bytes != 0, at least 2) and
 * common/results.cpp:8-15 (drop mode 0 when a0is0, at least 2 left).
 * Returns the number of indices written to idx, or -1 when too few modes.
 * mask == NULL: empty set (all modes are used by orc_entropies).
 */
int orc_entropy_modes(uint32_t n, const uint8_t *mask, int a0is0, uint32_t *idx) {
	if (!mask) return 0;
	int cnt = 0;
	for (uint32_t k = 0; k < n; k++)
		if (mask[k]) idx[cnt++] = k;
	if (cnt < 2) return -1;
	if (a0is0 && idx[0] == 0) {
		memmove(idx, idx + 1, sizeof(uint32_t) * (cnt - 1));
		cnt--;
		if (cnt < 2) return -1;
	}
	return cnt;
}

/*
 * WT and information entropies: common/results_entropy.cpp:21-46.
 * Over the mode set M (idx/nidx, or all modes but 0 when a0is0 if nidx == 0):
 *   E = sum e, S_WT = -(sum log e + |M| log(|M|/E)),
 *   S_INF = (|M|/E) (sum e log e + E log(|M|/E)).
 */
void orc_entropies(uint32_t n, const double *e, int a0is0, const uint32_t *idx, uint32_t nidx, double *wt,
                   double *inf) {
	double tot = 0, totlog = 0, totelog = 0;
	uint32_t cnt;
	if (!nidx) {
		cnt = n - (a0is0 ? 1 : 0);
		for (uint32_t k = a0is0 ? 1 : 0; k < n; k++) {
			double lg = log(e[k]);
			tot += e[k];
			totlog += lg;
			totelog += e[k] * lg;
		}
	} else {
		cnt = nidx;
		for (uint32_t j = 0; j < nidx; j++) {
			double ek = e[idx[j]], lg = log(ek);
			tot += ek;
			totlog += lg;
			totelog += ek * lg;
		}
	}
	double normalization = cnt / tot, lognorm = log(normalization);
	*wt = -(totlog + cnt * lognorm);
	*inf = normalization * (totelog + lognorm * tot);
}

/* ------------------------------------------------------------------------- */
/* dDNKG: eigensystem and projection energies                                */
/* ------------------------------------------------------------------------- */

/*
 * Eigensystem of the mass-weighted periodic Laplacian A = diag(mp2) - adjacency,
 * with A(0,n-1) = A(n-1,0) = -1: dDNKG/eigensystem.cpp:15-21 (Armadillo eig_sym,
 * omega = sqrt(lambda)).  Restated with cyclic Jacobi rotations (Golub & Van
 * Loan, Matrix Computations, section 8.5), then sorted ascending.
 * V is column-major n x n (column j = mode j), as dumped to "-eigenvectors".
 * Returns the number of sweeps, or -1 if not converged.
 */
int orc_eigensystem(uint32_t n, const double *mp2, double *omega, double *V) {
	double *A = (double *) calloc((size_t) n * n, sizeof(double));
	double *W = (double *) calloc((size_t) n * n, sizeof(double));
#define AA(i, j) A[(size_t) (j) * n + (i)]
#define WW(i, j) W[(size_t) (j) * n + (i)]
	for (uint32_t i = 0; i < n; i++) {
		AA(i, i) = mp2[i];
		WW(i, i) = 1;
		if (i + 1 < n) AA(i, i + 1) = AA(i + 1, i) = -1;
	}
	AA(0, n - 1) = AA(n - 1, 0) = -1;
	int sweep, converged = 0;
	for (sweep = 0; sweep < 100 && !converged; sweep++) {
		double off = 0, diag = 0;
		for (uint32_t j = 0; j < n; j++)
			for (uint32_t i = 0; i < n; i++) {
				if (i == j) diag += AA(i, i) * AA(i, i);
				else off += AA(i, j) * AA(i, j);
			}
		if (off <= 1e-30 * diag) { converged = 1; break; }
		for (uint32_t p = 0; p + 1 < n; p++)
			for (uint32_t q = p + 1; q < n; q++) {
				double apq = AA(p, q);
				if (apq == 0) continue;
				double app = AA(p, p), aqq = AA(q, q);
				double theta = (aqq - app) / (2 * apq);
				double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
				double cs = 1 / sqrt(t * t + 1), sn = t * cs;
				/* A <- J^T A J on rows/cols p, q */
				for (uint32_t k = 0; k < n; k++) {
					double akp = AA(k, p), akq = AA(k, q);
					AA(k, p) = cs * akp - sn * akq;
					AA(k, q) = sn * akp + cs * akq;
				}
				for (uint32_t k = 0; k < n; k++) {
					double apk = AA(p, k), aqk = AA(q, k);
					AA(p, k) = cs * apk - sn * aqk;
					AA(q, k) = sn * apk + cs * aqk;
				}
				AA(p, q) = AA(q, p) = 0;
				for (uint32_t k = 0; k < n; k++) {
					double wkp = WW(k, p), wkq = WW(k, q);
					WW(k, p) = cs * wkp - sn * wkq;
					WW(k, q) = sn * wkp + cs * wkq;
				}
			}
	}
	/* sort ascending (selection sort on indices) */
	uint32_t *ord = (uint32_t *) malloc(sizeof(uint32_t) * n);
	for (uint32_t i = 0; i < n; i++) ord[i] = i;
	for (uint32_t i = 0; i < n; i++) {
		uint32_t best = i;
		for (uint32_t j = i + 1; j < n; j++)
			if (AA(ord[j], ord[j]) < AA(ord[best], ord[best])) best = j;
		uint32_t t = ord[i]; ord[i] = ord[best]; ord[best] = t;
	}
	for (uint32_t j = 0; j < n; j++) {
		double lam = AA(ord[j], ord[j]);
		omega[j] = sqrt(lam);
		for (uint32_t i = 0; i < n; i++) V[(size_t) j * n + i] = WW(i, ord[j]);
	}
#undef AA
#undef WW
	free(ord);
	free(A);
	free(W);
	return converged ? sweep : -1;
}

/*
 * dDNKG projection energies: dDNKG/cpu-implementation/dDNKG.cpp:58-72.
 * e_raw[j] = sum_c (omega_j <q_c, V_j>)^2 + <p_c, V_j>^2.  Overwrites e_raw.
 */
void orc_linenergies_ddnkg(uint32_t n, uint32_t copies, const double *planar, const double *V, const double *omega,
                           double *e_raw) {
	for (uint32_t j = 0; j < n; j++) e_raw[j] = 0;
	for (uint32_t c = 0; c < copies; c++) {
		const double *chain = planar + (size_t) c * n * 2;
		for (uint32_t j = 0; j < n; j++) {
			const double *vj = V + (size_t) j * n;
			double dq = 0, dp = 0;
			for (uint32_t i = 0; i < n; i++) {
				dq += chain[2 * i] * vj[i];
				dp += chain[2 * i + 1] * vj[i];
			}
			double wq = omega[j] * dq;
			e_raw[j] += wq * wq + dp * dp;
		}
	}
}

/* ------------------------------------------------------------------------- */
/* DNLS split-step                                                           */
/* ------------------------------------------------------------------------- */

/*
 * Linear Fourier multipliers T[k][j] = exp(-i dt d_k omega_j) / n, k = 0..6:
 * DNLS/dispersion.cpp:19-26.  tab: double[7][n][2].
 */
void orc_dnls_tables(uint32_t n, double dt, double *tab) {
	double *omega = (double *) malloc(sizeof(double) * n);
	orc_dispersion_dnls(n, omega);
	double nrm = 1.0 / n;
	for (int k = 0; k < 7; k++)
		for (uint32_t j = 0; j < n; j++) {
			double ang = ((-dt) * orc_sym_d[k]) * omega[j];
			tab[((size_t) k * n + j) * 2] = cos(ang) * nrm;
			tab[((size_t) k * n + j) * 2 + 1] = sin(ang) * nrm;
		}
	free(omega);
}

/* psi <- psi * exp(-i theta |psi|^2): DNLS/cpu-implementation/DNLS.cpp:75-99 (true sin). */
static void orc_dnls_nonlinear(uint32_t n, double *re, double *im, double bdt) {
	for (uint32_t i = 0; i < n; i++) {
		double nrm = re[i] * re[i] + im[i] * im[i];
		double ang = -bdt * nrm;
		double er = cos(ang), ei = sin(ang);
		double r3 = re[i] * er - im[i] * ei, i3 = ei * re[i] + im[i] * er;
		re[i] = r3;
		im[i] = i3;
	}
}

/*
 * DNLS split step: DNLS/cpu-implementation/DNLS.cpp:110-128.  For each step and
 * k = 0..6: nonlinear rotation by beta dt c_k (2 c_0 merged after the first
 * step), forward DFT, multiply by T[k], backward DFT; after the batch a final
 * beta dt c_7 rotation.  psi: double[copies][n][2] (complex<double>).
 */
int orc_move_dnls(uint32_t n, uint32_t copies, double *psi, uint32_t steps, double dt, double beta, int threads) {
	if (n < 2) return -1;
	double *tab = (double *) malloc(sizeof(double) * 14 * n);
	orc_dnls_tables(n, dt, tab);
	(void) threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
	{
		double *re = (double *) malloc(sizeof(double) * n), *im = (double *) malloc(sizeof(double) * n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
		for (int64_t c = 0; c < (int64_t) copies; c++) {
			double *chain = psi + (size_t) c * n * 2;
			for (uint32_t i = 0; i < n; i++) re[i] = chain[2 * i], im[i] = chain[2 * i + 1];
			for (uint32_t s = 0; s < steps; s++)
				for (int k = 0; k < 7; k++) {
					orc_dnls_nonlinear(n, re, im, ((beta * dt) * (!k && s ? 2. : 1.)) * orc_sym_c[k]);
					orc_fft(n, re, im, -1);
					const double *t = tab + (size_t) k * n * 2;
					for (uint32_t i = 0; i < n; i++) {
						double r3 = re[i] * t[2 * i] - im[i] * t[2 * i + 1];
						double i3 = t[2 * i + 1] * re[i] + im[i] * t[2 * i];
						re[i] = r3;
						im[i] = i3;
					}
					orc_fft(n, re, im, +1);
				}
			orc_dnls_nonlinear(n, re, im, (beta * dt) * orc_sym_c[7]);
			for (uint32_t i = 0; i < n; i++) chain[2 * i] = re[i], chain[2 * i + 1] = im[i];
		}
		free(re);
		free(im);
	}
	free(tab);
	return 0;
}

/*
 * DNLS raw energies: DNLS/cpu-implementation/DNLS.cpp:66-73,103:
 * e_raw[k] = omega_k sum_c |psi_hat_k|^2 (forward DFT).  Overwrites e_raw.
 */
void orc_linenergies_dnls(uint32_t n, uint32_t copies, const double *psi, double *e_raw) {
	double *re = (double *) malloc(sizeof(double) * n), *im = (double *) malloc(sizeof(double) * n);
	double *omega = (double *) malloc(sizeof(double) * n);
	orc_dispersion_dnls(n, omega);
	for (uint32_t k = 0; k < n; k++) e_raw[k] = 0;
	for (uint32_t c = 0; c < copies; c++) {
		const double *chain = psi + (size_t) c * n * 2;
		for (uint32_t i = 0; i < n; i++) re[i] = chain[2 * i], im[i] = chain[2 * i + 1];
		orc_fft(n, re, im, -1);
		for (uint32_t k = 0; k < n; k++) e_raw[k] += re[k] * re[k] + im[k] * im[k];
	}
	for (uint32_t k = 0; k < n; k++) e_raw[k] *= omega[k];
	free(re); free(im); free(omega);
}

/* ------------------------------------------------------------------------- */
/* Total energy per chain: supporting-material/post-processing/nlchains.m:40-48 */
/* ------------------------------------------------------------------------- */

/*
 * E[c] for every copy.  d_x = q_{x+1} - q_x (cyclic).  masses: dDNKG m_x (not
 * m_x + 2).  DNLS: state is psi (complex), E = sum |psi_{x+1}-psi_x|^2 + beta/2 |psi|^4.
 */
void orc_chain_energy(int model, uint32_t n, uint32_t copies, const double *state, double m, double alpha,
                      double beta, const double *masses, double *E) {
	for (uint32_t c = 0; c < copies; c++) {
		const double *s = state + (size_t) c * n * 2;
		double tot = 0;
		for (uint32_t x = 0; x < n; x++) {
			uint32_t y = (x + 1) % n;
			if (model == ORC_DNLS) {
				double dr = s[2 * y] - s[2 * x], di = s[2 * y + 1] - s[2 * x + 1];
				double a2 = s[2 * x] * s[2 * x] + s[2 * x + 1] * s[2 * x + 1];
				tot += dr * dr + di * di + beta / 2 * a2 * a2;
				continue;
			}
			double q = s[2 * x], p = s[2 * x + 1], d = s[2 * y] - q;
			double e = p * p / 2;
			switch (model) {
			case ORC_DNKG: e += d * d / 2 + m * q * q / 2 + beta * q * q * q * q / 4; break;
			case ORC_DDNKG: e += d * d / 2 + masses[x] * q * q / 2 + beta * q * q * q * q / 4; break;
			case ORC_FPUT: e += d * d / 2 + alpha * d * d * d / 3 + beta * d * d * d * d / 4; break;
			case ORC_TODA:
				e += exp(2 * d * alpha) / (4 * alpha * alpha) - d / (2 * alpha) - 1 / (4 * alpha * alpha);
				break;
			}
			tot += e;
		}
		E[c] = tot;
	}
}

/* Norm per DNLS chain: sum_x |psi_x|^2. */
void orc_dnls_norm(uint32_t n, uint32_t copies, const double *psi, double *out) {
	for (uint32_t c = 0; c < copies; c++) {
		const double *s = psi + (size_t) c * n * 2;
		double tot = 0;
		for (uint32_t x = 0; x < n; x++) tot += s[2 * x] * s[2 * x] + s[2 * x + 1] * s[2 * x + 1];
		out[c] = tot;
	}
}

int orc_has_openmp(void) {
#ifdef _OPENMP
	return 1;
#else
	return 0;
#endif
}
