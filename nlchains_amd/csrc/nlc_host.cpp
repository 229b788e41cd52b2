// Host-side helpers of libnlchains_amd: Yoshida weights, dispersion relations,
// the dDNKG eigensystem and the seeded initial-condition generators.  Plain
// C++ (no device code); exported through include/nlc.h.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "../../include/nlc.h"

namespace nlc {

// Yoshida (1990) 6th-order weights, common/symplectic.hpp:5-8: c[8] drift, d[8] kick.
extern const double kYoshidaC[8];
extern const double kYoshidaD[8];
const double kYoshidaC[8] = {0.392256805238780,  0.510043411918458,  -0.471053385409757, 0.068753168252518,
                             0.068753168252518,  -0.471053385409757, 0.510043411918458,  0.392256805238780};
const double kYoshidaD[8] = {0.784513610477560,  0.235573213359357, -1.177679984178870, 1.315186320683906,
                             -1.177679984178870, 0.235573213359357, 0.784513610477560,  0};

namespace {

// Parallel for over [0, count) on the host's hardware threads.
template <class F>
void parallel_for(uint32_t count, F &&f) {
	unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 32u));
	if (count < 2 * nt) nt = 1;
	if (nt == 1) {
		for (uint32_t i = 0; i < count; ++i) f(i);
		return;
	}
	std::vector<std::thread> pool;
	for (unsigned t = 0; t < nt; ++t)
		pool.emplace_back([&, t] {
			for (uint32_t i = t; i < count; i += nt) f(i);
		});
	for (auto &th : pool) th.join();
}

// SplitMix64 (Steele, Lea, Flood 2014): a 64-bit counter-based generator.
struct SplitMix64 {
	uint64_t s;
	explicit SplitMix64(uint64_t seed) : s(seed) {}
	uint64_t next() {
		uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
		return z ^ (z >> 31);
	}
	double uniform() { return double(next() >> 11) * 0x1.0p-53; } // [0, 1), 53 bits
};

// Independent stream per copy: seed mixed with the copy index.
SplitMix64 copy_stream(uint64_t seed, uint64_t copy) {
	SplitMix64 h(seed ^ (0xD1B54A32D192ED03ULL * (copy + 1)));
	return SplitMix64(h.next());
}

// cos/sin(2 pi j / n), j < n
void unit_roots(uint32_t n, std::vector<double> &cs, std::vector<double> &sn) {
	cs.resize(n);
	sn.resize(n);
	for (uint32_t j = 0; j < n; ++j) {
		const double ang = 2.0 * M_PI * double(j) / double(n);
		cs[j] = std::cos(ang);
		sn[j] = std::sin(ang);
	}
}

// Householder reduction of the symmetric matrix a (column-major n x n) to
// tridiagonal form T = Q^T A Q (Golub & Van Loan, Matrix Computations, 8.3.1).
// On return d = diag(T), e[i] = T(i+1, i) (e[n-1] = 0) and q = Q (column-major).
void tridiagonalize(uint32_t n, std::vector<double> &a, std::vector<double> &q, std::vector<double> &d,
                    std::vector<double> &e) {
	auto A = [&](uint32_t i, uint32_t j) -> double & { return a[size_t(j) * n + i]; };
	std::vector<std::vector<double>> reflectors(n > 2 ? n - 2 : 0);
	std::vector<double> v(n), w(n);
	for (uint32_t k = 0; k + 2 < n; ++k) {
		// reflector mapping x = A(k+1.., k) onto alpha e_1
		double norm2 = 0;
		for (uint32_t i = k + 1; i < n; ++i) norm2 += A(i, k) * A(i, k);
		const double x0 = A(k + 1, k);
		const double alpha = -std::copysign(std::sqrt(norm2), x0);
		double vv = 0;
		for (uint32_t i = k + 1; i < n; ++i) {
			v[i] = A(i, k) - (i == k + 1 ? alpha : 0.0);
			vv += v[i] * v[i];
		}
		auto &ref = reflectors[k];
		if (vv == 0) continue; // column already reduced
		const double tau = 2.0 / vv;
		ref.assign(v.begin() + k + 1, v.end());
		// symmetric rank-2 update of the trailing block: A <- H A H
		for (uint32_t i = k + 1; i < n; ++i) {
			double s = 0;
			for (uint32_t j = k + 1; j < n; ++j) s += A(i, j) * v[j];
			w[i] = tau * s;
		}
		double kk = 0;
		for (uint32_t i = k + 1; i < n; ++i) kk += v[i] * w[i];
		kk *= 0.5 * tau;
		for (uint32_t i = k + 1; i < n; ++i) w[i] -= kk * v[i];
		for (uint32_t j = k + 1; j < n; ++j)
			for (uint32_t i = k + 1; i < n; ++i) A(i, j) -= v[i] * w[j] + w[i] * v[j];
		A(k + 1, k) = A(k, k + 1) = alpha;
		for (uint32_t i = k + 2; i < n; ++i) A(i, k) = A(k, i) = 0.0;
	}
	d.assign(n, 0.0);
	e.assign(n, 0.0);
	for (uint32_t i = 0; i < n; ++i) d[i] = A(i, i);
	for (uint32_t i = 0; i + 1 < n; ++i) e[i] = A(i + 1, i);
	// Q = H_0 H_1 ... H_{n-3}, accumulated backwards onto the identity
	q.assign(size_t(n) * n, 0.0);
	for (uint32_t i = 0; i < n; ++i) q[size_t(i) * n + i] = 1.0;
	for (int64_t k = int64_t(n) - 3; k >= 0; --k) {
		const auto &ref = reflectors[k];
		if (ref.empty()) continue;
		double vv = 0;
		for (double x : ref) vv += x * x;
		const double tau = 2.0 / vv;
		for (uint32_t j = uint32_t(k) + 1; j < n; ++j) {
			double s = 0;
			for (uint32_t i = uint32_t(k) + 1; i < n; ++i) s += ref[i - k - 1] * q[size_t(j) * n + i];
			s *= tau;
			for (uint32_t i = uint32_t(k) + 1; i < n; ++i) q[size_t(j) * n + i] -= s * ref[i - k - 1];
		}
	}
}

// Largest chain length solved on the host when a device is present.
constexpr uint32_t kHostEigenMax = 512;

// Symmetric QR algorithm on the tridiagonal (d, e) (Golub & Van Loan, Matrix
// Computations, 4th ed., Algorithms 8.3.2 and 8.3.3).  The active block
// [lo, hi] is the trailing unreduced part: off-diagonals below the tolerance
// are set to zero, which splits the matrix, and the block shrinks from the
// bottom as its last eigenvalue converges.  Each sweep is one implicit QR step
// with the Wilkinson shift: a plane rotation in (lo, lo + 1) that matches the
// first column of T - mu I, then rotations (k, k + 1) that chase the bulge it
// creates at T(k - 1, k + 1) down and out of the block.  Every rotation G is
// applied as T <- G^T T G to the 2 x 2 diagonal block it touches and to the
// two off-diagonal entries beside it, and accumulated into the eigenvectors as
// Z <- Z G (columns of z, column-major n x n).
struct Rotation {
	double c, s; // G^T (x, y) = (hypot(x, y), 0) with G^T = [c s; -s c]
	Rotation(double x, double y) {
		const double r = std::hypot(x, y);
		if (r == 0.0) c = 1.0, s = 0.0;
		else c = x / r, s = y / r;
	}
};

// The eigenvalue of [[d0, b], [b, d1]] nearer to d1 (Wilkinson's shift).
double wilkinson_shift(double d0, double b, double d1) {
	const double half_gap = 0.5 * (d0 - d1);
	if (b == 0.0) return d1;
	return d1 - b * b / (half_gap + std::copysign(std::hypot(half_gap, b), half_gap));
}

void qr_sweep(uint32_t n, uint32_t lo, uint32_t hi, std::vector<double> &d, std::vector<double> &e,
              std::vector<double> &z) {
	const double mu = wilkinson_shift(d[hi - 1], e[hi - 1], d[hi]);
	double x = d[lo] - mu, bulge = e[lo];
	for (uint32_t k = lo; k < hi; ++k) {
		const Rotation g(x, bulge);
		const double c = g.c, s = g.s;
		// columns k, k+1 of row k-1: (x, bulge) -> (r, 0)
		if (k > lo) e[k - 1] = c * x + s * bulge;
		// the 2 x 2 diagonal block of rows and columns k, k+1
		const double a = d[k], b = e[k], f = d[k + 1];
		const double cc = c * c, ss = s * s, cs = c * s;
		d[k] = cc * a + 2.0 * cs * b + ss * f;
		d[k + 1] = ss * a - 2.0 * cs * b + cc * f;
		e[k] = cs * (f - a) + (cc - ss) * b;
		if (k + 1 < hi) {
			// rows k, k+1 of column k+2: (0, e[k+1]) -> (s e[k+1], c e[k+1]); the
			// first is the new bulge, removed by the next rotation
			bulge = s * e[k + 1];
			e[k + 1] *= c;
			x = e[k];
		}
		double *zk = &z[size_t(k) * n], *zk1 = &z[size_t(k + 1) * n];
		for (uint32_t i = 0; i < n; ++i) {
			const double u = zk[i], v = zk1[i];
			zk[i] = c * u + s * v;
			zk1[i] = c * v - s * u;
		}
	}
}

bool negligible(double b, double d0, double d1) {
	return std::fabs(b) <= std::numeric_limits<double>::epsilon() * (std::fabs(d0) + std::fabs(d1));
}

bool tridiagonal_qr(uint32_t n, std::vector<double> &d, std::vector<double> &e, std::vector<double> &z) {
	const uint64_t max_sweeps = 30ull * n;
	uint64_t sweeps = 0;
	uint32_t hi = n - 1;
	while (hi > 0) {
		if (negligible(e[hi - 1], d[hi - 1], d[hi])) {
			e[hi - 1] = 0.0; // d[hi] has converged
			--hi;
			continue;
		}
		uint32_t lo = hi - 1;
		while (lo > 0 && !negligible(e[lo - 1], d[lo - 1], d[lo])) --lo;
		if (lo > 0) e[lo - 1] = 0.0;
		if (++sweeps > max_sweeps) return false;
		qr_sweep(n, lo, hi, d, e, z);
	}
	return true;
}

} // namespace

// rocSOLVER on the device (nlc_eigen.cpp); nonzero when unavailable.
int eigensystem_device(uint32_t n, const double *mp2, double *omegas, double *V);

int eigensystem(uint32_t n, const double *mp2, double *omegas, double *V) {
	if (n < 2 || !mp2 || !omegas || !V) return NLC_EARG;
	// long chains: the O(n^3) host solver takes seconds from n = 1024 on
	if (n > kHostEigenMax && eigensystem_device(n, mp2, omegas, V) == 0) return NLC_OK;
	// A = diag(m_x + 2) - adjacency, periodic corners (dDNKG/eigensystem.cpp:15-18)
	std::vector<double> a(size_t(n) * n, 0.0), q, d, e;
	for (uint32_t i = 0; i < n; ++i) {
		a[size_t(i) * n + i] = mp2[i];
		if (i + 1 < n) a[size_t(i + 1) * n + i] = a[size_t(i) * n + i + 1] = -1.0;
	}
	a[size_t(n - 1) * n + 0] = a[0 * size_t(n) + n - 1] = -1.0;
	tridiagonalize(n, a, q, d, e);
	if (!tridiagonal_qr(n, d, e, q)) return NLC_EUNSUPPORTED;
	std::vector<uint32_t> ord(n);
	for (uint32_t i = 0; i < n; ++i) ord[i] = i;
	std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return d[x] < d[y]; });
	for (uint32_t j = 0; j < n; ++j) {
		omegas[j] = std::sqrt(d[ord[j]]); // omega = sqrt(lambda), dDNKG/eigensystem.cpp:20
		std::memcpy(V + size_t(j) * n, &q[size_t(ord[j]) * n], sizeof(double) * n);
	}
	return NLC_OK;
}

} // namespace nlc

using namespace nlc;

extern "C" {

int nlc_yoshida(double *c, double *d) {
	if (!c || !d) return NLC_EARG;
	std::memcpy(c, kYoshidaC, sizeof kYoshidaC);
	std::memcpy(d, kYoshidaD, sizeof kYoshidaD);
	return NLC_OK;
}

int nlc_dispersion(int model, uint32_t n, double m, double *omega) {
	if (!omega || n < 2) return NLC_EARG;
	for (uint32_t k = 0; k < n; ++k) {
		const double s2 = 2 * std::sin(k * M_PI / n);
		if (model == NLC_DNLS) omega[k] = s2 * s2;              // DNLS/dispersion.cpp:10-17
		else if (model == NLC_DNKG || model == NLC_FPUT || model == NLC_TODA)
			omega[k] = std::sqrt((model == NLC_DNKG ? m : 0.0) + s2 * s2); // DNKG_FPUT_Toda/dispersion.cpp:9-16
		else return NLC_EARG;
	}
	return NLC_OK;
}

int nlc_eigensystem(uint32_t n, const double *mp2, double *omegas, double *eigenvectors) {
	return eigensystem(n, mp2, omegas, eigenvectors);
}

int nlc_ic_plane_waves(uint32_t n, uint32_t copies, double m, uint32_t k_lo, uint32_t k_hi, double e_density,
                       uint64_t seed, double *state) {
	if (n < 2 || !copies || !state || k_lo >= k_hi || k_hi > n || e_density < 0) return NLC_EARG;
	std::vector<double> cs, sn, omega(n);
	unit_roots(n, cs, sn);
	for (uint32_t k = 0; k < n; ++k) {
		const double s2 = 2 * std::sin(k * M_PI / n);
		omega[k] = std::sqrt(m + s2 * s2);
	}
	const double e_mode = e_density * n / double(k_hi - k_lo);
	parallel_for(copies, [&](uint32_t c) {
		SplitMix64 rng = copy_stream(seed, c);
		double *chain = state + size_t(c) * n * 2;
		for (uint32_t x = 0; x < 2 * n; ++x) chain[x] = 0.0;
		for (uint32_t k = k_lo; k < k_hi; ++k) {
			const double phi = 2.0 * M_PI * rng.uniform();
			if (!(omega[k] > 0)) continue;
			const double amp = std::sqrt(2.0 * e_mode / (n * omega[k] * omega[k]));
			const double cp = std::cos(phi), sp = std::sin(phi);
			uint64_t j = 0;
			for (uint32_t x = 0; x < n; ++x, j += k) {
				const uint32_t jj = uint32_t(j % n);
				// q = A cos(2 pi k x/n + phi), p = A w sin(2 pi k x/n + phi): a running wave
				chain[2 * x] += amp * (cs[jj] * cp - sn[jj] * sp);
				chain[2 * x + 1] += amp * omega[k] * (sn[jj] * cp + cs[jj] * sp);
			}
		}
	});
	return NLC_OK;
}

int nlc_ic_dnls(uint32_t n, uint32_t copies, uint32_t kmax, double norm_density, uint64_t seed, double *state) {
	if (n < 2 || !copies || !state || !kmax || 2 * kmax - 1 > n || norm_density < 0) return NLC_EARG;
	std::vector<double> cs, sn;
	unit_roots(n, cs, sn);
	const double amp = std::sqrt(norm_density / double(2 * kmax - 1));
	parallel_for(copies, [&](uint32_t c) {
		SplitMix64 rng = copy_stream(seed, c);
		double *chain = state + size_t(c) * n * 2;
		for (uint32_t x = 0; x < 2 * n; ++x) chain[x] = 0.0;
		for (int64_t k = -int64_t(kmax) + 1; k < int64_t(kmax); ++k) {
			const double phi = 2.0 * M_PI * rng.uniform();
			const double cp = std::cos(phi), sp = std::sin(phi);
			const uint64_t kk = uint64_t((k % int64_t(n) + int64_t(n)) % int64_t(n));
			uint64_t j = 0;
			for (uint32_t x = 0; x < n; ++x, j += kk) {
				const uint32_t jj = uint32_t(j % n);
				chain[2 * x] += amp * (cs[jj] * cp - sn[jj] * sp);
				chain[2 * x + 1] += amp * (sn[jj] * cp + cs[jj] * sp);
			}
		}
	});
	return NLC_OK;
}

int nlc_ic_eigenmodes(uint32_t n, uint32_t copies, const double *omegas, const double *V, uint32_t modes,
                      double e_density, uint64_t seed, double *state) {
	if (n < 2 || !copies || !state || !omegas || !V || !modes || modes > n || e_density < 0) return NLC_EARG;
	const double e_mode = e_density * n / double(modes);
	parallel_for(copies, [&](uint32_t c) {
		SplitMix64 rng = copy_stream(seed, c);
		double *chain = state + size_t(c) * n * 2;
		for (uint32_t x = 0; x < 2 * n; ++x) chain[x] = 0.0;
		for (uint32_t j = 0; j < modes; ++j) {
			const double phi = 2.0 * M_PI * rng.uniform();
			if (!(omegas[j] > 0)) continue;
			// standing mode q = A cos(phi) v_j, p = -A w sin(phi) v_j with A^2 w^2 / 2 = e_mode
			const double amp = std::sqrt(2.0 * e_mode) / omegas[j];
			const double cq = amp * std::cos(phi), cp = -amp * omegas[j] * std::sin(phi);
			const double *vj = V + size_t(j) * n;
			for (uint32_t x = 0; x < n; ++x) {
				chain[2 * x] += cq * vj[x];
				chain[2 * x + 1] += cp * vj[x];
			}
		}
	});
	return NLC_OK;
}

int nlc_ic_ragged_lengths(uint32_t nchains, const uint32_t *choices, uint32_t nchoices, uint64_t seed,
                          uint32_t *lengths) {
	if (!nchains || !choices || !nchoices || !lengths) return NLC_EARG;
	SplitMix64 rng = copy_stream(seed, 0xFFFFFFFEULL);
	for (uint32_t i = 0; i < nchains; ++i) lengths[i] = choices[rng.next() % nchoices];
	return NLC_OK;
}

int nlc_ic_ragged_plane_waves(const uint32_t *lengths, uint32_t nchains, double m, double mode_fraction, double e_lo,
                              double e_hi, uint64_t seed, double *state) {
	if (!lengths || !nchains || !state || !(e_hi >= e_lo) || e_lo < 0 || mode_fraction < 0) return NLC_EARG;
	std::vector<uint64_t> offset(nchains + 1, 0);
	for (uint32_t i = 0; i < nchains; ++i) {
		if (lengths[i] < 2) return NLC_EARG;
		offset[i + 1] = offset[i] + lengths[i];
	}
	parallel_for(nchains, [&](uint32_t c) {
		const uint32_t n = lengths[c];
		SplitMix64 rng = copy_stream(seed, c);
		const double dens = e_lo + (e_hi - e_lo) * rng.uniform();
		const uint32_t kmax = std::max<uint32_t>(1, uint32_t(std::floor(mode_fraction * n)));
		double *chain = state + offset[c] * 2;
		for (uint32_t x = 0; x < 2 * n; ++x) chain[x] = 0.0;
		const double e_mode = dens * n / double(kmax);
		for (uint32_t k = 1; k <= kmax && k < n; ++k) {
			const double phi = 2.0 * M_PI * rng.uniform();
			const double s2 = 2 * std::sin(k * M_PI / n), w = std::sqrt(m + s2 * s2);
			if (!(w > 0)) continue;
			const double amp = std::sqrt(2.0 * e_mode / (n * w * w));
			for (uint32_t x = 0; x < n; ++x) {
				const double th = 2.0 * M_PI * double((uint64_t(k) * x) % n) / double(n) + phi;
				chain[2 * x] += amp * std::cos(th);
				chain[2 * x + 1] += amp * w * std::sin(th);
			}
		}
	});
	return NLC_OK;
}

int nlc_ic_masses(uint32_t n, double lo, double hi, uint64_t seed, double *masses) {
	if (n < 1 || !masses || !(hi >= lo)) return NLC_EARG;
	SplitMix64 rng = copy_stream(seed, 0xFFFFFFFFULL);
	for (uint32_t x = 0; x < n; ++x) masses[x] = lo + (hi - lo) * rng.uniform();
	return NLC_OK;
}

} // extern "C"
