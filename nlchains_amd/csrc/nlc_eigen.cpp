// dDNKG eigensystem on the device: rocSOLVER's dsyevd (Householder
// tridiagonalisation and divide and conquer, in FP64), the counterpart of the
// reference's Armadillo eig_sym (dDNKG/eigensystem.cpp:13-36).  rocSOLVER and
// rocBLAS are loaded with dlopen on first use, so the library has no link-time
// dependency on them; without a device or without the libraries the host
// solver of nlc_host.cpp is used.
#include <dlfcn.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace nlc {
namespace {

struct Solver {
	decltype(&rocblas_create_handle) create = nullptr;
	decltype(&rocblas_destroy_handle) destroy = nullptr;
	decltype(&rocsolver_dsyevd) syevd = nullptr;
	bool ok = false;
};

void *open_first(const char *a, const char *b) {
	void *h = dlopen(a, RTLD_LAZY | RTLD_LOCAL);
	return h ? h : dlopen(b, RTLD_LAZY | RTLD_LOCAL);
}

const Solver &solver() {
	static const Solver s = [] {
		Solver r;
		void *blas = open_first("librocblas.so.5", "librocblas.so");
		void *sol = blas ? open_first("librocsolver.so.0", "librocsolver.so") : nullptr;
		if (!sol) return r;
		r.create = reinterpret_cast<decltype(r.create)>(dlsym(blas, "rocblas_create_handle"));
		r.destroy = reinterpret_cast<decltype(r.destroy)>(dlsym(blas, "rocblas_destroy_handle"));
		r.syevd = reinterpret_cast<decltype(r.syevd)>(dlsym(sol, "rocsolver_dsyevd"));
		r.ok = r.create && r.destroy && r.syevd;
		return r;
	}();
	return s;
}

template <class T>
struct DevBuf {
	T *p = nullptr;
	explicit DevBuf(size_t count) {
		if (hipMalloc((void **) &p, sizeof(T) * count) != hipSuccess) p = nullptr;
	}
	~DevBuf() {
		if (p) (void) hipFree(p);
	}
};

} // namespace

// omega_j = sqrt(lambda_j) ascending and V column-major (column j = mode j), as
// the host solver.  Returns 0 on success, nonzero when the device path is not
// available or did not converge (the caller then uses the host solver).
int eigensystem_device(uint32_t n, const double *mp2, double *omegas, double *V) {
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return -1;
	const Solver &s = solver();
	if (!s.ok) return -1;
	// A = diag(m_x + 2) - adjacency, periodic corners (dDNKG/eigensystem.cpp:15-18)
	std::vector<double> a(size_t(n) * n, 0.0), d(n);
	for (uint32_t i = 0; i < n; ++i) {
		a[size_t(i) * n + i] = mp2[i];
		if (i + 1 < n) a[size_t(i + 1) * n + i] = a[size_t(i) * n + i + 1] = -1.0;
	}
	a[size_t(n - 1) * n + 0] = a[0 * size_t(n) + n - 1] = -1.0;
	DevBuf<double> dA(a.size()), dD(n), dE(n);
	DevBuf<rocblas_int> dinfo(1);
	if (!dA.p || !dD.p || !dE.p || !dinfo.p) return -1;
	if (hipMemcpy(dA.p, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice) != hipSuccess) return -1;
	rocblas_handle h = nullptr;
	if (s.create(&h) != rocblas_status_success) return -1;
	const rocblas_status st =
	    s.syevd(h, rocblas_evect_original, rocblas_fill_upper, rocblas_int(n), dA.p, rocblas_int(n), dD.p, dE.p, dinfo.p);
	rocblas_int info = -1;
	const bool copied = st == rocblas_status_success &&
	                    hipMemcpy(&info, dinfo.p, sizeof info, hipMemcpyDeviceToHost) == hipSuccess &&
	                    hipMemcpy(d.data(), dD.p, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess &&
	                    hipMemcpy(V, dA.p, sizeof(double) * a.size(), hipMemcpyDeviceToHost) == hipSuccess;
	s.destroy(h);
	if (!copied || info != 0) return -1;
	for (uint32_t j = 0; j < n; ++j) omegas[j] = std::sqrt(d[j]); // omega = sqrt(lambda), dDNKG/eigensystem.cpp:20
	return 0;
}

} // namespace nlc
