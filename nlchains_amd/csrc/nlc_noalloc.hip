// nlc_linenergies_noalloc: the energy pass, normalisation and entropies of
// nlc_linenergies for a planar state already on the device, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  It has no context: every work buffer (the
// roots of unity, the dispersion, the per-block partial sums, the reduction
// scratch and the raw per-mode sums) lives in one file-scope device array,
// `scratch`, of kNoallocScratchBytes per GPU, held until the process ends.
//
// Host code never uses the array's name as an address (in host code it is a
// stand-in symbol, not device memory): the device address is looked up once
// per device with hipGetSymbolAddress, with that device current, and only that
// pointer reaches kernels.  Every call on a GPU shares its array; calls that
// overlap on one GPU race.  Nothing here allocates: the tables are computed by
// a kernel on each call, the partial sums are written (not accumulated) and
// the raw sums are rewritten, so no buffer has to start at zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "../../include/nlc.h"
#include "nlc_internal.hpp"

namespace nlc {
namespace {

constexpr size_t kNoallocScratchBytes = size_t(NLC_NOALLOC_SCRATCH_BYTES);
constexpr size_t kNoallocDoubles = kNoallocScratchBytes / sizeof(double);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoallocDoubles];

std::mutex g_lock;
double *g_scratch[kMaxDevices] = {};

// the message is what nlc_last_error(NULL) returns in this thread
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
	char buf[256];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	set_thread_error(buf);
	return code;
}

// Device address of this device's scratch array (looked up on first use).
int scratch_address(double **out) {
	int dev = -1;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess) return fail(NLC_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
	if (dev < 0 || dev >= kMaxDevices) return fail(NLC_EUNSUPPORTED, "device %d beyond %d devices", dev, kMaxDevices);
	std::lock_guard<std::mutex> lk(g_lock);
	if (!g_scratch[dev]) {
		void *ptr = nullptr;
		e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr)
			return fail(NLC_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
		g_scratch[dev] = static_cast<double *>(ptr);
	}
	*out = g_scratch[dev];
	return NLC_OK;
}

// tw[j] = e^{-2 pi i j/n}; omega[k] the model's dispersion (DNKG_FPUT_Toda/
// dispersion.cpp:9-16, DNLS/dispersion.cpp:10-17), computed with the
// accurate pi-scaled sin/cos.
__global__ void k_noalloc_tables(double2 *tw, double *omega, uint32_t n, int model, double m) {
	for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
		double s, c;
		sincospi(2.0 * double(j) / double(n), &s, &c);
		tw[j] = make_double2(c, -s);
		const double s2 = 2.0 * sinpi(double(j) / double(n));
		omega[j] = model == DNLS ? s2 * s2 : sqrt((model == DNKG ? m : 0.0) + s2 * s2);
	}
}

} // namespace
} // namespace nlc

extern "C" int nlc_linenergies_noalloc(int model, const void *d_state, uint32_t chain_length, uint32_t copies,
                                       double m, double norm, int a0is0, double *d_out, void *stream) {
	using namespace nlc;
	set_thread_error("");
	const uint32_t n = chain_length;
	if (model != NLC_DNKG && model != NLC_FPUT && model != NLC_TODA && model != NLC_DNLS)
		return fail(NLC_EUNSUPPORTED, "nlc_linenergies_noalloc: model %d has no plane-wave energy pass", model);
	if (!d_state || !d_out) return fail(NLC_EARG, "nlc_linenergies_noalloc: state or output pointer is NULL");
	if (n < 8 || n > NLC_NOALLOC_MAX_CHAIN || (n & (n - 1)) || !copies)
		return fail(NLC_EARG,
		            "nlc_linenergies_noalloc: chain_length must be a power of two from 8 to %d and copies >= 1 "
		            "(got %u x %u)",
		            NLC_NOALLOC_MAX_CHAIN, n, copies);
	if (uint64_t(n) * copies > (1ULL << 31))
		return fail(NLC_EARG, "nlc_linenergies_noalloc: chain_length * copies must be <= 2^31");
	double *base = nullptr;
	if (int rc = scratch_address(&base)) return rc;
	// layout in doubles: tw (2n), omega (n), raw sums (n), reduction scratch
	// (ceil(grid/64) rows of n), partial sums (grid rows of n); the grid is the
	// context path's row count when it fits, fewer rows otherwise (the energy
	// kernels stride over the copies)
	const size_t rows_fit = kNoallocDoubles / n - 4;
	const uint32_t grid = std::min<uint32_t>(energy_rows(n, copies), uint32_t(rows_fit * 64 / 65));
	double2 *tw = reinterpret_cast<double2 *>(base);
	double *omega = base + 2 * size_t(n);
	double *e_raw = omega + n;
	double *reduce = e_raw + n;
	double *partial = reduce + size_t(reduce_scratch_rows(grid)) * n;
	if (size_t(partial - base) + size_t(grid) * n > kNoallocDoubles)
		return fail(NLC_EARG, "nlc_linenergies_noalloc: work buffers exceed the %zu-byte scratch", kNoallocScratchBytes);
	if (!(norm > 0))
		norm = model == NLC_DNLS ? 1.0 / (double(n) * copies) // DNLS.cpp:150
		                         : (0.5 / copies) / n;         // DNKG_FPUT_Toda.cpp:129
	hipStream_t st = static_cast<hipStream_t>(stream);
	const double2 *state = static_cast<const double2 *>(d_state);
	hipLaunchKernelGGL(k_noalloc_tables, dim3((n + 255) / 256), dim3(256), 0, st, tw, omega, n, model,
	                   model == NLC_DNKG ? m : 0.0);
	hipError_t e = hipGetLastError();
	LongChainFft none;
	if (e == hipSuccess)
		e = model == NLC_DNLS ? launch_energy_dnls(st, state, n, copies, tw, partial, grid, none)
		                      : launch_energy_real(st, state, n, copies, tw, omega, partial, grid, none);
	if (e == hipSuccess)
		e = launch_reduce_results(st, partial, grid, n, model == NLC_DNLS ? omega : nullptr, e_raw, reduce, norm,
		                          a0is0 ? 1 : 0, nullptr, 0, d_out);
	if (e != hipSuccess) return fail(NLC_EHIP, "nlc_linenergies_noalloc: launch failed: %s", hipGetErrorString(e));
	return NLC_OK;
}
