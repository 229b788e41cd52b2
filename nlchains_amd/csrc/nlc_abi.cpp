// C-ABI layer of libnlchains_amd: context life cycle, kernel selection and the
// entry points declared in include/nlc.h.  No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nlc.h"
#include "nlc_internal.hpp"

namespace {

thread_local std::string g_create_error;

} // namespace

struct nlc_ctx {
	nlc_params p{};
	std::vector<double> masses; // dDNKG m_x
	hipStream_t stream = nullptr;
	nlc::StepArgs args{};
	uint32_t n = 0, copies = 0;
	size_t state_bytes = 0;

	double2 *state = nullptr;
	double *omega = nullptr;   // n
	double2 *tw = nullptr;     // n roots of unity exp(-2 pi i j/n)
	double *partial = nullptr; // energy partial sums
	double *reduce_scratch = nullptr;
	uint32_t partial_rows = 0;
	double *e_raw = nullptr;   // n
	double *results = nullptr; // n + 2
	uint32_t *modes = nullptr; // entropy mode indices
	uint32_t nmodes = 0;
	int a0is0 = 0;
	double *mp2 = nullptr;     // dDNKG m_x + 2
	double *masses_d = nullptr;
	double *eigvec = nullptr;  // dDNKG V, column-major
	double *split_q = nullptr, *split_p = nullptr;
	double *energy_d = nullptr, *norm_d = nullptr;
	double2 *dnls_table = nullptr; // DNLS Fourier multipliers
	double theta[9] = {};          // DNLS nonlinear angles
	double *h_results = nullptr; // pinned, n + 2

	nlc::StepLauncher launcher = nullptr;
	int L = 0, S = 0;
	bool exact = false, use_split = false;
	std::string err;

	int fail(int code, const char *fmt, ...) {
		char buf[512];
		va_list ap;
		va_start(ap, fmt);
		vsnprintf(buf, sizeof buf, fmt, ap);
		va_end(ap);
		err = buf;
		return code;
	}
	int hip_fail(hipError_t e, const char *what) {
		return fail(NLC_EHIP, "%s: %s (%s)", what, hipGetErrorName(e), hipGetErrorString(e));
	}
};

#define NLC_HIP(ctx, call)                                                                                            \
	do {                                                                                                             \
		hipError_t e_ = (call);                                                                                        \
		if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #call);                                                       \
	} while (0)

namespace {

bool is_pow2(uint32_t x) { return x && !(x & (x - 1)); }

nlc::StepLauncher lookup(int model, int L, int S, bool exact) {
	switch (model) {
	case NLC_DNKG: return nlc::lookup_step_dnkg(L, S, exact);
	case NLC_FPUT: return nlc::lookup_step_fput(L, S, exact);
	case NLC_TODA: return nlc::lookup_step_toda(L, S, exact);
	case NLC_DDNKG: return nlc::lookup_step_ddnkg(L, S, exact);
	}
	return nullptr;
}

// Default wave shape for a chain length: one chain per wavefront (L = 64) with
// ceil(n/64) sites per lane when n >= 64; for shorter power-of-two chains
// several chains share a wavefront with 8 sites per lane (thread-per-chain
// below 8 sites).
bool default_shape(uint32_t n, int &L, int &S, bool &exact) {
	if (n >= 64) {
		S = int((n + 63) / 64);
		L = 64;
		exact = (uint32_t(S) * 64 == n);
		return S <= 32;
	}
	if (!is_pow2(n)) return false;
	S = int(std::min<uint32_t>(n, 8));
	L = int(n) / S;
	exact = true;
	return true;
}

void free_all(nlc_ctx *c) {
	if (c->stream) hipStreamSynchronize(c->stream);
	for (void *ptr : {(void *) c->state, (void *) c->omega, (void *) c->tw, (void *) c->partial, (void *) c->reduce_scratch, (void *) c->e_raw,
	                  (void *) c->results, (void *) c->modes, (void *) c->mp2, (void *) c->masses_d,
	                  (void *) c->eigvec, (void *) c->split_q, (void *) c->split_p, (void *) c->energy_d,
	                  (void *) c->norm_d})
		if (ptr) hipFree(ptr);
	if (c->h_results) hipHostFree(c->h_results);
	if (c->stream) hipStreamDestroy(c->stream);
}

template <class T>
int dev_alloc(nlc_ctx *c, T **ptr, size_t count, const char *what) {
	hipError_t e = hipMalloc((void **) ptr, sizeof(T) * std::max<size_t>(count, 1));
	if (e != hipSuccess) {
		*ptr = nullptr;
		return c->fail(e == hipErrorOutOfMemory ? NLC_ENOMEM : NLC_EHIP, "hipMalloc(%s, %zu bytes): %s", what,
		               sizeof(T) * count, hipGetErrorString(e));
	}
	return NLC_OK;
}

int setup(nlc_ctx *c) {
	const nlc_params &p = c->p;
	const uint32_t n = c->n;
	// constants (DNKG_FPUT_Toda.cpp:76-84, dDNKG.cpp:93-97)
	for (int i = 0; i < 8; ++i) {
		c->args.dc[i] = nlc::kYoshidaC[i] * p.dt;
		c->args.dd[i] = nlc::kYoshidaD[i] * p.dt;
	}
	c->args.m = p.m;
	c->args.m2 = p.m + 2;
	c->args.beta = p.beta;
	c->args.alpha = p.alpha;
	c->args.a2 = 2 * p.alpha;
	c->args.a2i = 1 / (2 * p.alpha);
	c->a0is0 = (p.model == NLC_DNLS) ? 1 : (p.model == NLC_DDNKG ? 0 : (p.m == 0.0));

	NLC_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
	int rc;
	if ((rc = dev_alloc(c, &c->state, size_t(n) * c->copies, "shard"))) return rc;
	if ((rc = dev_alloc(c, &c->omega, n, "omega"))) return rc;
	if ((rc = dev_alloc(c, &c->tw, n, "twiddles"))) return rc;
	if ((rc = dev_alloc(c, &c->e_raw, n, "linenergies"))) return rc;
	if ((rc = dev_alloc(c, &c->results, n + 2, "results"))) return rc;
	if ((rc = dev_alloc(c, &c->modes, n, "entropy modes"))) return rc;
	if ((rc = dev_alloc(c, &c->energy_d, c->copies, "chain energies"))) return rc;
	if ((rc = dev_alloc(c, &c->norm_d, c->copies, "chain norms"))) return rc;
	{
		hipError_t e = hipHostMalloc((void **) &c->h_results, sizeof(double) * (n + 2), hipHostMallocDefault);
		if (e != hipSuccess) return c->fail(NLC_ENOMEM, "hipHostMalloc(results): %s", hipGetErrorString(e));
	}

	// twiddles and dispersion, computed on the host in FP64
	std::vector<double2> tw(n);
	for (uint32_t j = 0; j < n; ++j) {
		const double ang = 2.0 * M_PI * double(j) / double(n);
		tw[j] = make_double2(std::cos(ang), -std::sin(ang));
	}
	NLC_HIP(c, hipMemcpy(c->tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
	std::vector<double> omega(n);

	if (p.model == NLC_DDNKG) {
		std::vector<double> mp2(n), V(size_t(n) * n);
		for (uint32_t i = 0; i < n; ++i) mp2[i] = c->masses[i] + 2; // dDNKG.cpp:47
		rc = nlc::eigensystem(n, mp2.data(), omega.data(), V.data());
		if (rc) return c->fail(rc, "eigensystem of the dDNKG linear operator did not converge");
		if ((rc = dev_alloc(c, &c->mp2, n, "mp2"))) return rc;
		if ((rc = dev_alloc(c, &c->masses_d, n, "masses"))) return rc;
		if ((rc = dev_alloc(c, &c->eigvec, size_t(n) * n, "eigenvectors"))) return rc;
		NLC_HIP(c, hipMemcpy(c->mp2, mp2.data(), sizeof(double) * n, hipMemcpyHostToDevice));
		NLC_HIP(c, hipMemcpy(c->masses_d, c->masses.data(), sizeof(double) * n, hipMemcpyHostToDevice));
		NLC_HIP(c, hipMemcpy(c->eigvec, V.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
		c->partial_rows = nlc::ddnkg_partial_rows(c->copies);
	} else {
		nlc_dispersion(p.model, n, p.m, omega.data());
		c->partial_rows = std::min<uint32_t>(c->copies, 1024);
	}
	NLC_HIP(c, hipMemcpy(c->omega, omega.data(), sizeof(double) * n, hipMemcpyHostToDevice));
	if ((rc = dev_alloc(c, &c->partial, size_t(c->partial_rows) * n, "energy partials"))) return rc;
	if ((rc = dev_alloc(c, &c->reduce_scratch, size_t(nlc::reduce_scratch_rows(c->partial_rows)) * n,
	                    "reduction scratch")))
		return rc;

	if (p.model == NLC_DNLS) {
		// DNLS/dispersion.cpp:19-26 (multipliers) and DNLS.cpp:96-100,161-175 (angles)
		if (!is_pow2(n) || n > nlc::kEnergyFftMax)
			return c->fail(NLC_EUNSUPPORTED, "DNLS needs a power-of-two chain_length <= %u in this build",
			               nlc::kEnergyFftMax);
		const double bdt = p.beta * p.dt;
		for (int k = 0; k < 8; ++k) c->theta[k] = (bdt * 1.) * nlc::kYoshidaC[k];
		c->theta[8] = (bdt * 2.) * nlc::kYoshidaC[0];
		std::vector<double2> tab(size_t(7) * n), perm(size_t(7) * n);
		const double nrm = 1.0 / n;
		for (int k = 0; k < 7; ++k)
			for (uint32_t j = 0; j < n; ++j) {
				const double ang = ((-p.dt) * nlc::kYoshidaD[k]) * omega[j];
				tab[size_t(k) * n + j] = make_double2(std::cos(ang) * nrm, std::sin(ang) * nrm);
			}
		const double2 *src = tab.data();
		if (nlc::dnls_fast_path(n)) {
			nlc::dnls_permute_tables(n, tab.data(), perm.data());
			src = perm.data();
		}
		if ((rc = dev_alloc(c, &c->dnls_table, size_t(7) * n, "DNLS multipliers"))) return rc;
		NLC_HIP(c, hipMemcpy(c->dnls_table, src, sizeof(double2) * 7 * n, hipMemcpyHostToDevice));
		return nlc_set_entropy_mask(c, nullptr, c->a0is0);
	}

	// step kernel selection (the reference's move<> dispatch, DNKG_FPUT_Toda.cu:262-297)
	bool want_split = p.kernel == NLC_KERNEL_SPLIT;
	if (!want_split) {
		int L, S;
		bool exact;
		bool have;
		if (p.lanes_per_chain || p.sites_per_lane) {
			L = int(p.lanes_per_chain);
			S = int(p.sites_per_lane);
			exact = uint64_t(L) * S == n;
			have = L >= 1 && L <= 64 && is_pow2(uint32_t(L)) && S >= 1 && uint64_t(L) * S >= n &&
			       uint64_t(L) * (S - 1) < n;
			if (!have) return c->fail(NLC_EARG, "wave shape %d lanes x %d sites cannot hold chain_length %u", L, S, n);
		} else {
			have = default_shape(n, L, S, exact);
		}
		c->launcher = have ? lookup(p.model, L, S, exact) : nullptr;
		if (c->launcher) {
			c->L = L, c->S = S, c->exact = exact;
		} else if (p.kernel == NLC_KERNEL_WAVE || p.lanes_per_chain) {
			return c->fail(NLC_EUNSUPPORTED, "no wave-resident kernel for chain_length %u (shape %d x %d)", n, L, S);
		} else {
			want_split = true;
		}
	}
	if (want_split) {
		c->use_split = true;
		if ((rc = dev_alloc(c, &c->split_q, size_t(n) * c->copies, "split q"))) return rc;
		if ((rc = dev_alloc(c, &c->split_p, size_t(n) * c->copies, "split p"))) return rc;
	}
	return nlc_set_entropy_mask(c, nullptr, c->a0is0);
}

} // namespace

extern "C" {

int nlc_abi_version(void) { return NLC_ABI_VERSION; }

const char *nlc_strerror(int code) {
	switch (code) {
	case NLC_OK: return "success";
	case NLC_EARG: return "invalid argument";
	case NLC_EHIP: return "HIP runtime error";
	case NLC_ENOMEM: return "out of memory";
	case NLC_EUNSUPPORTED: return "unsupported";
	case NLC_ESTATE: return "invalid state";
	}
	return "unknown error";
}

const char *nlc_last_error(const nlc_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int nlc_params_init(nlc_params *p) {
	if (!p) return NLC_EARG;
	std::memset(p, 0, sizeof *p);
	p->struct_size = sizeof *p;
	p->model = NLC_DNKG;
	p->precision = NLC_FP64;
	p->kernel = NLC_KERNEL_AUTO;
	return NLC_OK;
}

int nlc_create(nlc_ctx **out, const nlc_params *p) {
	g_create_error.clear();
	auto reject = [](int code, const char *msg) {
		g_create_error = msg;
		return code;
	};
	if (!out) return reject(NLC_EARG, "out is NULL");
	*out = nullptr;
	if (!p || p->struct_size != sizeof(nlc_params)) return reject(NLC_EARG, "params NULL or struct_size mismatch");
	// argument rules of common/main.cpp:86-87,101-103 and the model options
	if (p->model < NLC_DNKG || p->model > NLC_DNLS) return reject(NLC_EARG, "unknown model");
	if (p->chain_length < 2) return reject(NLC_EARG, "--chain_length must be >= 2");
	if (!p->copies) return reject(NLC_EARG, "--copies must be a positive number");
	if (!(p->dt > 0)) return reject(NLC_EARG, "--dt must be a positive number");
	if (uint64_t(p->chain_length) * p->copies >= 0x40000000ULL)
		return reject(NLC_EARG, "(--chain_length * --copies) must be < 2^30");
	if (p->model == NLC_TODA && p->alpha == 0) return reject(NLC_EARG, "Toda needs --alpha != 0");
	if (p->model == NLC_DDNKG && !p->masses) return reject(NLC_EARG, "dDNKG needs the masses (-m file content)");
	if (p->precision != NLC_FP64) return reject(NLC_EUNSUPPORTED, "only FP64 precision is available in this build");
	if (p->kernel < NLC_KERNEL_AUTO || p->kernel > NLC_KERNEL_SPLIT) return reject(NLC_EARG, "unknown kernel family");

	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
		return reject(NLC_EHIP, "no HIP device available (the hot path has no CPU fallback)");

	nlc_ctx *c = new (std::nothrow) nlc_ctx;
	if (!c) return reject(NLC_ENOMEM, "context allocation failed");
	c->p = *p;
	c->n = p->chain_length;
	c->copies = p->copies;
	c->state_bytes = sizeof(double2) * size_t(c->n) * c->copies;
	if (p->masses) c->masses.assign(p->masses, p->masses + c->n);
	c->p.masses = nullptr;
	int rc = setup(c);
	if (rc) {
		g_create_error = c->err;
		free_all(c);
		delete c;
		return rc;
	}
	*out = c;
	return NLC_OK;
}

void nlc_destroy(nlc_ctx *ctx) {
	if (!ctx) return;
	free_all(ctx);
	delete ctx;
}

int nlc_upload(nlc_ctx *c, const void *state) {
	if (!c || !state) return NLC_EARG;
	NLC_HIP(c, hipMemcpyAsync(c->state, state, c->state_bytes, hipMemcpyHostToDevice, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	return NLC_OK;
}

int nlc_download(nlc_ctx *c, void *state) {
	if (!c || !state) return NLC_EARG;
	NLC_HIP(c, hipMemcpyAsync(state, c->state, c->state_bytes, hipMemcpyDeviceToHost, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	return NLC_OK;
}

int nlc_device_state(nlc_ctx *c, void **dptr, size_t *bytes) {
	if (!c || !dptr || !bytes) return NLC_EARG;
	*dptr = c->state;
	*bytes = c->state_bytes;
	return NLC_OK;
}

int nlc_move(nlc_ctx *c, uint32_t steps) {
	if (!c) return NLC_EARG;
	if (!steps) return c->fail(NLC_EARG, "--kernel_batching must be a positive number");
	if (c->p.model == NLC_DNLS) {
		NLC_HIP(c, nlc::launch_dnls(c->stream, c->state, c->n, c->copies, steps, c->theta, c->tw, c->dnls_table));
		return NLC_OK;
	}
	if (c->use_split) {
		NLC_HIP(c, nlc::launch_planar_to_split(c->stream, c->state, c->split_q, c->split_p, c->copies, c->n));
		NLC_HIP(c, nlc::launch_step_split(c->stream, c->p.model, c->split_q, c->split_p, c->copies, c->n, steps,
		                                  c->args, c->mp2));
		NLC_HIP(c, nlc::launch_split_to_planar(c->stream, c->split_q, c->split_p, c->state, c->copies, c->n));
		return NLC_OK;
	}
	NLC_HIP(c, c->launcher(c->stream, c->state, c->copies, c->n, steps, c->args, c->mp2));
	return NLC_OK;
}

int nlc_set_entropy_mask(nlc_ctx *c, const uint8_t *mask, int a0is0) {
	if (!c) return NLC_EARG;
	std::vector<uint32_t> idx;
	if (mask) {
		// common/main.cpp:119-128 then common/results.cpp:8-15
		for (uint32_t k = 0; k < c->n; ++k)
			if (mask[k]) idx.push_back(k);
		if (idx.size() < 2) return c->fail(NLC_EARG, "Too few modes for entropy calculation");
		if (a0is0 && idx[0] == 0) {
			idx.erase(idx.begin());
			if (idx.size() < 2) return c->fail(NLC_EARG, "Too few modes for entropy calculation");
		}
		NLC_HIP(c, hipMemcpyAsync(c->modes, idx.data(), sizeof(uint32_t) * idx.size(), hipMemcpyHostToDevice,
		                          c->stream));
		NLC_HIP(c, hipStreamSynchronize(c->stream));
	}
	c->nmodes = uint32_t(idx.size());
	c->a0is0 = a0is0 ? 1 : 0;
	return NLC_OK;
}

static int energy_pass(nlc_ctx *c) {
	const uint32_t n = c->n;
	switch (c->p.model) {
	case NLC_DNKG:
	case NLC_FPUT:
	case NLC_TODA:
		NLC_HIP(c, nlc::launch_energy_real(c->stream, c->state, n, c->copies, c->tw, c->omega, c->partial,
		                                   c->partial_rows));
		NLC_HIP(c, nlc::launch_reduce_partial(c->stream, c->partial, c->partial_rows, n, nullptr, c->e_raw,
		                                      c->reduce_scratch));
		break;
	case NLC_DDNKG:
		NLC_HIP(c, nlc::launch_energy_ddnkg(c->stream, c->state, n, c->copies, c->eigvec, c->omega, c->partial));
		NLC_HIP(c, nlc::launch_reduce_partial(c->stream, c->partial, c->partial_rows, n, nullptr, c->e_raw,
		                                      c->reduce_scratch));
		break;
	case NLC_DNLS:
		NLC_HIP(c, nlc::launch_energy_dnls(c->stream, c->state, n, c->copies, c->tw, c->partial, c->partial_rows));
		NLC_HIP(c, nlc::launch_reduce_partial(c->stream, c->partial, c->partial_rows, n, c->omega, c->e_raw,
		                                      c->reduce_scratch));
		break;
	default: return c->fail(NLC_EUNSUPPORTED, "no energy pass for this model");
	}
	return NLC_OK;
}

int nlc_linenergies_raw(nlc_ctx *c, double *e_raw) {
	if (!c || !e_raw) return NLC_EARG;
	int rc = energy_pass(c);
	if (rc) return rc;
	NLC_HIP(c, hipMemcpyAsync(c->h_results, c->e_raw, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	std::memcpy(e_raw, c->h_results, sizeof(double) * c->n);
	return NLC_OK;
}

int nlc_linenergies(nlc_ctx *c, double norm, double *linenergies, double *entropies) {
	if (!c) return NLC_EARG;
	const uint32_t n = c->n;
	if (!(norm > 0)) {
		// results::calc_linenergies normalisations of the drivers
		switch (c->p.model) {
		case NLC_DDNKG: norm = 0.5 / c->copies; break;                      // dDNKG.cpp:141
		case NLC_DNLS: norm = 1.0 / (double(n) * c->copies); break;         // DNLS.cpp:150
		default: norm = (0.5 / c->copies) / n; break;                       // DNKG_FPUT_Toda.cpp:129
		}
	}
	int rc = energy_pass(c);
	if (rc) return rc;
	NLC_HIP(c, nlc::launch_results(c->stream, c->e_raw, n, norm, c->a0is0, c->modes, c->nmodes, c->results));
	NLC_HIP(c, hipMemcpyAsync(c->h_results, c->results, sizeof(double) * (n + 2), hipMemcpyDeviceToHost, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	if (linenergies) std::memcpy(linenergies, c->h_results, sizeof(double) * n);
	if (entropies) entropies[0] = c->h_results[n], entropies[1] = c->h_results[n + 1];
	return NLC_OK;
}

int nlc_chain_energies(nlc_ctx *c, double *energies, double *norms) {
	if (!c || !energies) return NLC_EARG;
	NLC_HIP(c, nlc::launch_chain_energy(c->stream, c->p.model, c->state, c->n, c->copies, c->args, c->masses_d,
	                                    c->energy_d, c->norm_d));
	NLC_HIP(c, hipMemcpyAsync(energies, c->energy_d, sizeof(double) * c->copies, hipMemcpyDeviceToHost, c->stream));
	if (norms && c->p.model == NLC_DNLS)
		NLC_HIP(c, hipMemcpyAsync(norms, c->norm_d, sizeof(double) * c->copies, hipMemcpyDeviceToHost, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	return NLC_OK;
}

int nlc_sync(nlc_ctx *c) {
	if (!c) return NLC_EARG;
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	return NLC_OK;
}

void *nlc_stream(nlc_ctx *c) { return c ? (void *) c->stream : nullptr; }

int nlc_describe(const nlc_ctx *c, char *buf, size_t len) {
	if (!c || !buf || !len) return NLC_EARG;
	static const char *names[] = {"DNKG", "FPUT", "Toda", "dDNKG", "DNLS"};
	if (c->p.model == NLC_DNLS)
		snprintf(buf, len, "model=DNLS n=%u copies=%u kernel=%s", c->n, c->copies,
		         nlc::dnls_fast_path(c->n) ? "dnls_radix16_registers" : "dnls_lds_stockham");
	else if (c->use_split)
		snprintf(buf, len, "model=%s n=%u copies=%u kernel=split", names[c->p.model], c->n, c->copies);
	else
		snprintf(buf, len, "model=%s n=%u copies=%u kernel=wave lanes_per_chain=%d sites_per_lane=%d exact=%d",
		         names[c->p.model], c->n, c->copies, c->L, c->S, int(c->exact));
	return NLC_OK;
}

} // extern "C"
