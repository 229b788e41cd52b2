// C-ABI layer of libnlchains_amd: context life cycle, kernel selection,
// ragged ensembles, hipGraph batches, the autotuner and the entry points
// declared in include/nlc.h.  No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nlc.h"
#include "nlc_internal.hpp"

namespace {

thread_local std::string g_create_error;

// A length class: `count` chains of n sites, contiguous in device memory.
struct Klass {
	uint32_t n = 0, count = 0;
	uint64_t site0 = 0;  // first site of the class in the device shard
	uint64_t chain0 = 0; // first chain (internal order)
	uint64_t tab0 = 0;   // offset of the class in the per-class tables (sum of smaller n)
	uint32_t rows = 0;   // energy partial rows
	// per-class step kernel when the ragged persistent kernel does not apply
	nlc::StepLauncher launcher = nullptr;
	int L = 0, S = 0;
	bool exact = false;
	int side = 0; // per-class launches: 0 = context stream, 1 = the second stream
	int set = 0;   // parameter set (mixed ensembles: one per model; otherwise 0)
	int model = 0; // that set's model
	int a0is0 = 0; // the results(a0is0) flag of the class
};

} // namespace

void nlc::set_thread_error(const char *msg) { g_create_error = msg; }

struct nlc_ctx {
	nlc_params p{};                // the context's parameters (mixed ensembles: set 0's)
	std::vector<nlc_params> sets;  // parameter sets of the classes (one unless mixed)
	std::vector<double> masses; // dDNKG m_x
	hipStream_t stream = nullptr;
	nlc::StepArgs args{};
	uint32_t n = 0, copies = 0; // uniform: chain length and copies; ragged: max length and chain count
	uint64_t total_sites = 0;
	size_t state_bytes = 0;
	std::vector<Klass> cls;
	uint64_t table_len = 0; // sum of class lengths

	// ragged ensembles (chains in the caller's order vs class-major on the device)
	bool ragged = false;
	std::vector<uint32_t> lengths;       // caller order
	std::vector<uint64_t> caller_site0;  // caller order
	std::vector<uint64_t> device_site0;  // caller order -> device offset
	std::vector<uint32_t> device_chain;  // caller order -> internal chain index
	// persistent kernel: items sorted most expensive first, S sites per lane
	// (S/2 for the classes that fit in 64 lanes of S/2 when half_S), the
	// parameter sets of its kinds, and its two queue counters
	nlc::RaggedItem *items = nullptr;
	uint32_t nitems = 0;
	int ragged_S = 0;
	bool ragged_half = false;
	nlc::PersistentLauncher persistent = nullptr;
	nlc::PersistentArgs pargs{};
	uint32_t *queue = nullptr;
	// per-class launches run on two streams (fork and join by events), so that
	// the waves of one class fill the tail of another class's launch
	hipStream_t aux_stream = nullptr;
	hipEvent_t fork_ev = nullptr, join_ev = nullptr;

	double2 *state = nullptr;
	double *omega = nullptr;   // per-class dispersion, table_len
	double2 *tw = nullptr;     // per-class roots of unity exp(-2 pi i j/n), table_len
	double *partial = nullptr; // energy partial sums
	double *reduce_scratch = nullptr;
	double *e_raw = nullptr;   // table_len
	double2 *fft_scratch = nullptr; // two-level FFT of chains above 8192 sites, Bluestein above 4096
	size_t fft_scratch_sites = 0;
	double2 *blue_table = nullptr;  // Bluestein tables of a uniform length that is not a power of two
	uint32_t *modes = nullptr; // entropy mode indices
	uint32_t nmodes = 0;
	int a0is0 = 0;
	double *mp2 = nullptr;     // dDNKG m_x + 2
	double *masses_d = nullptr;
	double *eigvec = nullptr;  // dDNKG V, column-major
	int proj_variant = nlc::kProjMfmaWide; // dDNKG projection: FP64 matrix cores, vector ALU or split BF16
	uint16_t *vplanes = nullptr; // split BF16 form: the eigenvector splits
	double *split_q[2] = {nullptr, nullptr}, *split_p[2] = {nullptr, nullptr}; // ping-pong pair
	double *energy_d = nullptr, *norm_d = nullptr;
	double2 *dnls_table = nullptr; // DNLS Fourier multipliers
	float2 *dnls_tw32 = nullptr;   // FP32-compensated DNLS: FP32 twiddles
	float4 *dnls_mult32 = nullptr; // and double-single increment multipliers
	int dnls_variant = nlc::kDnlsRadix16; // n = 4096: register kernel, radix 16 or radix 8
	double theta[9] = {};          // DNLS nonlinear angles
	// results (per class: n energies + 2 entropies) are written by the results
	// kernel straight into pinned host memory: h_results for the synchronous
	// calls, one of the slots h_slots for queued batches (nlc_batch_submit)
	double *h_results = nullptr;
	double *h_slots = nullptr;
	hipEvent_t slot_ev[NLC_BATCH_SLOTS] = {};
	uint32_t slot_head = 0, slot_count = 0;
	// host transfers of the shard: pinned double buffer and the ragged chains in device order
	double2 *pinned[2] = {nullptr, nullptr};
	hipEvent_t pinned_ev[2] = {nullptr, nullptr};
	std::vector<uint32_t> device_order;

	nlc::StepLauncher launcher = nullptr; // uniform wave kernel
	int L = 0, S = 0;
	bool exact = false, use_split = false;
	nlc::TiledLauncher tiled = nullptr; // temporally blocked kernel for long chains
	double2 *tile_buf = nullptr;        // its second state buffer
	bool use_tiled = false;

	// cached graphs of one batch, one per result destination: [0, NLC_BATCH_SLOTS)
	// for the slots of queued batches, NLC_BATCH_SLOTS for nlc_batch
	struct BatchGraph {
		hipGraph_t graph = nullptr;
		hipGraphExec_t exec = nullptr;
		uint32_t steps = 0;
		double norm = 0;
	} graphs[NLC_BATCH_SLOTS + 1];
	// queued batches as direct kernel launches instead of graph replays (the
	// autotuner times both: a replay costs a few microseconds more than the
	// launches it holds, which matters only for batches of tens of microseconds)
	bool batch_direct = false;

	std::string err, tune_report;

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
	size_t results_len() const { return size_t(table_len) + 2 * cls.size(); }
};

#define NLC_HIP(ctx, call)                                                                                            \
	do {                                                                                                             \
		hipError_t e_ = (call);                                                                                        \
		if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #call);                                                       \
	} while (0)
#define NLC_TRY(call)                                                                                                 \
	do {                                                                                                             \
		int rc_ = (call);                                                                                              \
		if (rc_) return rc_;                                                                                           \
	} while (0)

namespace {

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
int log2u(uint32_t x) {
	int l = 0;
	while ((1u << l) < x) ++l;
	return l;
}

// A workgroup shape (L = 64 W lanes, W = 2..8) that the group launcher accepts for n.
bool group_shape_ok(uint32_t n, int L, int S) {
	if (L < 128 || L > 512 || L % 64 || S < 8 || S > 32) return false;
	const uint64_t G = 64u * ((uint64_t(n) + 64u * S - 1) / (64u * S));
	return G == uint64_t(L) && G * (S - 1) < n && G * S >= n;
}

nlc::StepLauncher lookup_group(int model, int S, bool exact) {
	nlc::StepLauncher (*f)(int, bool) = nullptr;
	switch (model) {
	case NLC_DNKG: f = nlc::lookup_group_dnkg; break;
	case NLC_FPUT: f = nlc::lookup_group_fput; break;
	case NLC_TODA: f = nlc::lookup_group_toda; break;
	case NLC_DDNKG: f = nlc::lookup_group_ddnkg; break;
	default: return nullptr;
	}
	// an exact fit without its own instance runs on the general one
	nlc::StepLauncher l = f(S, exact);
	return l ? l : (exact ? f(S, false) : nullptr);
}

nlc::StepLauncher lookup(int model, int L, int S, bool exact) {
	if (L > 64) return lookup_group(model, S, exact);
	nlc::StepLauncher (*f)(int, int, bool) = nullptr;
	nlc::StepLauncher (*f16)(int) = nullptr;
	switch (model) {
	case NLC_DNKG: f = nlc::lookup_step_dnkg, f16 = nlc::lookup_step16_dnkg; break;
	case NLC_FPUT: f = nlc::lookup_step_fput, f16 = nlc::lookup_step16_fput; break;
	case NLC_TODA: f = nlc::lookup_step_toda, f16 = nlc::lookup_step16_toda; break;
	case NLC_DDNKG: f = nlc::lookup_step_ddnkg, f16 = nlc::lookup_step16_ddnkg; break;
	default: return nullptr;
	}
	// an exact fit without its own instance (n = 64 S, 16 S or 4 S with S not a
	// power of two) runs on the ragged-lane one, whose lanes are then all full
	nlc::StepLauncher l = f(L, S, exact);
	if (!l && exact && S >= 2) l = f(L, S, false);
	if (!l && L == 16) l = f16(S);
	return l;
}

// FP32-compensated wave kernels: FPUT and Toda, exact power-of-two shapes.
nlc::StepLauncher lookup32(int model, int L, int S) {
	if (model == NLC_FPUT) return nlc::lookup_step32_fput(L, S);
	if (model == NLC_TODA) return nlc::lookup_step32_toda(L, S);
	return nullptr;
}

// The wave kernel of a shape in the given precision (real models).
nlc::StepLauncher lookup_prec(int model, int L, int S, bool exact, bool c32) {
	if (!c32) return lookup(model, L, S, exact);
	return exact && L <= 64 ? lookup32(model, L, S) : nullptr;
}

nlc::TiledLauncher lookup_tiled(int model) {
	switch (model) {
	case NLC_DNKG: return nlc::lookup_tiled_dnkg();
	case NLC_FPUT: return nlc::lookup_tiled_fput();
	case NLC_TODA: return nlc::lookup_tiled_toda();
	case NLC_DDNKG: return nlc::lookup_tiled_ddnkg();
	}
	return nullptr;
}

nlc::PersistentLauncher lookup_persistent(int S, uint32_t kind_mask, bool c32) {
	switch (S) {
	case 8: return c32 ? nlc::lookup_persistent32_s8(kind_mask) : nlc::lookup_persistent_s8(kind_mask);
	case 16: return c32 ? nlc::lookup_persistent32_s16(kind_mask) : nlc::lookup_persistent_s16(kind_mask);
	case 32: return c32 ? nlc::lookup_persistent32_s32(kind_mask) : nlc::lookup_persistent_s32(kind_mask);
	}
	return nullptr;
}

// Default wave shape of a chain of n <= 2048 sites in an ensemble of `copies`
// chains: 64 lanes from 64 sites up, except the powers of two 64..256, which
// take S = 16 sites per lane (on 2^24-site ensembles 4 x 16 runs 38 % faster
// than 64 x 1 at 64 sites, 16 x 16 6 % faster than 64 x 4 at 256), and the
// other lengths 65..384, which take 16 lanes; powers of two below 64 take
// S = min(n, 8).  A small ensemble trades sites per lane for lanes (L doubles,
// S halves; 16 lanes become 64) until it has 4096 wavefronts or L = 64: then
// the number of wavefronts, not their throughput, limits the step.
bool default_shape(uint32_t n, uint32_t copies, int &L, int &S, bool &exact) {
	if (is_pow2(n) && n <= 256) {
		S = int(std::min<uint32_t>(n, n >= 64 ? 16 : 8));
		L = int(n) / S;
		while (L < 64 && S > 1 && uint64_t(copies) * L < 4096ull * 64) L *= 2, S /= 2;
		exact = true;
		return true;
	}
	if (n > 64 && n <= 384 && uint64_t(copies) * 16 >= 4096ull * 64) {
		// other lengths up to 384: 16 lanes (a DPP row) with ragged lanes, four
		// chains to a wavefront (1.5x the rate of 64 x 2 at 100 sites, 1.13x of
		// 64 x 5 at 300; from 24 sites per lane on, 64 lanes are as fast or faster)
		S = int((n + 15) / 16);
		L = 16;
		exact = (uint32_t(S) * 16 == n);
		return true;
	}
	if (n >= 64) {
		S = int((n + 63) / 64);
		L = 64;
		exact = (uint32_t(S) * 64 == n);
		return S <= 32;
	}
	// short chains of other lengths: a thread per chain up to 32 sites, else 4
	// lanes with 9..16 sites (the last lanes one fewer)
	if (n <= 32) {
		L = 1, S = int(n), exact = true;
	} else {
		L = 4, S = int((n + 3) / 4), exact = false;
	}
	return true;
}

// Chains longer than a wavefront holds: the workgroup shape (measured on
// 2^24-site ensembles at 2500 to 16384 sites per chain).
bool group_default(int model, uint32_t n, int &L, int &S, bool &exact) {
	(void) model;
	// 2, 4 or 8 wavefronts spread evenly over the 4 SIMDs of a CU; with 3, 5, 6
	// or 7 every barrier waits for a SIMD that runs two of the workgroup's
	// wavefronts (measured 10-45 % slower).  The fewest such wavefronts with at
	// most 24 sites per lane, else 32; other counts only as a last resort.
	for (int smax : {24, 32})
		for (int W : {2, 4, 8}) {
			const int s = int((n + 64u * W - 1) / (64u * W));
			if (s > smax || !group_shape_ok(n, 64 * W, s)) continue;
			L = 64 * W, S = s, exact = uint64_t(L) * S == n;
			return true;
		}
	for (int W = 3; W <= 7; ++W) {
		const int s = int((n + 64u * W - 1) / (64u * W));
		if (!group_shape_ok(n, 64 * W, s)) continue;
		L = 64 * W, S = s, exact = uint64_t(L) * S == n;
		return true;
	}
	return false;
}

// Every instantiated wave and workgroup shape that holds a chain of n sites.
struct Shape {
	int L, S;
	bool exact;
};
std::vector<Shape> wave_shapes(int model, uint32_t n, bool c32) {
	std::vector<Shape> out;
	for (int L = 1; L <= 64; L *= 2) {
		const int S = int((n + L - 1) / L);
		if (S < 1 || S > 32) continue;
		const bool exact = uint32_t(L * S) == n;
		if (!exact && S < 2) continue;
		if (lookup_prec(model, L, S, exact, c32)) out.push_back({L, S, exact});
	}
	if (c32) return out;
	for (int L = 128; L <= 512; L += 64) {
		const int S = int((n + L - 1) / L);
		if (!group_shape_ok(n, L, S)) continue;
		const bool exact = uint64_t(L) * S == n;
		if (lookup(model, L, S, exact)) out.push_back({L, S, exact});
	}
	return out;
}

// Sites per lane of a class in the persistent kernel with S sites per lane:
// S, or with half_S the S/2 of a class that fits in 64 lanes of S/2; 0 if the
// class has no such shape (the kernel needs powers of two, n = L S').
int item_S(uint32_t n, int S, bool half_S) {
	if (!is_pow2(n) || n > 64u * S) return 0;
	if (half_S && S >= 2 && n <= 32u * S) return n >= uint32_t(S / 2) ? S / 2 : 0;
	return n >= uint32_t(S) ? S : 0;
}
bool ragged_fits(const std::vector<Klass> &cls, int S, bool half_S) {
	for (const auto &k : cls)
		if (!item_S(k.n, S, half_S)) return false;
	return true;
}
// VALU instructions per site and substep of each kind (its kick and drift),
// the cost weight of the persistent kernel's ordering; FP32-compensated FPUT
// and Toda kinds in the second row
constexpr double kKindCost[2][6] = {{5, 6, 6, 6, 7, 19}, {5, 6, 14, 14, 15, 30}};

void drop_graph(nlc_ctx *c) {
	// a queued replay may still be running
	if (c->slot_count) (void) hipStreamSynchronize(c->stream);
	for (auto &g : c->graphs) {
		if (g.exec) (void) hipGraphExecDestroy(g.exec);
		if (g.graph) (void) hipGraphDestroy(g.graph);
		g = nlc_ctx::BatchGraph();
	}
}

void free_all(nlc_ctx *c) {
	if (c->stream) (void) hipStreamSynchronize(c->stream);
	drop_graph(c);
	for (void *ptr :
	     {(void *) c->state, (void *) c->omega, (void *) c->tw, (void *) c->partial, (void *) c->reduce_scratch,
	      (void *) c->e_raw, (void *) c->modes, (void *) c->mp2, (void *) c->masses_d,
	      (void *) c->eigvec, (void *) c->split_q[0], (void *) c->split_p[0], (void *) c->split_q[1], (void *) c->split_p[1], (void *) c->energy_d, (void *) c->norm_d,
	      (void *) c->dnls_table, (void *) c->fft_scratch, (void *) c->blue_table, (void *) c->items, (void *) c->queue, (void *) c->vplanes,
	      (void *) c->tile_buf, (void *) c->dnls_tw32, (void *) c->dnls_mult32})
		if (ptr) (void) hipFree(ptr);
	if (c->h_results) (void) hipHostFree(c->h_results);
	if (c->h_slots) (void) hipHostFree(c->h_slots);
	for (hipEvent_t ev : c->slot_ev)
		if (ev) (void) hipEventDestroy(ev);
	for (int i = 0; i < 2; ++i) {
		if (c->pinned[i]) (void) hipHostFree(c->pinned[i]);
		if (c->pinned_ev[i]) (void) hipEventDestroy(c->pinned_ev[i]);
	}
	if (c->fork_ev) (void) hipEventDestroy(c->fork_ev);
	if (c->join_ev) (void) hipEventDestroy(c->join_ev);
	if (c->aux_stream) (void) hipStreamDestroy(c->aux_stream);
	if (c->stream) (void) hipStreamDestroy(c->stream);
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

// Step constants of one parameter set (DNKG_FPUT_Toda.cpp:76-84, dDNKG.cpp:93-97).
nlc::StepArgs step_args(const nlc_params &p) {
	nlc::StepArgs a{};
	for (int i = 0; i < 8; ++i) {
		a.dc[i] = nlc::kYoshidaC[i] * p.dt;
		a.dd[i] = nlc::kYoshidaD[i] * p.dt;
	}
	a.m = p.m;
	a.m2 = p.m + 2;
	a.beta = p.beta;
	a.alpha = p.alpha;
	a.a2 = 2 * p.alpha;
	a.a2i = 1 / (2 * p.alpha);
	a.form = nlc::force_form(p.model, p.alpha, p.beta, &a.scale);
	a.inv_scale = 1 / a.scale;
	return a;
}

// The reference's results(a0is0) default of a model (DNKG_FPUT_Toda.cpp:52,
// DNLS.cpp:49, dDNKG.cpp:83).
int default_a0is0(const nlc_params &p) { return p.model == NLC_DNLS ? 1 : p.model == NLC_DDNKG ? 0 : p.m == 0.0; }

// Buffers and tables shared by uniform and ragged contexts.
int setup_common(nlc_ctx *c) {
	const nlc_params &p = c->p;
	if (c->sets.empty()) c->sets.push_back(p);
	for (int s = 0; s < 3; ++s) {
		if (s < int(c->sets.size())) {
			c->pargs.a[s] = step_args(c->sets[s]);
			c->pargs.kind[s] = nlc::kind_of(c->sets[s].model, c->pargs.a[s].form);
		} else {
			c->pargs.a[s] = c->pargs.a[0];
			c->pargs.kind[s] = -1;
		}
	}
	c->args = c->pargs.a[0];
	c->a0is0 = default_a0is0(p);
	for (auto &k : c->cls) {
		k.model = c->sets[k.set].model;
		k.a0is0 = default_a0is0(c->sets[k.set]);
	}

	NLC_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
	c->state_bytes = sizeof(double2) * c->total_sites;
	NLC_TRY(dev_alloc(c, &c->state, c->total_sites, "shard"));
	NLC_TRY(dev_alloc(c, &c->omega, c->table_len, "omega"));
	NLC_TRY(dev_alloc(c, &c->tw, c->table_len, "twiddles"));
	NLC_TRY(dev_alloc(c, &c->e_raw, c->table_len, "linenergies"));
	NLC_TRY(dev_alloc(c, &c->modes, c->n, "entropy modes"));
	NLC_TRY(dev_alloc(c, &c->energy_d, c->copies, "chain energies"));
	NLC_TRY(dev_alloc(c, &c->norm_d, c->copies, "chain norms"));
	{
		hipError_t e = hipHostMalloc((void **) &c->h_results, sizeof(double) * c->results_len(), hipHostMallocCoherent);
		if (e != hipSuccess) return c->fail(NLC_ENOMEM, "hipHostMalloc(results): %s", hipGetErrorString(e));
	}
	// per-class twiddles and dispersion, computed on the host in FP64
	std::vector<double2> tw(c->table_len);
	std::vector<double> omega(c->table_len);
	size_t partial_len = 1, scratch_len = 1;
	for (auto &k : c->cls) {
		for (uint32_t j = 0; j < k.n; ++j) {
			const double ang = 2.0 * M_PI * double(j) / double(k.n);
			tw[k.tab0 + j] = make_double2(std::cos(ang), -std::sin(ang));
		}
		if (p.model != NLC_DDNKG) nlc_dispersion(k.model, k.n, c->sets[k.set].m, omega.data() + k.tab0);
		k.rows = p.model == NLC_DDNKG ? nlc::ddnkg_partial_rows(k.count) : nlc::energy_rows(k.n, k.count);
		if (p.model != NLC_DDNKG)
			c->fft_scratch_sites = std::max(c->fft_scratch_sites, nlc::energy_scratch_sites(k.n, k.count));
		partial_len = std::max(partial_len, size_t(k.rows) * k.n);
		scratch_len = std::max(scratch_len, size_t(nlc::reduce_scratch_rows(k.rows)) * k.n);
	}
	NLC_HIP(c, hipMemcpy(c->tw, tw.data(), sizeof(double2) * c->table_len, hipMemcpyHostToDevice));
	if (p.model == NLC_DDNKG) {
		const uint32_t n = c->n;
		std::vector<double> mp2(n), V(size_t(n) * n);
		for (uint32_t i = 0; i < n; ++i) mp2[i] = c->masses[i] + 2; // dDNKG.cpp:47
		const int rc = nlc::eigensystem(n, mp2.data(), omega.data(), V.data());
		if (rc) return c->fail(rc, "eigensystem of the dDNKG linear operator did not converge");
		NLC_TRY(dev_alloc(c, &c->mp2, n, "mp2"));
		NLC_TRY(dev_alloc(c, &c->masses_d, n, "masses"));
		NLC_TRY(dev_alloc(c, &c->eigvec, size_t(n) * n, "eigenvectors"));
		NLC_HIP(c, hipMemcpy(c->mp2, mp2.data(), sizeof(double) * n, hipMemcpyHostToDevice));
		NLC_HIP(c, hipMemcpy(c->masses_d, c->masses.data(), sizeof(double) * n, hipMemcpyHostToDevice));
		NLC_HIP(c, hipMemcpy(c->eigvec, V.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
		if (c->proj_variant == nlc::kProjBf16) {
			if (!nlc::ddnkg_bf16_supported(n))
				return c->fail(NLC_EUNSUPPORTED, "the split BF16 projection needs chain_length divisible by 32");
			std::vector<uint16_t> vp(size_t(3) * n * n);
			nlc::ddnkg_split_host(V.data(), size_t(n) * n, vp.data());
			NLC_TRY(dev_alloc(c, &c->vplanes, vp.size(), "eigenvector splits"));
			NLC_HIP(c, hipMemcpy(c->vplanes, vp.data(), sizeof(uint16_t) * vp.size(), hipMemcpyHostToDevice));
		}
	}
	NLC_HIP(c, hipMemcpy(c->omega, omega.data(), sizeof(double) * c->table_len, hipMemcpyHostToDevice));
	NLC_TRY(dev_alloc(c, &c->partial, partial_len, "energy partials"));
	if (c->fft_scratch_sites) NLC_TRY(dev_alloc(c, &c->fft_scratch, c->fft_scratch_sites, "FFT scratch"));
	if (const uint32_t L = nlc::bluestein_two_level_size(c->n);
	    L && !c->ragged && p.model != NLC_DDNKG && !nlc::mixed_radix_plan(c->n, nullptr)) {
		std::vector<double2> blue(size_t(2) * L + c->n);
		nlc::bluestein_two_level_tables(c->n, blue.data());
		NLC_TRY(dev_alloc(c, &c->blue_table, blue.size(), "Bluestein tables"));
		NLC_HIP(c, hipMemcpy(c->blue_table, blue.data(), sizeof(double2) * blue.size(), hipMemcpyHostToDevice));
	}
	NLC_TRY(dev_alloc(c, &c->reduce_scratch, scratch_len, "reduction scratch"));
	return NLC_OK;
}

// DNLS tables: DNLS/dispersion.cpp:19-26 (multipliers), DNLS.cpp:96-100,161-175 (angles)
int setup_dnls(nlc_ctx *c) {
	const uint32_t n = c->n;
	const nlc_params &p = c->p;
	if (c->copies >= (1u << 24))
		return c->fail(NLC_EUNSUPPORTED, "DNLS runs one workgroup per chain: at most 2^24 - 1 copies");
	if (!nlc::dnls_supported(n))
		return c->fail(NLC_EUNSUPPORTED, "DNLS needs a chain_length of 2 to %u, or a power of two up to %u",
		               nlc::kLargeFftMax / 2, nlc::kLargeFftMax);
	const double bdt = p.beta * p.dt;
	for (int k = 0; k < 8; ++k) c->theta[k] = (bdt * 1.) * nlc::kYoshidaC[k];
	c->theta[8] = (bdt * 2.) * nlc::kYoshidaC[0];
	std::vector<double> omega(n);
	nlc_dispersion(NLC_DNLS, n, 0, omega.data());
	std::vector<double2> tab(size_t(7) * n), perm(size_t(7) * n);
	const double nrm = 1.0 / n;
	for (int k = 0; k < 7; ++k)
		for (uint32_t j = 0; j < n; ++j) {
			const double ang = ((-p.dt) * nlc::kYoshidaD[k]) * omega[j];
			tab[size_t(k) * n + j] = make_double2(std::cos(ang) * nrm, std::sin(ang) * nrm);
		}
	// generic kernel: the 7 tables; n = 4096: the 4 distinct ones in the order of
	// each register kernel, radix 16 at [0, 4n), radix 8 at [4n, 8n); lengths
	// that are not a power of two: the 7 tables and the Bluestein tables
	if (const uint32_t M = nlc::dnls_bluestein_size(n)) {
		tab.resize(size_t(7) * n + size_t(2) * M + n);
		nlc::dnls_bluestein_tables(n, tab.data() + size_t(7) * n);
	} else if (const uint32_t L = nlc::bluestein_two_level_size(n); L && !nlc::mixed_radix_plan(n, nullptr)) {
		// Bluestein through the two-level FFT: the 7 natural tables (the
		// Bluestein tables are shared with the energy pass), the row-permuted
		// copy of bhat for 4096-point rows, and L values of scratch per chain
		if (const size_t extra = nlc::dnls_blue_row_table_size(n)) {
			tab.resize(size_t(7) * n + extra);
			nlc::dnls_blue_row_tables(n, tab.data() + size_t(7) * n);
		}
		const size_t need = size_t(L) * c->copies;
		if (c->fft_scratch_sites < need) {
			if (c->fft_scratch) (void) hipFree(c->fft_scratch);
			c->fft_scratch = nullptr;
			c->fft_scratch_sites = 0;
			NLC_TRY(dev_alloc(c, &c->fft_scratch, need, "DNLS Bluestein buffer"));
			c->fft_scratch_sites = need;
		}
	} else if (nlc::dnls_two_level(n)) {
		// the two-level FFT keeps the spectrum in HBM in [r][k] order: the
		// multipliers in that order, and a scratch of the state's size
		perm.resize(nlc::dnls_two_level_table_size(n));
		nlc::dnls_two_level_tables(n, tab.data(), perm.data());
		tab.swap(perm);
		if (c->fft_scratch_sites < c->total_sites) {
			if (c->fft_scratch) (void) hipFree(c->fft_scratch);
			c->fft_scratch = nullptr;
			c->fft_scratch_sites = 0;
			NLC_TRY(dev_alloc(c, &c->fft_scratch, c->total_sites, "DNLS spectrum"));
			c->fft_scratch_sites = c->total_sites;
		}
	} else if (nlc::dnls_fast_path(n)) {
		perm.resize(size_t(8) * n);
		nlc::dnls_permute_tables(n, nlc::kDnlsRadix16, tab.data(), perm.data());
		nlc::dnls_permute_tables(n, nlc::kDnlsRadix8, tab.data(), perm.data() + size_t(4) * n);
		tab.swap(perm);
	}
	if (p.precision == NLC_FP32C) {
		// FP32 fields with compensated accumulation: the radix-16 register kernel
		if (!nlc::dnls_fast_path(n) || (p.sites_per_lane && p.sites_per_lane != 16))
			return c->fail(NLC_EUNSUPPORTED, "DNLS in FP32C precision runs 4096-site chains on the radix-16 kernel "
			                                 "(chain_length %u, sites_per_lane %u)", n, p.sites_per_lane);
		std::vector<float2> tw32(n);
		std::vector<float4> mult32(size_t(4) * n);
		nlc::dnls_fp32c_tables(n, p.dt, tw32.data(), mult32.data());
		NLC_TRY(dev_alloc(c, &c->dnls_tw32, tw32.size(), "DNLS FP32 twiddles"));
		NLC_TRY(dev_alloc(c, &c->dnls_mult32, mult32.size(), "DNLS FP32C multipliers"));
		NLC_HIP(c, hipMemcpy(c->dnls_tw32, tw32.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
		NLC_HIP(c, hipMemcpy(c->dnls_mult32, mult32.data(), sizeof(float4) * mult32.size(), hipMemcpyHostToDevice));
		c->dnls_variant = nlc::kDnlsRadix16;
		return NLC_OK;
	}
	// sites_per_lane selects the register kernel: values per thread, 16 or 8
	if (p.sites_per_lane == 8 && nlc::dnls_fast_path(n)) c->dnls_variant = nlc::kDnlsRadix8;
	else if (p.sites_per_lane == 0 || (p.sites_per_lane == 16 && nlc::dnls_fast_path(n))) c->dnls_variant = nlc::kDnlsRadix16;
	else return c->fail(NLC_EUNSUPPORTED, "DNLS: sites_per_lane %u has no kernel for chain_length %u", p.sites_per_lane, n);
	NLC_TRY(dev_alloc(c, &c->dnls_table, tab.size(), "DNLS multipliers"));
	NLC_HIP(c, hipMemcpy(c->dnls_table, tab.data(), sizeof(double2) * tab.size(), hipMemcpyHostToDevice));
	return NLC_OK;
}

// Step kernel of a uniform real-model context (the reference's move<>
// dispatch, DNKG_FPUT_Toda.cu:254-293).
// FP32-compensated real models (the DNLS form has its own kernel).
bool is_c32(const nlc_ctx *c) { return c->p.precision == NLC_FP32C && c->p.model != NLC_DNLS; }

int select_uniform(nlc_ctx *c) {
	const nlc_params &p = c->p;
	const uint32_t n = c->n;
	const bool c32 = is_c32(c);
	bool want_split = p.kernel == NLC_KERNEL_SPLIT, want_tiled = p.kernel == NLC_KERNEL_TILED;
	c->use_split = c->use_tiled = false;
	if (c32 && (want_split || want_tiled))
		return c->fail(NLC_EUNSUPPORTED, "FP32C precision runs on the wave kernels (no split or tiled form)");
	if (c32 && !(is_pow2(n) && n <= 2048))
		return c->fail(NLC_EUNSUPPORTED, "FP32C precision of FPUT and Toda needs a power-of-two chain_length <= 2048");
	if (!want_split && !want_tiled) {
		int L = 0, S = 0;
		bool exact = false, have;
		if (p.lanes_per_chain || p.sites_per_lane) {
			L = int(p.lanes_per_chain);
			S = int(p.sites_per_lane);
			exact = uint64_t(L) * S == n;
			have = (L >= 1 && L <= 64 && is_pow2(uint32_t(L)) && S >= 1 && uint64_t(L) * S >= n &&
			        uint64_t(L) * (S - 1) < n) ||
			       group_shape_ok(n, L, S);
			if (!have) return c->fail(NLC_EARG, "wave shape %d lanes x %d sites cannot hold chain_length %u", L, S, n);
		} else {
			have = default_shape(n, c->copies, L, S, exact) || (!c32 && group_default(p.model, n, L, S, exact));
		}
		c->launcher = have ? lookup_prec(p.model, L, S, exact, c32) : nullptr;
		if (c->launcher) {
			c->L = L, c->S = S, c->exact = exact;
		} else if (p.lanes_per_chain || c32) {
			return c->fail(NLC_EUNSUPPORTED, "no wave-resident kernel for chain_length %u (shape %d x %d)", n, L, S);
		} else {
			// longer than a workgroup holds: segments in registers with temporal blocking
			want_tiled = true;
		}
	}
	if (want_tiled) {
		if (!(c->tiled = lookup_tiled(p.model)) || n < 256)
			return c->fail(NLC_EUNSUPPORTED, "no temporally blocked kernel for chain_length %u", n);
		if (!c->tile_buf) NLC_TRY(dev_alloc(c, &c->tile_buf, size_t(n) * c->copies, "second state buffer"));
		c->use_tiled = true;
		return NLC_OK;
	}
	if (want_split) {
		c->use_split = true;
		for (int b = 0; b < 2; ++b) {
			if (!c->split_q[b]) NLC_TRY(dev_alloc(c, &c->split_q[b], size_t(n) * c->copies, "split q"));
			if (!c->split_p[b]) NLC_TRY(dev_alloc(c, &c->split_p[b], size_t(n) * c->copies, "split p"));
		}
	}
	return NLC_OK;
}

// Ragged: the persistent kernel with S sites per lane (items built here), or
// one uniform kernel per class when S is 0 or no kernel fits every class.
int select_ragged(nlc_ctx *c, int S, bool half_S) {
	if (c->items) {
		(void) hipFree(c->items);
		c->items = nullptr;
	}
	uint32_t kind_mask = 0;
	for (int s = 0; s < 3; ++s)
		if (c->pargs.kind[s] >= 0) kind_mask |= 1u << c->pargs.kind[s];
	const bool c32 = is_c32(c);
	c->persistent = S && ragged_fits(c->cls, S, half_S) ? lookup_persistent(S, kind_mask, c32) : nullptr;
	if (c->persistent) {
		struct Costed {
			nlc::RaggedItem it;
			double cost;
		};
		std::vector<Costed> all;
		for (const auto &k : c->cls) {
			const int s_item = item_S(k.n, S, half_S);
			const uint32_t L = k.n / uint32_t(s_item), cpw = 64 / L;
			const uint8_t flags = uint8_t(k.set | (s_item < S ? nlc::kItemHalfS : 0));
			const double cost = s_item * kKindCost[c32][c->pargs.kind[k.set]];
			for (uint32_t i = 0; i < k.count; i += cpw)
				all.push_back({{k.site0 + uint64_t(i) * k.n, k.n, uint16_t(std::min(cpw, k.count - i)),
				                uint8_t(log2u(L)), flags},
				               cost});
		}
		// most expensive first (stable: within a cost, device order)
		std::stable_sort(all.begin(), all.end(), [](const Costed &a, const Costed &b) { return a.cost > b.cost; });
		std::vector<nlc::RaggedItem> items(all.size());
		for (size_t i = 0; i < all.size(); ++i) items[i] = all[i].it;
		c->nitems = uint32_t(items.size());
		NLC_TRY(dev_alloc(c, &c->items, items.size(), "ragged items"));
		NLC_HIP(c, hipMemcpy(c->items, items.data(), sizeof(nlc::RaggedItem) * items.size(), hipMemcpyHostToDevice));
		if (!c->queue) {
			NLC_TRY(dev_alloc(c, &c->queue, 2, "persistent queue"));
			NLC_HIP(c, hipMemset(c->queue, 0, 2 * sizeof(uint32_t)));
		}
		c->ragged_S = S;
		c->ragged_half = half_S;
		return NLC_OK;
	}
	c->ragged_S = 0;
	c->ragged_half = false;
	c->nitems = 0;
	for (auto &k : c->cls) {
		if (!default_shape(k.n, k.count, k.L, k.S, k.exact) ||
		    !(k.launcher = lookup_prec(k.model, k.L, k.S, k.exact, c32)))
			return c->fail(NLC_EUNSUPPORTED,
			               c32 ? "FP32C ragged ensembles need power-of-two chain lengths <= 2048 (got %u)"
			                   : "ragged ensembles need chain lengths with a wave shape (got %u)",
			               k.n);
	}
	return NLC_OK;
}

int enqueue_move(nlc_ctx *c, uint32_t steps) {
	if (!steps) return c->fail(NLC_EARG, "--kernel_batching must be a positive number");
	if (c->p.model == NLC_DNLS && c->p.precision == NLC_FP32C) {
		NLC_HIP(c, nlc::launch_dnls_fp32c(c->stream, c->state, c->copies, steps, c->theta, c->dnls_tw32,
		                                  c->dnls_mult32));
	} else if (c->p.model == NLC_DNLS) {
		const bool r8 = nlc::dnls_fast_path(c->n) && c->dnls_variant == nlc::kDnlsRadix8;
		NLC_HIP(c, nlc::launch_dnls(c->stream, c->state, c->n, c->copies, steps, c->theta, c->tw,
		                            c->dnls_table + (r8 ? size_t(4) * c->n : 0), c->dnls_variant, c->fft_scratch,
		                            c->blue_table));
	} else if (c->ragged) {
		if (c->persistent) {
			NLC_HIP(c, c->persistent(c->stream, c->state, c->items, c->nitems, steps, c->pargs, c->queue, 0));
		} else {
			const bool fork = c->aux_stream != nullptr;
			if (fork) {
				NLC_HIP(c, hipEventRecord(c->fork_ev, c->stream));
				NLC_HIP(c, hipStreamWaitEvent(c->aux_stream, c->fork_ev, 0));
			}
			for (const auto &k : c->cls)
				NLC_HIP(c, k.launcher(fork && k.side ? c->aux_stream : c->stream, c->state + k.site0, k.count, k.n,
				                      steps, c->pargs.a[k.set], nullptr));
			if (fork) {
				NLC_HIP(c, hipEventRecord(c->join_ev, c->aux_stream));
				NLC_HIP(c, hipStreamWaitEvent(c->stream, c->join_ev, 0));
			}
		}
	} else if (c->use_tiled) {
		NLC_HIP(c, c->tiled(c->stream, c->state, c->tile_buf, c->copies, c->n, steps, c->args, c->mp2));
	} else if (c->use_split) {
		int cur = 0;
		NLC_HIP(c, nlc::launch_planar_to_split(c->stream, c->state, c->split_q[0], c->split_p[0], c->copies, c->n));
		NLC_HIP(c, nlc::launch_step_split(c->stream, c->p.model, c->split_q, c->split_p, cur, c->copies, c->n, steps,
		                                  c->args, c->mp2));
		NLC_HIP(c, nlc::launch_split_to_planar(c->stream, c->split_q[cur], c->split_p[cur], c->state, c->copies, c->n));
	} else {
		NLC_HIP(c, c->launcher(c->stream, c->state, c->copies, c->n, steps, c->args, c->mp2));
	}
	return NLC_OK;
}

nlc::LongChainFft long_chain_fft(const nlc_ctx *c) {
	nlc::LongChainFft lf;
	lf.scratch = c->fft_scratch;
	lf.scratch_sites = c->fft_scratch_sites;
	lf.blue = c->blue_table;
	return lf;
}

// results::calc_linenergies normalisation of the drivers for a class
double default_norm(int model, const Klass &k) {
	switch (model) {
	case NLC_DDNKG: return 0.5 / k.count;                // dDNKG.cpp:141
	case NLC_DNLS: return 1.0 / (double(k.n) * k.count); // DNLS.cpp:150
	default: return (0.5 / k.count) / k.n;               // DNKG_FPUT_Toda.cpp:129
	}
}

// Energy pass of every class: raw sums over copies into e_raw and, with dst,
// the normalisation and entropies written to dst (pinned host memory,
// results_len doubles).  The classes share the partial rows, so each class is
// reduced before the next one's pass.
int enqueue_energy(nlc_ctx *c, double norm = 0.0, double *dst = nullptr) {
	for (size_t i = 0; i < c->cls.size(); ++i) {
		const Klass &k = c->cls[i];
		const double2 *st = c->state + k.site0;
		const double *scale = nullptr;
		switch (c->p.model) {
		case NLC_DNKG:
		case NLC_FPUT:
		case NLC_TODA:
			NLC_HIP(c, nlc::launch_energy_real(c->stream, st, k.n, k.count, c->tw + k.tab0, c->omega + k.tab0,
			                                   c->partial, k.rows, long_chain_fft(c)));
			break;
		case NLC_DDNKG:
			NLC_HIP(c, nlc::launch_energy_ddnkg(c->stream, st, k.n, k.count, c->eigvec, c->omega, c->partial,
			                                    c->proj_variant, c->vplanes));
			break;
		case NLC_DNLS:
			NLC_HIP(c, nlc::launch_energy_dnls(c->stream, st, k.n, k.count, c->tw + k.tab0, c->partial, k.rows,
			                                   long_chain_fft(c)));
			scale = c->omega + k.tab0;
			break;
		default: return c->fail(NLC_EUNSUPPORTED, "no energy pass for this model");
		}
		double *er = c->e_raw + k.tab0;
		if (dst) {
			const double nm = norm > 0 ? norm : default_norm(k.model, k);
			NLC_HIP(c, nlc::launch_reduce_results(c->stream, c->partial, k.rows, k.n, scale, er, c->reduce_scratch, nm,
			                                      k.a0is0, c->modes, c->ragged ? 0 : c->nmodes,
			                                      dst + k.tab0 + 2 * i));
		} else {
			NLC_HIP(c, nlc::launch_reduce_partial(c->stream, c->partial, k.rows, k.n, scale, er, c->reduce_scratch));
		}
	}
	return NLC_OK;
}

void unpack_results(const nlc_ctx *c, const double *src, double *linenergies, double *entropies) {
	for (size_t i = 0; i < c->cls.size(); ++i) {
		const Klass &k = c->cls[i];
		const double *r = src + k.tab0 + 2 * i;
		if (linenergies) std::memcpy(linenergies + k.tab0, r, sizeof(double) * k.n);
		if (entropies) entropies[2 * i] = r[k.n], entropies[2 * i + 1] = r[k.n + 1];
	}
}

int reject(int code, const char *msg) {
	g_create_error = msg;
	return code;
}

int validate_params(const nlc_params *p, bool ragged) {
	if (!p || p->struct_size != sizeof(nlc_params)) return reject(NLC_EARG, "params NULL or struct_size mismatch");
	if (p->model < NLC_DNKG || p->model > NLC_DNLS) return reject(NLC_EARG, "unknown model");
	if (!ragged) {
		if (p->chain_length < 2) return reject(NLC_EARG, "--chain_length must be >= 2");
		if (!p->copies) return reject(NLC_EARG, "--copies must be a positive number");
		// every launch has at most about one work-item per site (per-chain kernels
		// with more are grid-strided or capped), so launch sizes stay below 2^32
		if (uint64_t(p->chain_length) * p->copies > (1ULL << 31))
			return reject(NLC_EARG, "(--chain_length * --copies) must be <= 2^31");
	}
	if (!(p->dt > 0)) return reject(NLC_EARG, "--dt must be a positive number");
	if (p->model == NLC_TODA && p->alpha == 0) return reject(NLC_EARG, "Toda needs --alpha != 0");
	if (p->model == NLC_DDNKG && !p->masses) return reject(NLC_EARG, "dDNKG needs the masses (-m file content)");
	if (p->precision != NLC_FP64 && p->precision != NLC_FP32C) return reject(NLC_EARG, "unknown precision");
	if (p->precision == NLC_FP32C && p->model != NLC_DNLS && p->model != NLC_FPUT && p->model != NLC_TODA)
		return reject(NLC_EUNSUPPORTED, "FP32C precision is implemented for FPUT, Toda and DNLS");
	if (p->kernel < NLC_KERNEL_AUTO || p->kernel > NLC_KERNEL_TILED) return reject(NLC_EARG, "unknown kernel family");
	if (p->projection < NLC_PROJ_AUTO || p->projection > NLC_PROJ_MFMA_F64_WIDE)
		return reject(NLC_EARG, "unknown projection form");
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
		return reject(NLC_EHIP, "no HIP device available (the hot path has no CPU fallback)");
	return NLC_OK;
}

int finish_create(nlc_ctx *c, nlc_ctx **out, int rc) {
	if (rc) {
		g_create_error = c->err;
		free_all(c);
		delete c;
		return rc;
	}
	*out = c;
	return NLC_OK;
}


// The cached graph of one batch for result destination `which`: move, energy
// pass, normalisation and entropies.
int ensure_batch_graph(nlc_ctx *c, uint32_t which, uint32_t steps, double norm, double *dst) {
	if (!steps) return c->fail(NLC_EARG, "--kernel_batching must be a positive number");
	nlc_ctx::BatchGraph &bg = c->graphs[which];
	if (!bg.exec || bg.steps != steps || bg.norm != norm) {
		if (bg.exec) {
			if (c->slot_count) NLC_HIP(c, hipStreamSynchronize(c->stream));
			(void) hipGraphExecDestroy(bg.exec);
			(void) hipGraphDestroy(bg.graph);
			bg = nlc_ctx::BatchGraph();
		}
		NLC_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
		int rc = enqueue_move(c, steps);
		if (!rc) rc = enqueue_energy(c, norm, dst);
		hipGraph_t g = nullptr;
		const hipError_t e = hipStreamEndCapture(c->stream, &g);
		if (rc) {
			if (g) (void) hipGraphDestroy(g);
			return rc;
		}
		if (e != hipSuccess) return c->hip_fail(e, "hipStreamEndCapture");
		bg.graph = g;
		NLC_HIP(c, hipGraphInstantiate(&bg.exec, bg.graph, nullptr, nullptr, 0));
		bg.steps = steps;
		bg.norm = norm;
	}
	return NLC_OK;
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
	if (!out) return reject(NLC_EARG, "out is NULL");
	*out = nullptr;
	if (int rc = validate_params(p, false)) return rc;
	nlc_ctx *c = new (std::nothrow) nlc_ctx;
	if (!c) return reject(NLC_ENOMEM, "context allocation failed");
	c->p = *p;
	c->proj_variant = p->projection == NLC_PROJ_VALU_F64   ? nlc::kProjValu
	                  : p->projection == NLC_PROJ_SPLIT_BF16 ? nlc::kProjBf16
	                  : p->projection == NLC_PROJ_MFMA_F64   ? nlc::kProjMfma
	                                                         : nlc::kProjMfmaWide; // auto: the faster measured form
	c->n = p->chain_length;
	c->copies = p->copies;
	c->total_sites = uint64_t(c->n) * c->copies;
	c->table_len = c->n;
	Klass k;
	k.n = c->n;
	k.count = c->copies;
	c->cls.push_back(k);
	if (p->masses) c->masses.assign(p->masses, p->masses + c->n);
	c->p.masses = nullptr;
	int rc = setup_common(c);
	if (!rc) rc = p->model == NLC_DNLS ? setup_dnls(c) : select_uniform(c);
	if (!rc) rc = nlc_set_entropy_mask(c, nullptr, c->a0is0);
	return finish_create(c, out, rc);
}

} // extern "C"

namespace {

// Ragged and mixed ensembles: chain i has length chain_lengths[i] and
// parameter set chain_set[i] (NULL: all set 0).  Classes are the distinct
// (set, length) pairs in ascending order, each contiguous on the device.
int create_classes(nlc_ctx **out, const nlc_params *params, uint32_t nsets, const uint32_t *chain_lengths,
                   const uint32_t *chain_set, uint32_t nchains) {
	g_create_error.clear();
	if (!out) return reject(NLC_EARG, "out is NULL");
	*out = nullptr;
	if (!chain_lengths || !nchains) return reject(NLC_EARG, "no chains");
	if (!params || nsets < 1 || nsets > 3) return reject(NLC_EARG, "1 to 3 parameter sets");
	for (uint32_t s = 0; s < nsets; ++s) {
		if (int rc = validate_params(&params[s], true)) return rc;
		if (params[s].model != NLC_DNKG && params[s].model != NLC_FPUT && params[s].model != NLC_TODA)
			return reject(NLC_EUNSUPPORTED, "ragged and mixed ensembles are available for DNKG, FPUT and Toda");
		if (params[s].precision != params[0].precision)
			return reject(NLC_EARG, "the parameter sets of a mixed ensemble share one precision");
		if (params[s].precision == NLC_FP32C && params[s].model == NLC_DNKG)
			return reject(NLC_EUNSUPPORTED, "FP32C ragged and mixed ensembles hold FPUT and Toda chains");
	}
	uint64_t total = 0;
	for (uint32_t i = 0; i < nchains; ++i) {
		if (chain_lengths[i] < 2) return reject(NLC_EARG, "every chain length must be >= 2");
		if (chain_set && chain_set[i] >= nsets) return reject(NLC_EARG, "chain parameter set out of range");
		total += chain_lengths[i];
	}
	if (total > (1ULL << 31)) return reject(NLC_EARG, "total number of sites must be <= 2^31");
	nlc_ctx *c = new (std::nothrow) nlc_ctx;
	if (!c) return reject(NLC_ENOMEM, "context allocation failed");
	c->p = params[0];
	c->p.masses = nullptr;
	c->sets.assign(params, params + nsets);
	c->ragged = true;
	c->lengths.assign(chain_lengths, chain_lengths + nchains);
	c->copies = nchains;
	c->total_sites = total;
	// classes in ascending (set, length); inside a class the caller's order
	auto key = [&](uint32_t i) { return (uint64_t(chain_set ? chain_set[i] : 0) << 32) | chain_lengths[i]; };
	std::vector<uint64_t> distinct(nchains);
	for (uint32_t i = 0; i < nchains; ++i) distinct[i] = key(i);
	std::sort(distinct.begin(), distinct.end());
	distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
	for (uint64_t kv : distinct) {
		Klass k;
		k.n = uint32_t(kv);
		k.set = int(kv >> 32);
		k.tab0 = c->table_len;
		c->table_len += k.n;
		c->n = std::max(c->n, k.n);
		c->cls.push_back(k);
	}
	std::vector<size_t> class_of(nchains);
	for (uint32_t i = 0; i < nchains; ++i) {
		class_of[i] = std::lower_bound(distinct.begin(), distinct.end(), key(i)) - distinct.begin();
		c->cls[class_of[i]].count++;
	}
	uint64_t site = 0, chain = 0;
	for (auto &k : c->cls) {
		k.site0 = site;
		k.chain0 = chain;
		site += uint64_t(k.n) * k.count;
		chain += k.count;
	}
	c->caller_site0.resize(nchains);
	c->device_site0.resize(nchains);
	c->device_chain.resize(nchains);
	std::vector<uint32_t> filled(c->cls.size(), 0);
	uint64_t caller = 0;
	for (uint32_t i = 0; i < nchains; ++i) {
		Klass &k = c->cls[class_of[i]];
		const uint32_t f = filled[class_of[i]]++;
		c->caller_site0[i] = caller;
		c->device_site0[i] = k.site0 + uint64_t(f) * k.n;
		c->device_chain[i] = uint32_t(k.chain0 + f);
		caller += c->lengths[i];
	}
	int rc = setup_common(c);
	if (!rc && c->cls.size() > 1) {
		// split the classes over the two streams, longest processing time first
		// (cost ~ sites of the class times its model's instructions per site)
		std::vector<size_t> order(c->cls.size());
		for (size_t i = 0; i < order.size(); ++i) order[i] = i;
		auto sites = [&](size_t i) {
			return double(c->cls[i].n) * c->cls[i].count * kKindCost[is_c32(c)][c->pargs.kind[c->cls[i].set]];
		};
		std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return sites(a) > sites(b); });
		double load[2] = {0, 0};
		for (size_t i : order) {
			const int side = load[1] < load[0];
			c->cls[i].side = side;
			load[side] += sites(i);
		}
		if (hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking) != hipSuccess ||
		    hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
		    hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess)
			rc = c->fail(NLC_EHIP, "stream or event creation failed");
	}
	if (!rc) {
		// the persistent kernel with the smallest S that fits (most chains per
		// wave), items at S / 2 sites per lane where they fit; else per class
		int S = 0;
		for (int s : {32, 16, 8})
			if (ragged_fits(c->cls, s, true)) S = s;
		rc = select_ragged(c, S, true);
	}
	return finish_create(c, out, rc);
}

} // namespace

extern "C" {

int nlc_create_ragged(nlc_ctx **out, const nlc_params *p, const uint32_t *chain_lengths, uint32_t nchains) {
	return create_classes(out, p, 1, chain_lengths, nullptr, nchains);
}

int nlc_create_mixed(nlc_ctx **out, const nlc_params *params, uint32_t nsets, const uint32_t *chain_lengths,
                     const uint32_t *chain_set, uint32_t nchains) {
	if (!chain_set) return reject(NLC_EARG, "chain_set is NULL");
	return create_classes(out, params, nsets, chain_lengths, chain_set, nchains);
}

int nlc_ragged_class_sets(const nlc_ctx *c, uint32_t *sets, uint32_t cap) {
	if (!c || (!sets && cap)) return NLC_EARG;
	for (uint32_t i = 0; i < std::min<uint32_t>(cap, uint32_t(c->cls.size())); ++i) sets[i] = uint32_t(c->cls[i].set);
	return NLC_OK;
}



int nlc_ragged_classes(const nlc_ctx *c, uint32_t *nclasses, uint32_t *lengths, uint32_t *counts, uint32_t cap) {
	if (!c || !nclasses) return NLC_EARG;
	*nclasses = uint32_t(c->cls.size());
	for (uint32_t i = 0; i < std::min<uint32_t>(cap, *nclasses); ++i) {
		if (lengths) lengths[i] = c->cls[i].n;
		if (counts) counts[i] = c->cls[i].count;
	}
	return NLC_OK;
}

void nlc_destroy(nlc_ctx *ctx) {
	if (!ctx) return;
	free_all(ctx);
	delete ctx;
}

// Host <-> device copies of the shard.  Chunks of kXferSites sites alternate
// between two pinned buffers: the DMA of one chunk overlaps the host-side copy
// (several threads; for ragged ensembles the permutation between the caller's
// chain order and the device's class-major order) of the other.
constexpr size_t kXferSites = size_t(1) << 19; // 8 MiB of double2

static int ensure_pinned(nlc_ctx *c) {
	for (int i = 0; i < 2; ++i) {
		if (!c->pinned[i] && hipHostMalloc((void **) &c->pinned[i], sizeof(double2) * kXferSites) != hipSuccess)
			return c->fail(NLC_ENOMEM, "pinned staging buffer allocation failed");
		if (!c->pinned_ev[i] && hipEventCreateWithFlags(&c->pinned_ev[i], hipEventDisableTiming) != hipSuccess)
			return c->fail(NLC_EHIP, "event creation failed");
	}
	if (c->ragged && c->device_order.empty()) {
		c->device_order.resize(c->lengths.size());
		for (uint32_t i = 0; i < c->device_order.size(); ++i) c->device_order[i] = i;
		std::sort(c->device_order.begin(), c->device_order.end(),
		          [c](uint32_t a, uint32_t b) { return c->device_site0[a] < c->device_site0[b]; });
	}
	return NLC_OK;
}

// Copies the device-order sites [d0, d0 + len) between the pinned chunk and the
// caller's buffer (to_host: pinned -> caller), on up to 8 threads.
static void host_chunk(const nlc_ctx *c, double2 *pinned, double2 *caller, uint64_t d0, size_t len, bool to_host) {
	auto copy = [&](uint64_t a, uint64_t b) { // device-order sub-range [a, b)
		if (!c->ragged) {
			if (to_host) std::memcpy(caller + a, pinned + (a - d0), sizeof(double2) * (b - a));
			else std::memcpy(pinned + (a - d0), caller + a, sizeof(double2) * (b - a));
			return;
		}
		// chains overlapping [a, b), in device order
		auto it = std::upper_bound(c->device_order.begin(), c->device_order.end(), a,
		                           [c](uint64_t x, uint32_t i) { return x < c->device_site0[i]; });
		if (it != c->device_order.begin()) --it;
		for (; it != c->device_order.end() && c->device_site0[*it] < b; ++it) {
			const uint32_t i = *it;
			const uint64_t s0 = std::max<uint64_t>(a, c->device_site0[i]);
			const uint64_t s1 = std::min<uint64_t>(b, c->device_site0[i] + c->lengths[i]);
			if (s0 >= s1) continue;
			double2 *u = caller + c->caller_site0[i] + (s0 - c->device_site0[i]);
			if (to_host) std::memcpy(u, pinned + (s0 - d0), sizeof(double2) * (s1 - s0));
			else std::memcpy(pinned + (s0 - d0), u, sizeof(double2) * (s1 - s0));
		}
	};
	const unsigned nt = len >= (size_t(1) << 16) ? std::max(1u, std::min(8u, std::thread::hardware_concurrency())) : 1u;
	if (nt == 1) {
		copy(d0, d0 + len);
		return;
	}
	std::vector<std::thread> pool;
	for (unsigned t = 0; t < nt; ++t)
		pool.emplace_back([&, t] { copy(d0 + len * t / nt, d0 + len * (t + 1) / nt); });
	for (auto &th : pool) th.join();
}

int nlc_upload(nlc_ctx *c, const void *state) {
	if (!c || !state) return NLC_EARG;
	NLC_TRY(ensure_pinned(c));
	double2 *src = const_cast<double2 *>(static_cast<const double2 *>(state));
	const size_t chunks = (c->total_sites + kXferSites - 1) / kXferSites;
	for (size_t k = 0; k < chunks; ++k) {
		const int b = int(k & 1);
		const uint64_t d0 = k * kXferSites;
		const size_t len = std::min<uint64_t>(kXferSites, c->total_sites - d0);
		if (k >= 2) NLC_HIP(c, hipEventSynchronize(c->pinned_ev[b])); // the DMA of chunk k - 2 has read it
		host_chunk(c, c->pinned[b], src, d0, len, false);
		NLC_HIP(c, hipMemcpyAsync(c->state + d0, c->pinned[b], sizeof(double2) * len, hipMemcpyHostToDevice, c->stream));
		NLC_HIP(c, hipEventRecord(c->pinned_ev[b], c->stream));
	}
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	return NLC_OK;
}

int nlc_download(nlc_ctx *c, void *state) {
	if (!c || !state) return NLC_EARG;
	NLC_TRY(ensure_pinned(c));
	double2 *dst = static_cast<double2 *>(state);
	const size_t chunks = (c->total_sites + kXferSites - 1) / kXferSites;
	auto drain = [&](size_t k) -> int { // host side of chunk k, once its DMA is done
		const int b = int(k & 1);
		const uint64_t d0 = k * kXferSites;
		NLC_HIP(c, hipEventSynchronize(c->pinned_ev[b]));
		host_chunk(c, c->pinned[b], dst, d0, std::min<uint64_t>(kXferSites, c->total_sites - d0), true);
		return NLC_OK;
	};
	for (size_t k = 0; k < chunks; ++k) {
		const int b = int(k & 1);
		const uint64_t d0 = k * kXferSites;
		const size_t len = std::min<uint64_t>(kXferSites, c->total_sites - d0);
		NLC_HIP(c, hipMemcpyAsync(c->pinned[b], c->state + d0, sizeof(double2) * len, hipMemcpyDeviceToHost, c->stream));
		NLC_HIP(c, hipEventRecord(c->pinned_ev[b], c->stream));
		if (k >= 1) NLC_TRY(drain(k - 1)); // overlaps the DMA of chunk k
	}
	if (chunks) NLC_TRY(drain(chunks - 1));
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
	return enqueue_move(c, steps);
}

int nlc_set_entropy_mask(nlc_ctx *c, const uint8_t *mask, int a0is0) {
	if (!c) return NLC_EARG;
	std::vector<uint32_t> idx;
	if (mask) {
		if (c->ragged) return c->fail(NLC_EUNSUPPORTED, "entropy masks apply to uniform ensembles only");
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
	for (auto &k : c->cls) k.a0is0 = c->a0is0;
	drop_graph(c);
	return NLC_OK;
}

int nlc_linenergies_raw(nlc_ctx *c, double *e_raw) {
	if (!c || !e_raw) return NLC_EARG;
	NLC_TRY(enqueue_energy(c));
	NLC_HIP(c, hipMemcpyAsync(c->h_results, c->e_raw, sizeof(double) * c->table_len, hipMemcpyDeviceToHost, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	std::memcpy(e_raw, c->h_results, sizeof(double) * c->table_len);
	return NLC_OK;
}

int nlc_linenergies(nlc_ctx *c, double norm, double *linenergies, double *entropies) {
	if (!c) return NLC_EARG;
	NLC_TRY(enqueue_energy(c, norm, c->h_results));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	unpack_results(c, c->h_results, linenergies, entropies);
	return NLC_OK;
}


int nlc_batch(nlc_ctx *c, uint32_t steps, double norm, double *linenergies, double *entropies) {
	if (!c) return NLC_EARG;
	NLC_TRY(ensure_batch_graph(c, NLC_BATCH_SLOTS, steps, norm, c->h_results));
	NLC_HIP(c, hipGraphLaunch(c->graphs[NLC_BATCH_SLOTS].exec, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	unpack_results(c, c->h_results, linenergies, entropies);
	return NLC_OK;
}

int nlc_batch_submit(nlc_ctx *c, uint32_t steps, double norm) {
	if (!c) return NLC_EARG;
	if (c->slot_count == NLC_BATCH_SLOTS)
		return c->fail(NLC_EARG, "%d batches are already queued; collect one first", NLC_BATCH_SLOTS);
	if (!c->h_slots) {
		hipError_t e = hipHostMalloc((void **) &c->h_slots, sizeof(double) * c->results_len() * NLC_BATCH_SLOTS,
		                             hipHostMallocCoherent);
		if (e != hipSuccess) {
			c->h_slots = nullptr;
			return c->fail(NLC_ENOMEM, "hipHostMalloc(batch result slots): %s", hipGetErrorString(e));
		}
		for (hipEvent_t &ev : c->slot_ev) NLC_HIP(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
	}
	const uint32_t slot = c->slot_head;
	double *dst = c->h_slots + c->results_len() * slot;
	if (c->batch_direct) {
		NLC_TRY(enqueue_move(c, steps));
		NLC_TRY(enqueue_energy(c, norm, dst));
	} else {
		NLC_TRY(ensure_batch_graph(c, slot, steps, norm, dst));
		NLC_HIP(c, hipGraphLaunch(c->graphs[slot].exec, c->stream));
	}
	NLC_HIP(c, hipEventRecord(c->slot_ev[slot], c->stream));
	c->slot_head = (slot + 1) % NLC_BATCH_SLOTS;
	++c->slot_count;
	return NLC_OK;
}

int nlc_set_batch_mode(nlc_ctx *c, int direct) {
	if (!c || (direct != 0 && direct != 1)) return NLC_EARG;
	if (c->slot_count) return c->fail(NLC_ESTATE, "nlc_set_batch_mode: %u batches are queued", c->slot_count);
	c->batch_direct = direct == 1;
	return NLC_OK;
}

int nlc_batch_collect(nlc_ctx *c, double *linenergies, double *entropies) {
	if (!c) return NLC_EARG;
	if (!c->slot_count) return c->fail(NLC_EARG, "no batch is queued");
	const uint32_t slot = (c->slot_head + NLC_BATCH_SLOTS - c->slot_count) % NLC_BATCH_SLOTS;
	NLC_HIP(c, hipEventSynchronize(c->slot_ev[slot]));
	--c->slot_count;
	unpack_results(c, c->h_slots + c->results_len() * slot, linenergies, entropies);
	return NLC_OK;
}

int nlc_chain_energies(nlc_ctx *c, double *energies, double *norms) {
	if (!c || !energies) return NLC_EARG;
	for (const auto &k : c->cls)
		NLC_HIP(c, nlc::launch_chain_energy(c->stream, k.model, c->state + k.site0, k.n, k.count, c->pargs.a[k.set],
		                                    c->masses_d, c->energy_d + k.chain0, c->norm_d + k.chain0));
	std::vector<double> e(c->copies), nr(c->copies);
	NLC_HIP(c, hipMemcpyAsync(e.data(), c->energy_d, sizeof(double) * c->copies, hipMemcpyDeviceToHost, c->stream));
	NLC_HIP(c, hipMemcpyAsync(nr.data(), c->norm_d, sizeof(double) * c->copies, hipMemcpyDeviceToHost, c->stream));
	NLC_HIP(c, hipStreamSynchronize(c->stream));
	for (uint32_t i = 0; i < c->copies; ++i) {
		const uint32_t j = c->ragged ? c->device_chain[i] : i;
		energies[i] = e[j];
		if (norms && c->p.model == NLC_DNLS) norms[i] = nr[j];
	}
	return NLC_OK;
}

// Queued batches (nlc_batch_submit): time a pipeline of batches of
// trial_steps, three in flight as the CLI keeps them, as graph replays and as
// direct launches, on a copy of the state, and keep the faster.  Skipped while
// batches are queued.
static int tune_submission(nlc_ctx *c, uint32_t steps) {
	if (c->slot_count) return NLC_OK;
	double2 *backup = nullptr;
	NLC_TRY(dev_alloc(c, &backup, c->total_sites, "autotune backup"));
	int rc = NLC_OK;
	if (hipMemcpyAsync(backup, c->state, c->state_bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
		rc = c->fail(NLC_EHIP, "autotune setup failed");
	double us[2] = {0, 0};
	constexpr int kBatches = 16;
	for (int mode = 0; mode < 2 && !rc; ++mode) {
		c->batch_direct = mode == 1;
		for (int rep = 0; rep < 2 && !rc; ++rep) { // the first run builds the graphs and warms up
			if (hipStreamSynchronize(c->stream) != hipSuccess) rc = c->fail(NLC_EHIP, "autotune sync failed");
			const auto t0 = std::chrono::steady_clock::now();
			for (int i = 0; i < kBatches && !rc; ++i) {
				if (c->slot_count == 3) rc = nlc_batch_collect(c, nullptr, nullptr);
				if (!rc) rc = nlc_batch_submit(c, steps, 0.0);
			}
			while (c->slot_count && !rc) rc = nlc_batch_collect(c, nullptr, nullptr);
			us[mode] = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() /
			           kBatches;
		}
	}
	if (c->slot_count) { // only after an error: drop what is still queued
		(void) hipStreamSynchronize(c->stream);
		c->slot_count = 0;
	}
	if (hipMemcpyAsync(c->state, backup, c->state_bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess && !rc)
		rc = c->fail(NLC_EHIP, "autotune restore failed");
	(void) hipStreamSynchronize(c->stream);
	(void) hipFree(backup);
	if (rc) return rc;
	c->batch_direct = us[1] < us[0];
	char buf[96];
	snprintf(buf, sizeof buf, "; queued batches: graph %.1f us%s direct %.1f us%s", us[0], c->batch_direct ? "" : "*",
	         us[1], c->batch_direct ? "*" : "");
	c->tune_report += buf;
	return NLC_OK;
}

// dDNKG: time the energy pass with each projection form (it only reads the
// state) and keep the faster one.
static int tune_projection(nlc_ctx *c) {
	hipEvent_t e0 = nullptr, e1 = nullptr;
	int rc = NLC_OK;
	const int forms[3] = {nlc::kProjMfma, nlc::kProjMfmaWide, nlc::kProjValu};
	const char *names[3] = {"mfma", "mfma_wide", "valu"};
	float ms[3] = {0, 0, 0};
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
		rc = c->fail(NLC_EHIP, "autotune: event creation failed");
	for (int f = 0; f < 3 && !rc; ++f) {
		c->proj_variant = forms[f];
		for (int rep = 0; rep < 2 && !rc; ++rep) { // the first run warms up
			rc = enqueue_energy(c);
			if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = c->fail(NLC_EHIP, "event record");
			for (int i = 0; i < 3 && !rc; ++i) rc = enqueue_energy(c);
			if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
			            hipEventElapsedTime(&ms[f], e0, e1) != hipSuccess))
				rc = c->fail(NLC_EHIP, "autotune timing failed");
		}
	}
	if (e0) (void) hipEventDestroy(e0);
	if (e1) (void) hipEventDestroy(e1);
	if (rc) return rc;
	int best = 0;
	for (int f = 1; f < 3; ++f)
		if (ms[f] < ms[best]) best = f;
	c->proj_variant = forms[best];
	std::string rep = "; projection:";
	char buf[64];
	for (int f = 0; f < 3; ++f) {
		snprintf(buf, sizeof buf, " %s %.3f ms%s", names[f], ms[f] / 3, f == best ? "*" : "");
		rep += buf;
	}
	c->tune_report += rep;
	return NLC_OK;
}

int nlc_autotune(nlc_ctx *c, uint32_t trial_steps, char *report, size_t len) {
	if (!c || !trial_steps) return NLC_EARG;
	struct Cand {
		int L, S;
		bool exact, split;
		int ragged_S;
		float ms;
		int dnls;
		bool tiled;
		bool half; // persistent kernel: the classes that fit at S/2 run there
	};
	std::vector<Cand> cands;
	if (c->p.model == NLC_DNLS) {
		if (nlc::dnls_fast_path(c->n) && c->p.precision != NLC_FP32C)
			for (int v : {nlc::kDnlsRadix16, nlc::kDnlsRadix8}) cands.push_back({0, 0, false, false, 0, 0, v});
		else
			c->tune_report = "DNLS: a single kernel for this chain length, nothing to tune";
	} else if (c->ragged) {
		for (int S : {8, 16, 32})
			for (bool half : {false, true}) {
				uint32_t kinds = 0;
				for (int s = 0; s < 3; ++s)
					if (c->pargs.kind[s] >= 0) kinds |= 1u << c->pargs.kind[s];
				if (ragged_fits(c->cls, S, half) && lookup_persistent(S, kinds, is_c32(c)))
					cands.push_back({0, 0, false, false, S, 0, 0, false, half});
			}
		bool per_class = true; // one uniform kernel per length class, launched in sequence
		for (const auto &k : c->cls) {
			int L, S;
			bool exact;
			per_class = per_class && default_shape(k.n, k.count, L, S, exact) &&
			            lookup_prec(k.model, L, S, exact, is_c32(c));
		}
		if (per_class) cands.push_back({0, 0, false, false, 0, 0, 0});
	} else {
		for (const Shape &s : wave_shapes(c->p.model, c->n, is_c32(c)))
			cands.push_back({s.L, s.S, s.exact, false, 0, 0, 0});
		if (!is_c32(c)) {
			if (c->n > 2048 && lookup_tiled(c->p.model)) cands.push_back({0, 0, false, false, 0, 0, 0, true});
			cands.push_back({0, 0, false, true, 0, 0, 0});
		}
	}
	if (!cands.empty()) {
		drop_graph(c);
		double2 *backup = nullptr;
		NLC_TRY(dev_alloc(c, &backup, c->total_sites, "autotune backup"));
		hipEvent_t e0 = nullptr, e1 = nullptr;
		int rc = NLC_OK;
		if (hipEventCreate(&e0) || hipEventCreate(&e1) ||
		    hipMemcpyAsync(backup, c->state, c->state_bytes, hipMemcpyDeviceToDevice, c->stream))
			rc = c->fail(NLC_EHIP, "autotune setup failed");
		const nlc_params saved = c->p;
		for (auto &cd : cands) {
			if (rc) break;
			if (c->p.model == NLC_DNLS) {
				c->dnls_variant = cd.dnls;
			} else if (c->ragged) {
				rc = select_ragged(c, cd.ragged_S, cd.half);
			} else {
				c->p.kernel = cd.split ? NLC_KERNEL_SPLIT : cd.tiled ? NLC_KERNEL_TILED : NLC_KERNEL_WAVE;
				c->p.lanes_per_chain = uint32_t(cd.L);
				c->p.sites_per_lane = uint32_t(cd.S);
				rc = select_uniform(c);
			}
			// the first run warms up and sizes the second: short trials (small
			// ensembles) are repeated back to back until the timed region is
			// about 0.2 ms, so that launch gaps and timer resolution do not
			// decide between shapes; ms is the time per trial launch
			int nrep = 1;
			for (int rep = 0; rep < 2 && !rc; ++rep) {
				rc = enqueue_move(c, trial_steps);
				if (!rc && (hipEventRecord(e0, c->stream) != hipSuccess)) rc = c->fail(NLC_EHIP, "event record");
				for (int i = 0; i < nrep && !rc; ++i) rc = enqueue_move(c, trial_steps);
				float ms = 0;
				if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
				            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
					rc = c->fail(NLC_EHIP, "autotune timing failed");
				cd.ms = ms / float(nrep);
				if (rep == 0 && cd.ms > 0) nrep = std::max(1, std::min(256, int(0.2f / cd.ms + 0.5f)));
			}
			if (hipMemcpyAsync(c->state, backup, c->state_bytes, hipMemcpyDeviceToDevice, c->stream) && !rc)
				rc = c->fail(NLC_EHIP, "autotune restore failed");
		}
		(void) hipStreamSynchronize(c->stream);
		if (e0) (void) hipEventDestroy(e0);
		if (e1) (void) hipEventDestroy(e1);
		(void) hipFree(backup);
		c->p = saved;
		if (rc) return rc;
		const Cand *best = &cands[0];
		for (const auto &cd : cands)
			if (cd.ms < best->ms) best = &cd;
		std::string rep = c->p.model == NLC_DNLS ? "DNLS 4096-point kernels:"
		                  : c->ragged            ? "ragged sites-per-lane:"
		                                         : "wave shapes (lanes x sites):";
		char buf[96];
		for (const auto &cd : cands) {
			const char *mark = &cd == best ? "*" : "";
			if (c->p.model == NLC_DNLS)
				snprintf(buf, sizeof buf, " %s %.3f ms%s", cd.dnls == nlc::kDnlsRadix8 ? "radix8x512" : "radix16x256", cd.ms,
				         mark);
			else if (c->ragged && cd.ragged_S)
				snprintf(buf, sizeof buf, " S=%d%s %.3f ms%s", cd.ragged_S, cd.half ? "/half" : "", cd.ms, mark);
			else if (c->ragged) snprintf(buf, sizeof buf, " per-class %.3f ms%s", cd.ms, mark);
			else if (cd.split) snprintf(buf, sizeof buf, " split %.3f ms%s", cd.ms, mark);
			else if (cd.tiled) snprintf(buf, sizeof buf, " tiled %.3f ms%s", cd.ms, mark);
			else snprintf(buf, sizeof buf, " %dx%d%s %.3f ms%s", cd.L, cd.S, cd.exact ? "" : "r", cd.ms, mark);
			rep += buf;
		}
		c->tune_report = rep;
		if (c->p.model == NLC_DNLS) {
			c->dnls_variant = best->dnls;
		} else if (c->ragged) {
			NLC_TRY(select_ragged(c, best->ragged_S, best->half));
		} else {
			c->p.kernel = best->split ? NLC_KERNEL_SPLIT : best->tiled ? NLC_KERNEL_TILED : NLC_KERNEL_WAVE;
			c->p.lanes_per_chain = uint32_t(best->L);
			c->p.sites_per_lane = uint32_t(best->S);
			NLC_TRY(select_uniform(c));
		}
	}
	if (c->p.model == NLC_DDNKG && c->p.projection == NLC_PROJ_AUTO) NLC_TRY(tune_projection(c));
	NLC_TRY(tune_submission(c, trial_steps));
	if (report && len) snprintf(report, len, "%s", c->tune_report.c_str());
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
		         nlc::mixed_radix_plan(c->n, nullptr)       ? "dnls_lds_mixed_radix"
		         : nlc::bluestein_two_level_size(c->n)     ? "dnls_bluestein_two_level"
		         : nlc::dnls_two_level(c->n)               ? "dnls_two_level_fft"
		         : nlc::dnls_bluestein_size(c->n)          ? "dnls_lds_bluestein"
		         : !nlc::dnls_fast_path(c->n)              ? "dnls_lds_stockham"
		         : c->p.precision == NLC_FP32C        ? "dnls_radix16_registers_fp32c"
		         : c->dnls_variant == nlc::kDnlsRadix8 ? "dnls_radix8_registers"
		                                                : "dnls_radix16_registers");
	else if (c->ragged) {
		std::string models;
		for (const auto &ps : c->sets) models += std::string(models.empty() ? "" : "+") + names[ps.model];
		snprintf(buf, len,
		         "model=%s %s chains=%u classes=%zu sites=%llu kernel=%s sites_per_lane=%d%s items=%u",
		         models.c_str(), c->sets.size() > 1 ? "mixed" : "ragged", c->copies, c->cls.size(),
		         (unsigned long long) c->total_sites, c->persistent ? "ragged_persistent" : "per_class_wave",
		         c->ragged_S, c->persistent && c->ragged_half ? "/half" : "", c->nitems);
	}
	else if (c->use_split)
		snprintf(buf, len, "model=%s n=%u copies=%u kernel=split", names[c->p.model], c->n, c->copies);
	else if (c->use_tiled)
		snprintf(buf, len, "model=%s n=%u copies=%u kernel=tiled", names[c->p.model], c->n, c->copies);
	else
		snprintf(buf, len, "model=%s n=%u copies=%u kernel=%s lanes_per_chain=%d sites_per_lane=%d exact=%d",
		         names[c->p.model], c->n, c->copies, c->L > 64 ? "group" : "wave", c->L, c->S, int(c->exact));
	if (c->p.model != NLC_DNLS && !c->use_split) {
		// force form of the register-resident kernels
		std::string forms;
		for (size_t s = 0; s < std::max<size_t>(1, c->sets.size()); ++s) {
			const int f = c->pargs.a[s].form;
			forms += std::string(s ? "+" : "") +
			         (f == nlc::kFormCubic ? "cubic_unit" : f == nlc::kFormQuad ? "quadratic_unit" : "plain");
		}
		const size_t at = strlen(buf);
		if (at + 1 < len) snprintf(buf + at, len - at, " force=%s%s", forms.c_str(), is_c32(c) ? " precision=fp32c" : "");
	}
	if (c->p.model == NLC_DDNKG) {
		const size_t at = strlen(buf);
		if (at + 1 < len)
			snprintf(buf + at, len - at, " projection=%s",
			         c->proj_variant == nlc::kProjMfma       ? "mfma_f64"
			         : c->proj_variant == nlc::kProjMfmaWide ? "mfma_wide_f64"
			         : c->proj_variant == nlc::kProjBf16     ? "split_bf16"
			                                                 : "valu_f64");
	}
	return NLC_OK;
}

} // extern "C"
