// nlchains: command-line driver of the MI355X hot path.
//
//   nlchains <model> <common options> <model-specific options>
//
// Keeps the reference's subprogram registry (common/main.cpp:33-36,142-176),
// common options and validation (common/main.cpp:53-134), model options
// (DNKG_FPUT_Toda.cpp:26-46, dDNKG.cpp:27-47, DNLS.cpp:30-41), batch cadence
// and termination (DNKG_FPUT_Toda.cpp:108-151, common/loop_control.hpp:16-37)
// and dump formats (README.md:74-82, common/results.cpp:23-61,
// dDNKG/eigensystem.cpp:23-29).  One process, one GPU, no MPI: every device
// operation goes through the C ABI of include/nlc.h.
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <future>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <system_error>
#include <type_traits>
#include <vector>

#include "nlc.h"
#include "options.hpp"

using nlc_cli::Options;

namespace {

volatile std::sig_atomic_t quit_requested = 0;

std::string process_ident() {
	char host[256] = "localhost";
	gethostname(host, sizeof host - 1);
	return std::string("nlchains/hostname: ") + host;
}

struct io_error : std::runtime_error {
	using std::runtime_error::runtime_error;
};

struct help_quit {
	std::string text;
};

// Common options (common/main.cpp:53-70) and the parsed configuration.
struct Config {
	bool verbose = false, autotune = false;
	std::string initial, entropymask, prefix;
	uint32_t n = 0, copies = 0, batch = 0, dump_interval = 0;
	double dt = 0;
	uint64_t steps = 0, time_offset = 0;
	enum { NONE, INFORMATION, WT } limit_type = NONE;
	double entropy_limit = 0;
	std::vector<uint8_t> mask;
	std::vector<double> shard; // double[copies][n][2] or complex<double>[copies][n]
};

void add_common(Options &o) {
	o.add("verbose", 'v', false, false, "print extra informations")
	    .add("initial", 'i', true, true, "initial state file")
	    .add("chain_length", 'n', true, true, "length of the chain")
	    .add("copies", 'c', true, true, "number of realizations of the chain in the ensemble")
	    .add("dt", 0, true, true, "time step value")
	    .add("kernel_batching", 'b', true, true, "number of steps per kernel invocation")
	    .add("steps", 's', true, false, "number of steps in total (0 for infinite)", "0")
	    .add("time_offset", 'o', true, false, "time offset in dump files", "0")
	    .add("entropy", 'e', true, false, "entropy limit")
	    .add("WTlimit", 0, false, false, "limit WT entropy instead of information entropy")
	    .add("entropymask", 0, true, false, "mask of linenergies to include in entropy calculation")
	    .add("prefix", 'p', true, true, "prefix for dump files")
	    .add("dump_interval", 0, true, false,
	         "number of steps between full state dumps (defaults to same value as --batch, does not affect time "
	         "granularity of entropy)")
	    .add("kernel_shape", 0, true, false, "wave shape LANESxSITES of the step kernel (default: automatic)")
	    .add("autotune", 0, false, false, "time the step kernel forms on the initial state and keep the fastest");
}

// Validation and input reading: common/main.cpp:72-134 (single process, so
// the shard is the whole ensemble).
void run_common(Options &o, int argc, char **argv, Config &c) {
	if (argc == 1) throw help_quit{o.usage()};
	try {
		o.parse(argc, argv);
		c.verbose = o.count("verbose");
		c.autotune = o.count("autotune");
		c.n = o.get<uint32_t>("chain_length");
		c.copies = o.get<uint32_t>("copies");
		c.dt = o.get<double>("dt");
		c.batch = o.get<uint32_t>("kernel_batching");
		c.steps = o.get<uint64_t>("steps");
		c.time_offset = o.get<uint64_t>("time_offset");
		c.prefix = o.get<std::string>("prefix");
		c.initial = o.get<std::string>("initial");
		if (c.n < 2 || !c.copies || !(c.dt > 0) || !c.batch)
			throw std::invalid_argument(
			    "--chain_length must be >= 2, --copies, --dt and --kernel_batching must be positive numbers");
		if (o.count("entropy")) {
			c.entropy_limit = o.get<double>("entropy");
			c.limit_type = o.count("WTlimit") ? Config::WT : Config::INFORMATION;
		}
		if (!o.count("dump_interval")) c.dump_interval = c.batch;
		else if (!(c.dump_interval = o.get<uint32_t>("dump_interval")))
			throw std::invalid_argument("--dump_interval must be > 0");
		if (c.prefix.empty()) throw std::invalid_argument("--prefix must not be an empty string");
		if (c.dump_interval % c.batch) throw std::invalid_argument("--dump_interval must be a multiple of --kernel_batching");
		if (c.steps % c.batch) throw std::invalid_argument("--steps must be a multiple of --kernel_batching");
		if (c.time_offset && c.steps && c.time_offset >= c.steps)
			throw std::invalid_argument("--time_offset must be less than --steps");
		if (uint64_t(c.n) * c.copies > (1ULL << 31))
			throw std::invalid_argument("(--chain_length * --copies) must be <= 2^31");
	} catch (const std::invalid_argument &e) {
		throw std::invalid_argument(std::string(e.what()) + "\n" + o.usage());
	}
	const size_t doubles = size_t(c.n) * c.copies * 2;
	c.shard.resize(doubles);
	{
		std::ifstream in(c.initial, std::ios::binary);
		if (!in || !in.read(reinterpret_cast<char *>(c.shard.data()), std::streamsize(doubles * sizeof(double))))
			throw io_error("could not read initial state (" + c.initial + ")");
	}
	if (o.count("entropymask")) {
		const std::string f = o.get<std::string>("entropymask");
		std::ifstream in(f, std::ios::binary);
		c.mask.resize(c.n);
		if (!in || !in.read(reinterpret_cast<char *>(c.mask.data()), c.n))
			throw io_error("could not read entropy mask file (" + f + ")");
		size_t cnt = 0;
		for (uint8_t b : c.mask) cnt += b != 0;
		if (cnt < 2) throw std::invalid_argument("Too few modes for entropy calculation");
	}
}

// RAII wrapper of an nlc_ctx; every failure becomes an exception.
struct Context {
	nlc_ctx *ctx = nullptr;
	explicit Context(const nlc_params &p) {
		const int rc = nlc_create(&ctx, &p);
		if (rc) throw std::runtime_error(std::string("nlc_create: ") + nlc_strerror(rc) + ": " + nlc_last_error(nullptr));
	}
	~Context() { nlc_destroy(ctx); }
	void check(int rc, const char *what) const {
		if (rc) throw std::runtime_error(std::string(what) + ": " + nlc_strerror(rc) + ": " + nlc_last_error(ctx));
	}
};

void write_file(const std::string &name, const void *data, size_t bytes, const char *what) {
	std::ofstream out(name, std::ios::binary | std::ios::trunc);
	if (!out || !out.write(static_cast<const char *>(data), std::streamsize(bytes)))
		throw io_error(std::string("cannot open ") + what + " file (" + name + ")");
}

// Results files and checks: common/results.cpp:8-73.
class Results {
public:
	Results(const Config &c) : c_(c), entropy_(c.prefix + "-entropy", std::ios::binary | std::ios::trunc) {
		if (!entropy_) throw io_error("cannot open entropy file (" + c.prefix + "-entropy)");
	}
	// {t dt, S_WT, S_INF} appended and flushed every batch
	void write_entropy(uint64_t t, const double s[2]) {
		const double row[3] = {double(t) * c_.dt, s[0], s[1]};
		if (!entropy_.write(reinterpret_cast<const char *>(row), sizeof row).flush())
			throw io_error("cannot write entropy file");
		if (c_.verbose) std::cout << "Entropy step " << t << ": " << s[0] << '/' << s[1] << std::endl;
	}
	void write_linenergies(uint64_t t, const std::vector<double> &e) const {
		write_file(c_.prefix + "-linenergies-" + std::to_string(t), e.data(), e.size() * sizeof(double), "linenergies");
	}
	// the shard is written by a background thread while the GPU continues
	void write_shard(uint64_t t, std::vector<double> shard) {
		wait();
		pending_ = std::async(std::launch::async, [name = c_.prefix + "-" + std::to_string(t), s = std::move(shard)] {
			write_file(name, s.data(), s.size() * sizeof(double), "dump");
		});
	}
	void wait() {
		if (pending_.valid()) pending_.get();
	}
	// non-finite or below-limit entropy requests termination
	void check_entropy(const double s[2]) const {
		if (!std::isfinite(s[0]) || !std::isfinite(s[1])) {
			quit_requested = 1;
			return;
		}
		if (c_.limit_type == Config::NONE) return;
		const double checked = c_.limit_type == Config::WT ? s[0] : s[1];
		if (checked < c_.entropy_limit) quit_requested = 1;
	}

private:
	const Config &c_;
	std::ofstream entropy_;
	std::future<void> pending_;
};

// The batch loop shared by every model (DNKG_FPUT_Toda.cpp:108-151,
// dDNKG.cpp:115-163, DNLS.cpp:139-188, loop_control.hpp:16-37).
int run_loop(Config &c, nlc_params p, bool split_kernel, const std::string &shape) {
	p.chain_length = c.n;
	p.copies = c.copies;
	p.dt = c.dt;
	if (split_kernel) p.kernel = NLC_KERNEL_SPLIT;
	if (!shape.empty()) {
		unsigned L = 0, S = 0;
		if (std::sscanf(shape.c_str(), "%ux%u", &L, &S) != 2)
			throw std::invalid_argument("--kernel_shape must be LANESxSITES, e.g. 64x16");
		p.lanes_per_chain = L;
		p.sites_per_lane = S;
	}
	Context ctx(p);
	if (c.verbose) {
		char desc[256];
		nlc_describe(ctx.ctx, desc, sizeof desc);
		std::cerr << process_ident() << ": " << desc << std::endl;
	}
	ctx.check(nlc_upload(ctx.ctx, c.shard.data()), "upload");
	// a0is0 is the reference's per-model default, set by nlc_create
	if (!c.mask.empty()) {
		const int a0is0 = p.model == NLC_DNLS || (p.model != NLC_DDNKG && p.m == 0.0);
		ctx.check(nlc_set_entropy_mask(ctx.ctx, c.mask.data(), a0is0), "entropy mask");
	}
	Results res(c);
	std::vector<double> e(c.n);
	double s[2];
	uint64_t t = c.time_offset;
	auto dump = [&](uint64_t tt) {
		res.write_linenergies(tt, e);
		std::vector<double> shard(c.shard.size());
		ctx.check(nlc_download(ctx.ctx, shard.data()), "download");
		res.write_shard(tt, std::move(shard));
	};
	if (c.autotune) {
		char report[1024];
		ctx.check(nlc_autotune(ctx.ctx, std::max<uint32_t>(1, c.batch / 10), report, sizeof report), "autotune");
		if (c.verbose) std::cerr << process_ident() << ": autotune: " << report << std::endl;
	}
	bool synched = false;
	// energies at the start; afterwards every batch is one replayed hipGraph
	// (move, energy pass, entropies), up to two queued ahead of the results
	// being written (loop_control.hpp:46-57); tq is the time the state reaches
	// after the queued batches.  A batch that starts at a dump time is queued
	// only once that dump has been taken.
	constexpr uint32_t kAhead = 2;
	uint64_t tq = t;
	uint32_t queued = 0;
	ctx.check(nlc_linenergies(ctx.ctx, 0.0, e.data(), s), "linenergies");
	while (true) {
		res.write_entropy(t, s);
		res.check_entropy(s);
		if (t % c.dump_interval == 0) dump(t);
		// loop_control::break_now: on a quit request the end step becomes the
		// step the queued work reaches, as in the reference
		if (quit_requested && !synched) c.steps = tq, synched = true;
		if (c.steps == t && (synched || c.steps)) break;
		while (queued < kAhead && !(c.steps && tq >= c.steps) && (tq == t || tq % c.dump_interval != 0)) {
			ctx.check(nlc_batch_submit(ctx.ctx, c.batch, 0.0), "batch");
			tq += c.batch;
			++queued;
		}
		ctx.check(nlc_batch_collect(ctx.ctx, e.data(), s), "batch");
		--queued;
		t += c.batch;
	}
	if (t % c.dump_interval != 0) dump(t);
	res.wait();
	return 0;
}

using Program = std::function<int(int, char **)>;

std::map<std::string, Program> &programs() {
	static std::map<std::string, Program> progs;
	return progs;
}

// DNKG / FPUT / Toda: DNKG_FPUT_Toda.cpp:18-46
int main_real(int model, int argc, char **argv) {
	Options o(std::string("Options for ") + argv[0]);
	add_common(o);
	switch (model) {
	case NLC_DNKG:
		o.add("", 'm', true, true, "linear parameter m").add("beta", 0, true, true, "fourth order nonlinearity");
		break;
	case NLC_FPUT:
		o.add("alpha", 0, true, true, "third order nonlinearity").add("beta", 0, true, true, "fourth order nonlinearity");
		break;
	case NLC_TODA:
		o.add("alpha", 0, true, true, "steepness of exponential, V(dx)=e^(2*alpha*dx)/(4*alpha^2)-dx/(2*alpha)");
		break;
	}
	o.add("split_kernel", 0, false, false, "force use of split kernel");
	if (model != NLC_DNKG)
		o.add("fp32c", 0, false, false,
		      "FP32 fields with compensated accumulation, FP64 reductions (power-of-two chain_length <= 2048)");
	Config c;
	run_common(o, argc, argv, c);
	nlc_params p;
	nlc_params_init(&p);
	p.model = model;
	if (model != NLC_DNKG && o.count("fp32c")) p.precision = NLC_FP32C;
	if (model == NLC_DNKG) p.m = o.get<double>("-m");
	if (model != NLC_TODA) p.beta = o.get<double>("beta");
	if (model != NLC_DNKG) p.alpha = o.get<double>("alpha");
	if (model == NLC_TODA && p.alpha == 0) throw std::invalid_argument("--alpha must not be zero for Toda");
	return run_loop(c, p, o.count("split_kernel"), o.has("kernel_shape") ? o.get<std::string>("kernel_shape") : "");
}

// dDNKG: dDNKG.cpp:27-47 (masses file), eigensystem dumps eigensystem.cpp:23-29
int main_ddnkg(int argc, char **argv) {
	Options o(std::string("Options for ") + argv[0]);
	add_common(o);
	o.add("", 'm', true, true, "linear parameter m filename")
	    .add("beta", 0, true, true, "fourth order nonlinearity")
	    .add("split_kernel", 0, false, false, "force use of split kernel");
	Config c;
	run_common(o, argc, argv, c);
	const std::string mfile = o.get<std::string>("-m");
	std::vector<double> masses(c.n);
	{
		std::ifstream in(mfile, std::ios::binary);
		if (!in || !in.read(reinterpret_cast<char *>(masses.data()), std::streamsize(c.n * sizeof(double))))
			throw io_error("could not read m file (" + mfile + ")");
	}
	std::vector<double> mp2(c.n), omegas(c.n), V(size_t(c.n) * c.n);
	for (uint32_t i = 0; i < c.n; ++i) mp2[i] = masses[i] + 2;
	if (nlc_eigensystem(c.n, mp2.data(), omegas.data(), V.data())) throw std::runtime_error("Cannot calculate eigensystem!");
	write_file(c.prefix + "-omegas", omegas.data(), omegas.size() * sizeof(double), "omegas");
	write_file(c.prefix + "-eigenvectors", V.data(), V.size() * sizeof(double), "eigenvectors");
	nlc_params p;
	nlc_params_init(&p);
	p.model = NLC_DDNKG;
	p.beta = o.get<double>("beta");
	p.masses = masses.data();
	return run_loop(c, p, o.count("split_kernel"), o.has("kernel_shape") ? o.get<std::string>("kernel_shape") : "");
}

// DNLS: DNLS.cpp:30-41.  The cuFFT-callback switches are accepted for
// compatibility; the library has no callbacks (the FFTs are fused on chip).
int main_dnls(int argc, char **argv) {
	Options o(std::string("Options for ") + argv[0]);
	add_common(o);
	o.add("no_linear_callback", 0, false, false, "do not use cuFFT callback for linear evolution (no effect)")
	    .add("no_nonlinear_callback", 0, false, false, "do not use cuFFT callback for nonlinear evolution (no effect)")
	    .add("beta", 0, true, true, "fourth order nonlinearity")
	    .add("fp32c", 0, false, false,
	         "FP32 fields with compensated accumulation, FP64 reductions (chain_length 4096)");
	Config c;
	run_common(o, argc, argv, c);
	nlc_params p;
	nlc_params_init(&p);
	p.model = NLC_DNLS;
	p.beta = o.get<double>("beta");
	if (o.count("fp32c")) p.precision = NLC_FP32C;
	return run_loop(c, p, false, "");
}

int fatal(const std::string &msg) {
	std::cerr << process_ident() << ": " << msg << std::endl;
	return 1;
}

} // namespace

int main(int argc, char **argv) try {
	for (int sig : {SIGINT, SIGTERM}) std::signal(sig, [](int) { quit_requested = 1; });
	auto &progs = programs();
	progs["DNKG"] = [](int ac, char **av) { return main_real(NLC_DNKG, ac, av); };
	progs["FPUT"] = [](int ac, char **av) { return main_real(NLC_FPUT, ac, av); };
	progs["Toda"] = [](int ac, char **av) { return main_real(NLC_TODA, ac, av); };
	progs["dDNKG"] = main_ddnkg;
	progs["DNLS"] = main_dnls;
	std::string names;
	for (auto &kv : progs) names += (names.empty() ? "" : ", ") + kv.first;
	if (argc == 1) {
		std::cout << "Please specify as first argument the subprogram to run: " << names << std::endl;
		return 0;
	}
	auto it = progs.find(argv[1]);
	if (it == progs.end()) throw std::invalid_argument("please specific a valid program among " + names);
	return it->second(argc - 1, argv + 1);
} catch (const help_quit &h) {
	std::cout << h.text;
	return 0;
} catch (const io_error &e) {
	return fatal(std::string("I/O error, ") + e.what());
} catch (const std::invalid_argument &e) {
	return fatal(std::string("invalid argument, ") + e.what());
} catch (const std::exception &e) {
	return fatal(std::string("exception, ") + e.what());
} catch (...) {
	return fatal("unspecified fatal exception");
}
