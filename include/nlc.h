/*
 * nlc.h -- C ABI of libnlchains_amd.so, the MI355X (gfx950) hot path of nlchains.
 *
 * The reference exposes its hot path as C++ device entry points per model
 * (reference paths, relative to the reference tree):
 *   DNKG_FPUT_Toda/DNKG_FPUT_Toda.hpp:19-23  move<Model>(plane2split*&, cudaStream_t),
 *                                            make_linenergies(fft_phis, fft_pis, omegas, stream)
 *   dDNKG/dDNKG.hpp:12-14                    move(plane2split&, bool& use_split, mp2, stream),
 *                                            make_linenergies(proj_phi, proj_pi, stream)
 *   DNLS/DNLS.hpp:16-20                      evolve_nonlinear / evolve_linear / make_linenergies
 *   common/results.hpp:26-31                 results::calc_linenergies(norm).calc_entropies()
 * with state in the globals gres.shard_gpu / gres.linenergies_gpu
 * (common/configuration.hpp:67-78) and parameters in __constant__ symbols.
 * This ABI replaces that seam with an opaque context: plain pointers and sizes,
 * no exceptions (a negative return code plus nlc_last_error()).
 *
 * State layout is the reference shard/dump layout (README.md:74-78):
 *   real models: double[copies][chain_length][2] = {q, p}
 *   DNLS:        complex<double>[copies][chain_length]  (= double[..][..][2] {re, im})
 *
 * Threading: a context is used by one host thread at a time.  Device work is
 * enqueued on the context's own HIP stream; calls marked "asynchronous" return
 * before the work completes, all others synchronise the stream before returning.
 */
#ifndef NLC_H
#define NLC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLC_ABI_VERSION 1

/* Return codes.  The reference throws instead (common/main.cpp:162-176 maps
 * exceptions to exit code 1); here every failure is a negative code. */
#define NLC_OK 0
#define NLC_EARG (-1)        /* invalid argument: std::invalid_argument, common/main.cpp:83-103 */
#define NLC_EHIP (-2)        /* HIP runtime failure: cuda_error from assertcu, common/utilities_cuda.cuh:19-37 */
#define NLC_ENOMEM (-3)      /* device or pinned host allocation failed */
#define NLC_EUNSUPPORTED (-4)/* valid request the library has no kernel for */
#define NLC_ESTATE (-5)      /* call not valid in the context's current state */

/* Subprograms of the reference registry (common/main.cpp:33-36, DNKG_FPUT_Toda.cpp:156-162,
 * dDNKG.cpp:168-170, DNLS.cpp:193-195). */
typedef enum { NLC_DNKG = 0, NLC_FPUT = 1, NLC_TODA = 2, NLC_DDNKG = 3, NLC_DNLS = 4 } nlc_model;

/* Arithmetic of the step.  NLC_FP64 is the reference's all-double method
 * (DNLS.cpp:159-175 for DNLS, DNKG_FPUT_Toda.cu:33-66 for the FPUT and Toda
 * kicks).  NLC_FP32C keeps the state on chip as two FP32 values (hi + lo),
 * forms every substep's increment in FP32 and adds it with a compensated
 * sum; the state in HBM, the energies, norms and entropies stay FP64:
 *   DNLS with 4096-site chains: nonlinear phase, FFTs and Fourier multiplier
 *   in FP32;
 *   FPUT and Toda with power-of-two chain lengths up to 2048 (uniform, ragged,
 *   and mixed FPUT + Toda ensembles, every set FP32C): FPUT bond forces in
 *   FP32, the Toda bond exponential in FP64, q and p updated by Kahan's
 *   compensated summation.
 * Other models and lengths: NLC_EUNSUPPORTED.  On gfx950 without packed FP32
 * an FP32 instruction issues at the FP64 rate, so FP32C is faster only where
 * the FP64 step is not bound by its instructions (DNLS). */
typedef enum { NLC_FP64 = 0, NLC_FP32C = 1 } nlc_precision;

/* Step-kernel family.  NLC_KERNEL_SPLIT is the reference's --split_kernel
 * (DNKG_FPUT_Toda.cpp:43,91-94; dDNKG.cpp:34): HBM-streaming over the
 * ensemble-major (split) layout.  NLC_KERNEL_WAVE keeps each chain in the
 * registers of one or more wavefront lanes (the reference's move_chain_in_warp /
 * move_chain_in_thread), or of a workgroup of 2-8 wavefronts for chains of
 * 2049-16384 sites.  NLC_KERNEL_TILED keeps overlapping segments of 2048
 * sites (1024 for Toda) in the registers of one wavefront for a few steps per
 * launch (temporal blocking, any chain length from 256 sites; it needs a
 * second buffer of the state's size).  NLC_KERNEL_AUTO picks WAVE when a shape
 * exists for the chain length (every length up to 16384), else TILED (the
 * reference's move<> dispatch, DNKG_FPUT_Toda.cu:254-293, which falls back to
 * its split kernel there); WAVE without a shape override does the same. */
typedef enum { NLC_KERNEL_AUTO = 0, NLC_KERNEL_WAVE = 1, NLC_KERNEL_SPLIT = 2, NLC_KERNEL_TILED = 3 } nlc_kernel;

/* dDNKG energy projection (the two cublasDgemm of dDNKG.cpp:118-135 and the
 * reduction of dDNKG.cu:237-265, fused).  NLC_PROJ_AUTO = FP64 matrix cores
 * with 64 x 256 tiles until nlc_autotune times the three FP64 forms (matrix
 * cores with 64 x 128 or 64 x 256 tiles, vector ALU).  NLC_PROJ_SPLIT_BF16 (chain
 * lengths divisible by 32, chosen explicitly, never by the autotuner) runs the
 * products on BF16 matrix cores with three-term splits of both operands and
 * FP32 accumulation: mode energies differ from the FP64 forms at the 1e-7
 * level of the largest mode energy. */
typedef enum {
	NLC_PROJ_AUTO = 0,
	NLC_PROJ_MFMA_F64 = 1,
	NLC_PROJ_VALU_F64 = 2,
	NLC_PROJ_SPLIT_BF16 = 3,
	NLC_PROJ_MFMA_F64_WIDE = 4
} nlc_projection;

/* Model and ensemble parameters: the reference's common options (gconf,
 * common/configuration.hpp:20-39) and model options (DNKG_FPUT_Toda.cpp:26-43,
 * dDNKG.cpp:31-34, DNLS.cpp:34-37). */
typedef struct nlc_params {
	uint32_t struct_size;     /* sizeof(nlc_params), set by nlc_params_init */
	int32_t model;            /* nlc_model */
	int32_t precision;        /* nlc_precision */
	int32_t kernel;           /* nlc_kernel */
	uint32_t chain_length;    /* -n, >= 2; DNLS: at most 65536, or a power of two up to 131072
	                             (NLC_EUNSUPPORTED otherwise) */
	uint32_t copies;          /* -c, >= 1; chain_length * copies <= 2^31 (32 GiB of state);
	                             DNLS: copies < 2^24 */
	double dt;                /* --dt, > 0 */
	double m;                 /* -m (DNKG linear parameter) */
	double alpha;             /* --alpha (FPUT cubic term; Toda steepness, != 0) */
	double beta;              /* --beta (DNKG, FPUT, dDNKG, DNLS) */
	const double *masses;     /* dDNKG: m_x, chain_length doubles (content of the -m file); copied */
	uint32_t lanes_per_chain; /* wave shape override (0 = default); used by the autotuner and tests.
	                             1..64 (power of two): lanes of one wavefront; 128..512 (multiple
	                             of 64): one workgroup of lanes_per_chain/64 wavefronts per chain */
	uint32_t sites_per_lane;  /* wave shape override (0 = default); DNLS with chain_length 4096:
	                             16 or 8 complex values per thread (radix-16 or radix-8 kernel) */
	int32_t projection;       /* nlc_projection (dDNKG only) */
} nlc_params;

typedef struct nlc_ctx nlc_ctx;

/* ---- library ---------------------------------------------------------- */

int nlc_abi_version(void);
const char *nlc_strerror(int code);
/* Message of the last failure on ctx; with ctx == NULL, of the last failed
 * nlc_create (or nlc_create_ragged, nlc_create_mixed, nlc_linenergies_noalloc)
 * in the calling thread. */
const char *nlc_last_error(const nlc_ctx *ctx);

/* Defaults: FP64, AUTO kernel, everything else zero. */
int nlc_params_init(nlc_params *p);

/* ---- context life cycle ------------------------------------------------ */

/* Validates p (the rules of common/main.cpp:83-103 that concern the hot path),
 * selects the step kernel, allocates the device shard and work buffers on the
 * current HIP device, creates the stream.  Replaces cuda_ctx.activate
 * (common/utilities_cuda.cpp:13-28) plus the per-driver setup
 * (DNKG_FPUT_Toda.cpp:47-95, dDNKG.cpp:36-70 incl. the eigensystem of
 * dDNKG/eigensystem.cpp:13-36, DNLS.cpp:55-126). */
int nlc_create(nlc_ctx **out, const nlc_params *p);
void nlc_destroy(nlc_ctx *ctx);

/* ---- state ------------------------------------------------------------- */

/* Host -> device copy of the whole shard (the initial-state read of
 * common/main.cpp:112-118 plus the H2D of utilities_cuda.cpp:22). */
int nlc_upload(nlc_ctx *ctx, const void *state);
/* Device -> host copy of the shard (the dumper of DNKG_FPUT_Toda.cpp:98-104). */
int nlc_download(nlc_ctx *ctx, void *state);
/* Device address and size of the shard (reference layout), for callers that
 * move it with their own device copies. */
int nlc_device_state(nlc_ctx *ctx, void **dptr, size_t *bytes);

/* ---- hot path ---------------------------------------------------------- */

/* Asynchronous.  Advances every chain by `steps` Yoshida-6 steps: the
 * reference's move<Model> (DNKG_FPUT_Toda.cu:254-293), dDNKG::move
 * (dDNKG.cu:199-234), and the DNLS split-step batch (DNLS.cpp:159-175), each
 * with kernel_batching = steps.  Drift/kick merging as the reference: the
 * c[7] drift of a step is fused with the c[0] drift of the next, and one final
 * c[7] drift closes the batch. */
int nlc_move(nlc_ctx *ctx, uint32_t steps);

/* Entropy mode set: --entropymask semantics (common/main.cpp:119-128,
 * common/results.cpp:8-15).  mask: chain_length bytes, non-zero = mode included,
 * or NULL for "all modes" (minus mode 0 when a0is0).  a0is0: the reference's
 * results(a0is0) flag; nlc_create sets the reference default per model
 * (m == 0 for DNKG/FPUT/Toda, DNKG_FPUT_Toda.cpp:52; true for DNLS, DNLS.cpp:49;
 * false for dDNKG, dDNKG.cpp:83).  Returns NLC_EARG for fewer than 2 modes. */
int nlc_set_entropy_mask(nlc_ctx *ctx, const uint8_t *mask, int a0is0);

/* Synchronous.  Linear-mode energy pass plus results::calc_linenergies(norm)
 * and calc_entropies (common/results_entropy.cpp:11-46), all FP64 on the GPU:
 *   linenergies[k] = norm * sum_copies e_k   (e_0 := 0 when a0is0)
 *   entropies[0] = S_WT, entropies[1] = S_INF
 * norm <= 0 selects the reference's normalisation for the model:
 *   0.5/(copies*n) (DNKG_FPUT_Toda.cpp:129), 0.5/copies (dDNKG.cpp:141),
 *   1/(n*copies) (DNLS.cpp:150).
 * Either output pointer may be NULL.  Replaces the per-batch energy chain
 * DNKG_FPUT_Toda.cpp:113-131, dDNKG.cpp:118-143, DNLS.cpp:142-152. */
int nlc_linenergies(nlc_ctx *ctx, double norm, double *linenergies, double *entropies);

/* Synchronous.  Raw per-mode sums over copies, before normalisation (the
 * content of gres.linenergies after make_linenergies, e.g.
 * DNKG_FPUT_Toda.cu:296-323). */
int nlc_linenergies_raw(nlc_ctx *ctx, double *e_raw);

/* Synchronous.  Total Hamiltonian of every copy, FP64 on the GPU, with the
 * energy functions of supporting-material/post-processing/nlchains.m:40-48.
 * For DNLS also norms[c] = sum |psi|^2 (norms may be NULL; ignored otherwise). */
int nlc_chain_energies(nlc_ctx *ctx, double *energies, double *norms);

/* ---- batched form ------------------------------------------------------ */

/* Ragged ensemble: nchains chains of individual lengths chain_lengths[i]
 * (DNKG, FPUT, Toda; FP32C for FPUT and Toda with power-of-two lengths up to
 * 2048).  The reference runs one chain length per process
 * (gconf.chain_length, common/configuration.hpp:24); here chains of all lengths
 * share one context and one persistent launch per batch (powers of two from 4
 * to 2048 sites; other lengths run one wave-shaped launch per class).  State layout for
 * upload/download: the chains concatenated in the caller's order, each in the
 * reference layout.  p->chain_length and p->copies are ignored.  Energies and
 * entropies are computed per length class: nlc_linenergies writes, for every
 * class in ascending length, that class's linear energies (sum of the class
 * lengths doubles in total) and its two entropies (2 doubles per class). */
int nlc_create_ragged(nlc_ctx **out, const nlc_params *p, const uint32_t *chain_lengths, uint32_t nchains);
/* Mixed ensemble: chains of several models and lengths in one context, one
 * persistent launch per batch.  params[0..nsets) (1 <= nsets <= 3): a model
 * (DNKG, FPUT or Toda) with its parameters, all sets FP64 or all FP32C (FPUT
 * and Toda only, powers of two up to 2048 sites); chain i has length
 * chain_lengths[i] and follows params[chain_set[i]].  The reference runs one
 * model per process (common/main.cpp:157-160, a program per model in the
 * registry of DNKG_FPUT_Toda.cpp:156-162); here the models share the state, the
 * queue of the persistent kernel and the batch graph.  Length classes are the
 * distinct (set, length) pairs in ascending order; energies, entropies and the
 * state layout are as for nlc_create_ragged, each class with its own model's
 * dispersion, normalisation and a0is0 default (m == 0). */
int nlc_create_mixed(nlc_ctx **out, const nlc_params *params, uint32_t nsets, const uint32_t *chain_lengths,
                     const uint32_t *chain_set, uint32_t nchains);
/* Length classes of a context (ascending (set, length)); a uniform context has one. */
int nlc_ragged_classes(const nlc_ctx *ctx, uint32_t *nclasses, uint32_t *lengths, uint32_t *counts, uint32_t cap);
/* The parameter set of each class (0 unless mixed), in the order of nlc_ragged_classes. */
int nlc_ragged_class_sets(const nlc_ctx *ctx, uint32_t *sets, uint32_t cap);

/* Synchronous.  One dump interval: nlc_move(steps) followed by the energy pass,
 * normalisation and entropies, replayed from a hipGraph captured on the first
 * call (and recaptured when steps or norm change).  Outputs as nlc_linenergies.
 * The batched counterpart of one iteration of DNKG_FPUT_Toda.cpp:108-143. */
int nlc_batch(nlc_ctx *ctx, uint32_t steps, double norm, double *linenergies, double *entropies);

/* Pipelined batches.  nlc_batch_submit queues one batch as nlc_batch does and
 * returns without waiting: its results are copied to one of NLC_BATCH_SLOTS
 * pinned slots.  nlc_batch_collect waits for the oldest queued batch and
 * returns its outputs (as nlc_batch's), so results come back in submission
 * order.  At most NLC_BATCH_SLOTS batches may be queued (NLC_EARG beyond,
 * or for a collect with none queued).  Every other call is ordered after the
 * queued batches on the context's stream: nlc_download, nlc_linenergies and
 * nlc_chain_energies see the state after all of them.
 * Replaces the reference's batch throttle (at most two batches in flight) and
 * its result callbacks: common/loop_control.hpp:40-57, DNKG_FPUT_Toda.cpp:127-143. */
#define NLC_BATCH_SLOTS 4
int nlc_batch_submit(nlc_ctx *ctx, uint32_t steps, double norm);
int nlc_batch_collect(nlc_ctx *ctx, double *linenergies, double *entropies);

/* How nlc_batch_submit issues a batch: 0 = replay of the batch's cached
 * hipGraph (the default), 1 = direct launches of the same kernels (the same
 * results).  A graph replay costs a few microseconds more than the launches it
 * holds, which matters for batches of tens of microseconds (64 x 1024 chains);
 * nlc_autotune times both and sets the faster.  No counterpart in the
 * reference (its batches are direct launches and cuFFT calls).  NLC_ESTATE
 * while batches are queued. */
int nlc_set_batch_mode(nlc_ctx *ctx, int direct);

/* Times every applicable step-kernel shape (wave and workgroup shapes lanes x
 * sites and the split kernel for uniform ensembles; sites per lane of the persistent kernel
 * for ragged ones) on the context's own state for trial_steps steps, keeps the
 * fastest, restores the state, and writes a one-line report.  Replaces the
 * reference's compile-time optimized_chain_length choice (CMakeLists.txt:74-78)
 * and its launch heuristic (common/utilities_cuda.cpp:46-72). */
int nlc_autotune(nlc_ctx *ctx, uint32_t trial_steps, char *report, size_t len);

/* ---- allocation-free form ------------------------------------------------ */

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The energy pass, normalisation and entropies of nlc_linenergies
 * (all modes, minus mode 0 when a0is0) for a planar state already on the
 * current device: d_state = double[copies][chain_length][2] (DNKG, FPUT, Toda:
 * {q, p}; DNLS: psi), d_out = chain_length + 2 doubles of device memory
 * (linear energies, S_WT, S_INF), m the DNKG linear parameter, norm <= 0 the
 * model's default normalisation.  Asynchronous on `stream` (a hipStream_t, NULL
 * for the null stream).  It allocates nothing: its work buffers live in one
 * static device array of NLC_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; calls that overlap on one GPU race.  Chain lengths: powers of
 * two from 8 to NLC_NOALLOC_MAX_CHAIN (NLC_EARG beyond, with the reason in
 * nlc_last_error(NULL)). */
#define NLC_NOALLOC_SCRATCH_BYTES (8u << 20)
#define NLC_NOALLOC_MAX_CHAIN 4096
int nlc_linenergies_noalloc(int model, const void *d_state, uint32_t chain_length, uint32_t copies, double m,
                            double norm, int a0is0, double *d_out, void *stream);

/* Waits for all work enqueued on the context. */
int nlc_sync(nlc_ctx *ctx);
/* The context's hipStream_t (for events and graph capture by the caller). */
void *nlc_stream(nlc_ctx *ctx);
/* Human-readable description of the selected kernels and launch shapes, as
 * space-separated key=value fields: model, n, copies, kernel, the shape
 * (lanes_per_chain, sites_per_lane, exact), for the register-resident real
 * kernels force=cubic_unit|quadratic_unit|plain (the state is held scaled so
 * that the leading nonlinear coefficient is 1: sqrt(beta) for DNKG / dDNKG
 * with beta > 0 and FPUT with alpha = 0, alpha for FPUT with beta = 0; the
 * results are the reference's up to rounding), and projection= for dDNKG. */
int nlc_describe(const nlc_ctx *ctx, char *buf, size_t len);

/* ---- host helpers (no device needed) ------------------------------------ */

/* Yoshida-6 weights c[8], d[8]: common/symplectic.hpp:5-8. */
int nlc_yoshida(double *c, double *d);
/* Linear dispersion omega[n]: DNKG_FPUT_Toda/dispersion.cpp:9-16 for
 * DNKG/FPUT/Toda (FPUT/Toda use m = 0), DNLS/dispersion.cpp:10-17 for DNLS. */
int nlc_dispersion(int model, uint32_t n, double m, double *omega);
/* Eigensystem of diag(mp2) - periodic adjacency (dDNKG/eigensystem.cpp:13-36):
 * omegas ascending = sqrt(eigenvalues); eigenvectors column-major n x n.
 * n <= 512, or no HIP device: Householder tridiagonalisation + implicit QR on
 * the host; larger n with a device: rocSOLVER dsyevd (loaded on first use). */
int nlc_eigensystem(uint32_t n, const double *mp2, double *omegas, double *eigenvectors);

/* Seeded initial conditions (SplitMix64 streams, one per copy).  The
 * reference has no generator (it only reads -i files); these define the
 * synthetic workloads.  All write the reference layout. */
/* Running waves of the harmonic chain with omega_k^2 = m + 4 sin^2(pi k/n):
 * modes k in [k_lo, k_hi) get energy e_density*n/(k_hi-k_lo) each, with
 * uniform random phases; modes with omega_k == 0 are skipped. */
int nlc_ic_plane_waves(uint32_t n, uint32_t copies, double m, uint32_t k_lo, uint32_t k_hi, double e_density,
                       uint64_t seed, double *state);
/* DNLS: psi_x = sum_{|k|<kmax} a e^{i(2 pi k x/n + phi_k)}, |a|^2 = norm_density/(2kmax-1). */
int nlc_ic_dnls(uint32_t n, uint32_t copies, uint32_t kmax, double norm_density, uint64_t seed, double *state);
/* dDNKG: standing eigenmodes 0..modes-1 of (omegas, eigenvectors), energy
 * e_density*n/modes each, uniform random phases. */
int nlc_ic_eigenmodes(uint32_t n, uint32_t copies, const double *omegas, const double *eigenvectors, uint32_t modes,
                      double e_density, uint64_t seed, double *state);
/* Ragged chain lengths: each drawn uniformly from choices[0..nchoices). */
int nlc_ic_ragged_lengths(uint32_t nchains, const uint32_t *choices, uint32_t nchoices, uint64_t seed,
                          uint32_t *lengths);
/* Ragged running waves (m = 0 dispersion for FPUT/Toda, m > 0 for DNKG): chain
 * i of length n_i gets energy density ~ U[e_lo, e_hi) spread equally over modes
 * 1..max(1, floor(mode_fraction n_i)).  Chains concatenated in order. */
int nlc_ic_ragged_plane_waves(const uint32_t *lengths, uint32_t nchains, double m, double mode_fraction, double e_lo,
                              double e_hi, uint64_t seed, double *state);
/* Per-site masses m_x ~ U[lo, hi) (dDNKG masses file content). */
int nlc_ic_masses(uint32_t n, double lo, double hi, uint64_t seed, double *masses);

#ifdef __cplusplus
}
#endif
#endif /* NLC_H */
