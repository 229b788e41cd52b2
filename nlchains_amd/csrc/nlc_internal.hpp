// Internal declarations shared by the gfx950 kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <initializer_list>

#include "nlc_fft_large.hpp"

namespace nlc {

// Message returned by nlc_last_error(NULL) in the calling thread (set by the
// context-free entry points when they fail).
void set_thread_error(const char *msg);

enum Model { DNKG = 0, FPUT = 1, TODA = 2, DDNKG = 3, DNLS = 4 };

// Scaled Yoshida weights and model constants, passed by value as a kernel
// argument (they land in SGPRs).  dc[k] = c_k*dt, dd[k] = d_k*dt, computed on
// the host exactly as the reference does (DNKG_FPUT_Toda.cpp:76-77).
struct StepArgs {
	double dc[8];
	double dd[8];
	double m;     // DNKG linear parameter
	double m2;    // DNKG: m + 2
	double beta;
	double alpha; // FPUT cubic coefficient
	double a2;    // Toda: 2*alpha
	double a2i;   // Toda: 1/(2*alpha)
	double scale; // sigma of the scaled force forms (see force_form)
	double inv_scale; // 1 / sigma
	int form;     // kFormPlain, kFormCubic or kFormQuad
};

// Force forms of the register-resident step kernels.  The scaled forms hold
// sigma q and sigma p, which turns the leading nonlinear coefficient into 1:
//   kFormCubic  DNKG / dDNKG with beta > 0, FPUT with alpha = 0 < beta:
//               sigma = sqrt(beta);
//   kFormQuad   FPUT with beta = 0 != alpha: sigma = alpha;
//   kFormPlain  every other case; Toda keeps v = sigma q with unscaled p,
//               sigma = 256 alpha/ln2 (its bond exponent in 1/128 octaves;
//               kernels with 25+ sites per lane use sigma/128, octaves).
enum ForceForm { kFormPlain = 0, kFormCubic = 1, kFormQuad = 2 };

// The form and sigma for a model's parameters.
inline ForceForm force_form(int model, double alpha, double beta, double *sigma) {
	*sigma = model == TODA ? alpha * 0x1.71547652b82fep+8 : 1.0; // 256/ln2
	if ((model == DNKG || model == DDNKG || (model == FPUT && alpha == 0.0)) && beta > 0.0) {
		*sigma = __builtin_sqrt(beta);
		return kFormCubic;
	}
	if (model == FPUT && beta == 0.0 && alpha != 0.0) {
		*sigma = alpha;
		return kFormQuad;
	}
	return kFormPlain;
}

// Enqueues `steps` Yoshida steps of a wave-resident step kernel.
using StepLauncher = hipError_t (*)(hipStream_t, double2 *state, uint32_t copies, uint32_t n, uint32_t steps,
                                    const StepArgs &args, const double *mp2);

// Wave shapes: a chain occupies L lanes of a wavefront, S consecutive sites
// per lane (64/L chains per wavefront).  `exact` means n == L*S; otherwise the
// last L*S-n lanes hold S-1 sites (requires (S-1)*L < n).
StepLauncher lookup_step_dnkg(int L, int S, bool exact);
StepLauncher lookup_step_fput(int L, int S, bool exact);
StepLauncher lookup_step_toda(int L, int S, bool exact);
StepLauncher lookup_step_ddnkg(int L, int S, bool exact);
// FP32-compensated wave kernels (FPUT, Toda; p.precision = NLC_FP32C): exact
// power-of-two shapes n = L S, L <= 64, S <= 32 (StepLauncher's mp2 unused).
StepLauncher lookup_step32_fput(int L, int S);
StepLauncher lookup_step32_toda(int L, int S);
// Temporally blocked step for chains longer than a workgroup holds: `steps`
// steps in launches of a few steps each, ping-ponging between state and
// scratch (same size as state); the result ends in state.
using TiledLauncher = hipError_t (*)(hipStream_t, double2 *state, double2 *scratch, uint32_t copies, uint32_t n,
                                     uint32_t steps, const StepArgs &args, const double *mp2);
TiledLauncher lookup_tiled_dnkg();
TiledLauncher lookup_tiled_fput();
TiledLauncher lookup_tiled_toda();
TiledLauncher lookup_tiled_ddnkg();
// 16 lanes with S = 5..32 sites per lane for any n in (16(S-1), 16S] (ragged lanes).
StepLauncher lookup_step16_dnkg(int S);
StepLauncher lookup_step16_fput(int S);
StepLauncher lookup_step16_toda(int S);
StepLauncher lookup_step16_ddnkg(int S);

// Workgroup shapes for long chains: one chain per workgroup of W = 2..8
// wavefronts (L = 64 W lanes), S sites per lane, W = ceil(n / (64 S)) and
// 64 W (S - 1) < n <= 64 W S; exact instances for S = 8, 16, 32, others for
// S = 9..32.
StepLauncher lookup_group_dnkg(int S, bool exact);
StepLauncher lookup_group_fput(int S, bool exact);
StepLauncher lookup_group_toda(int S, bool exact);
StepLauncher lookup_group_ddnkg(int S, bool exact);

// Ragged and mixed ensembles: the persistent kernel (nlc_persist_s<S>.hip).
// A kind is a model with its force form, compiled as its own code path.
constexpr int kKindDnkgCubic = 0, kKindDnkgPlain = 1, kKindFputCubic = 2, kKindFputQuad = 3, kKindFputPlain = 4,
              kKindToda = 5;
constexpr uint32_t kAllKinds = 0x3F;
__host__ __device__ constexpr int kind_model(int kind) { return kind <= 1 ? DNKG : kind <= 4 ? FPUT : TODA; }
__host__ __device__ constexpr int kind_form(int kind) {
	return (kind == kKindDnkgCubic || kind == kKindFputCubic) ? kFormCubic : kind == kKindFputQuad ? kFormQuad
	                                                                                                : kFormPlain;
}
// The kind of a model with parameters giving force form `form`; -1 if none.
inline int kind_of(int model, int form) {
	if (model == DNKG) return form == kFormCubic ? kKindDnkgCubic : form == kFormPlain ? kKindDnkgPlain : -1;
	if (model == FPUT)
		return form == kFormCubic ? kKindFputCubic : form == kFormQuad ? kKindFputQuad : kKindFputPlain;
	if (model == TODA) return form == kFormPlain ? kKindToda : -1;
	return -1;
}
// Parameter sets of a persistent launch (at most 3: one per model of a mixed
// ensemble), each with its kind.
struct PersistentArgs {
	StepArgs a[3];
	int kind[3];
};
// An item: one wavefront's worth (nchains <= 64/L) of consecutive chains of n
// sites of one class, L = 2^log2_lanes lanes per chain; flags bits 0-1 the
// parameter set, bit 2: S/2 sites per lane instead of the kernel's S.
constexpr uint8_t kItemSetMask = 3, kItemHalfS = 4;
struct RaggedItem {
	uint64_t site0;
	uint32_t n;
	uint16_t nchains;
	uint8_t log2_lanes;
	uint8_t flags;
};
// queue: two zeroed counters owned by the caller (the kernel leaves them zero).
using PersistentLauncher = hipError_t (*)(hipStream_t, double2 *state, const RaggedItem *items, uint32_t nitems,
                                          uint32_t steps, const PersistentArgs &pa, uint32_t *queue,
                                          uint32_t max_blocks);
// S in {8, 16, 32}; kind_mask: one kind, or (S >= 16) any set of kinds
PersistentLauncher lookup_persistent_s8(uint32_t kind_mask);
PersistentLauncher lookup_persistent_s16(uint32_t kind_mask);
PersistentLauncher lookup_persistent_s32(uint32_t kind_mask);
// The same, FP32-compensated (FPUT and Toda kinds).
PersistentLauncher lookup_persistent32_s8(uint32_t kind_mask);
PersistentLauncher lookup_persistent32_s16(uint32_t kind_mask);
PersistentLauncher lookup_persistent32_s32(uint32_t kind_mask);

// HBM-streaming step over the split layout q[n][copies], p[n][copies]: one
// launch per substep from buffer pair cur to pair cur ^ 1 (cur is updated; the
// result is in q[cur], p[cur]).
hipError_t launch_step_split(hipStream_t, int model, double *q[2], double *p[2], int &cur, uint32_t copies,
                             uint32_t n, uint32_t steps, const StepArgs &args, const double *mp2);
// planar double2[copies][n] <-> split q[n][copies], p[n][copies]
hipError_t launch_planar_to_split(hipStream_t, const double2 *planar, double *q, double *p, uint32_t copies,
                                  uint32_t n);
hipError_t launch_split_to_planar(hipStream_t, const double *q, const double *p, double2 *planar, uint32_t copies,
                                  uint32_t n);

// Energy pass of the plane-wave models (DNKG/FPUT/Toda): per-block partial sums
// of e_k over copies into partial[grid][n], grid = energy_rows(n, copies).
// tw[j] = exp(-2 pi i j/n).
// Chains of 8192 < n <= 131072 sites (powers of two) go through a two-level
// FFT in chunks of copies; scratch holds energy_scratch_sites(n, copies)
// values (0: not needed; without it such chains use the O(n^2) direct DFT).
// Lengths that are not powers of two, 4096 < n <= 65536, use Bluestein's
// convolution through the same FFT when the Bluestein tables are given.
struct LongChainFft {
	double2 *scratch = nullptr;    // energy_scratch_sites(n, copies) values
	size_t scratch_sites = 0;
	const double2 *blue = nullptr; // bluestein_two_level_tables(n), or null
};
uint32_t energy_rows(uint32_t n, uint32_t copies);
size_t energy_scratch_sites(uint32_t n, uint32_t copies);
hipError_t launch_energy_real(hipStream_t, const double2 *state, uint32_t n, uint32_t copies, const double2 *tw,
                              const double *omega, double *partial, uint32_t grid, const LongChainFft &lf);
// DNLS energies: partial[g][k] = sum |psi_hat_k|^2 (omega applied in the reduction).
hipError_t launch_energy_dnls(hipStream_t, const double2 *state, uint32_t n, uint32_t copies, const double2 *tw,
                              double *partial, uint32_t grid, const LongChainFft &lf);
// e_raw[k] = scale_k * sum_g partial[g][k] (scale = omega for DNLS, else NULL = 1),
// fixed-order; scratch holds reduce_scratch_rows(grid) * n doubles.
uint32_t reduce_scratch_rows(uint32_t grid);
hipError_t launch_reduce_partial(hipStream_t, const double *partial, uint32_t grid, uint32_t n,
                                 const double *scale, double *e_raw, double *scratch);
// out[0..n) = normalised linenergies, out[n] = S_WT, out[n+1] = S_INF.
hipError_t launch_results(hipStream_t, const double *e_raw, uint32_t n, double norm, int a0is0,
                          const uint32_t *idx, uint32_t nidx, double *out);
// launch_reduce_partial followed by launch_results, fused into one workgroup
// for n <= 2048.
hipError_t launch_reduce_results(hipStream_t, const double *partial, uint32_t grid, uint32_t n, const double *scale,
                                 double *e_raw, double *scratch, double norm, int a0is0, const uint32_t *idx,
                                 uint32_t nidx, double *out);
// Total energy per copy (and DNLS norm).
hipError_t launch_chain_energy(hipStream_t, int model, const double2 *state, uint32_t n, uint32_t copies,
                               const StepArgs &args, const double *masses, double *energy, double *norm);

// Largest chain length the LDS FFT of the energy pass handles (power of two).
constexpr uint32_t kEnergyFftMax = 4096;

} // namespace nlc

namespace nlc {
// dDNKG projection energies: partial[rows][n] with rows = ddnkg_partial_rows(copies),
// on FP64 matrix cores (kProjMfma: 64 copies x 128 modes per workgroup;
// kProjMfmaWide: 64 x 256 with the state tile shared through LDS), the FP64
// vector ALU (kProjValu) or BF16 matrix cores with three-term operand splits
// (kProjBf16: n % 32 == 0; vplanes = ddnkg_split_host of V, 3 n^2 values; the
// state is split inside the kernel).
constexpr int kProjMfma = 0, kProjValu = 1, kProjBf16 = 2, kProjMfmaWide = 3;
uint32_t ddnkg_partial_rows(uint32_t copies);
bool ddnkg_bf16_supported(uint32_t n);
void ddnkg_split_host(const double *x, size_t count, uint16_t *planes);
hipError_t launch_energy_ddnkg(hipStream_t, const double2 *state, uint32_t n, uint32_t copies, const double *V,
                               const double *omega, double *partial, int variant, const uint16_t *vplanes);
// Host eigensystem (nlc_host.cpp).
int eigensystem(uint32_t n, const double *mp2, double *omegas, double *V);
extern const double kYoshidaC[8], kYoshidaD[8];
} // namespace nlc

namespace nlc {
// DNLS split step (nlc_dnls.hip).  theta9[k] = beta dt c_k (k <= 7),
// theta9[8] = beta dt 2 c_0; table: 7 Fourier multipliers of n entries, or for
// the register kernels (dnls_fast_path: n = 4096, radix 16 with 256 threads or
// radix 8 with 512 threads per chain) the 4 distinct ones permuted by
// dnls_permute_tables for that variant.  Chain lengths: powers of two up to
// kDnlsLdsMax; any other length whose Bluestein convolution size
// (dnls_bluestein_size, the power of two >= 2n - 1) is at most
// kDnlsBluesteinMax, with dnls_bluestein_tables' 2M + n entries after the 7
// multipliers.
constexpr int kDnlsRadix16 = 0, kDnlsRadix8 = 1;
constexpr uint32_t kDnlsLdsMax = 8192, kDnlsBluesteinMax = 8192;
// Powers of two above 8192 up to 131072 (dnls_two_level) use the two-level
// FFT in HBM with the multipliers permuted by dnls_two_level_tables and a
// scratch of copies x n values.
bool dnls_supported(uint32_t n);
uint32_t dnls_bluestein_size(uint32_t n);
// Stockham passes (nlc_fft_mixed.hpp) of a length n <= 8192 that is not a
// power of two and has no prime factor above 7: radix 4 first, then 2, 3, 5,
// 7 (at most 8 passes); false for every other n.
struct MixedPlan {
	uint8_t radix[15];
	uint8_t passes;
};
inline bool mixed_radix_plan(uint32_t n, MixedPlan *plan) {
	if (n < 2 || (n & (n - 1)) == 0 || n > 8192) return false;
	MixedPlan p{};
	uint32_t r = n;
	while (r % 4 == 0) p.radix[p.passes++] = 4, r /= 4;
	if (r % 2 == 0) p.radix[p.passes++] = 2, r /= 2;
	for (const uint32_t f : {3u, 5u, 7u})
		while (r % f == 0) p.radix[p.passes++] = uint8_t(f), r /= f;
	if (r != 1) return false;
	if (plan) *plan = p;
	return true;
}
void dnls_bluestein_tables(uint32_t n, double2 *out);
bool dnls_two_level(uint32_t n);
size_t dnls_two_level_table_size(uint32_t n);
size_t dnls_blue_row_table_size(uint32_t n);
void dnls_blue_row_tables(uint32_t n, double2 *out);
void dnls_two_level_tables(uint32_t n, const double2 *table, double2 *out);
// Row pass of the two-level FFT for rows of 4096 points on the radix-16
// register FFT (tk: the row-permuted multipliers of dnls_two_level_tables;
// twn: the n-point roots for the inverse column twiddle).
hipError_t launch_dnls_rows4096(hipStream_t, double2 *rows, uint32_t nrows, uint32_t R, const double2 *tw4096,
                                const double2 *twn, const double2 *tk);
bool dnls_fast_path(uint32_t n);
void dnls_permute_tables(uint32_t n, int variant, const double2 *table, double2 *out);
// FP32-compensated DNLS step of 4096-site chains (nlc_dnls_fp32c.hip): the
// FP32 twiddles tw[n] and the double-single increment multipliers
// mult[4][n] = {hi.x, hi.y, lo.x, lo.y} of (e^{-i dt d_k omega_j} - 1)/n in
// the radix-16 kernel's order; theta9 as for launch_dnls.
void dnls_fp32c_tables(uint32_t n, double dt, float2 *tw, float4 *mult);
hipError_t launch_dnls_fp32c(hipStream_t, double2 *state, uint32_t copies, uint32_t steps, const double *theta9,
                             const float2 *tw, const float4 *mult);
hipError_t launch_dnls(hipStream_t, double2 *state, uint32_t n, uint32_t copies, uint32_t steps,
                       const double *theta9, const double2 *tw, const double2 *table, int variant,
                       double2 *scratch, const double2 *blue);
// Lengths that are not powers of two, 4096 < n <= 65536: Bluestein's
// convolution of length L = bluestein_two_level_size(n) (a power of two
// >= 2n - 1) through the two-level FFT, with the tables of
// bluestein_two_level_tables (2L + n values) and a scratch of L values per
// chain (DNLS step: every copy; energy pass: chunks of copies).
uint32_t bluestein_two_level_size(uint32_t n);
void bluestein_two_level_tables(uint32_t n, double2 *out);
// bhat = FFT_M(b)/M of the chirp b_j = e^{i pi j^2/n} (|j| < n, cyclic), the
// M-point roots and the chirp itself (M, M and n values; bhat in the
// two-level [r][k] order when R > 1).
void bluestein_tables(uint32_t n, uint32_t M, uint32_t R, double2 *out);
// The DNLS launchers of the other chain lengths, called by launch_dnls:
// LDS-resident kernels (nlc_dnls_generic.hip) and the HBM paths
// (nlc_dnls_long.hip).
hipError_t launch_dnls_generic(hipStream_t, double2 *state, uint32_t n, uint32_t copies, uint32_t steps,
                               const double *theta9, const double2 *tw, const double2 *table);
hipError_t launch_dnls_long(hipStream_t, double2 *state, uint32_t n, uint32_t copies, uint32_t steps,
                            const double *theta9, const double2 *tw, const double2 *table, double2 *scratch,
                            const double2 *blue);
hipError_t launch_energy_bluestein(hipStream_t, bool dnls, const double2 *state, uint32_t n, uint32_t copies,
                                   const double2 *blue, const double *omega, double *partial, uint32_t grid,
                                   double2 *scratch, size_t scratch_sites);
} // namespace nlc
