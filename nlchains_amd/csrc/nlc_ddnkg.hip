// dDNKG linear-mode energies: projection of q and p on the eigenvectors of the
// mass-weighted periodic Laplacian, squared and summed over copies.
//
// Reference behaviour followed: dDNKG.cpp:118-135 (plane2split transpose, two
// cublasDgemm with V*diag(omega) and V, make_linenergies) and dDNKG.cu:237-265
// (warp per mode: sum over copies of proj_phi^2 + proj_pi^2); CPU form
// dDNKG/cpu-implementation/dDNKG.cpp:58-72.
//
// Design: one fused kernel reads the planar shard directly (no transpose, no
// projection matrices in HBM) and its epilogue squares, applies omega and
// reduces 64 copies into one partial row; the per-tile rows are reduced in a
// fixed order afterwards.  Forms of the product:
//  * k_proj_mfma: FP64 matrix cores, v_mfma_f64_16x16x4f64.  A wavefront owns
//    64 copies x 32 modes; the K index of the instruction runs over sites in a
//    lane-private order (lane group g = lane/16 takes sites 4g..4g+3 of each
//    16-site chunk, one per instruction), the same order for the state operand
//    and for V, so every lane loads 64 contiguous bytes of its copy's (q,p)
//    pairs and 32 contiguous bytes of its mode's eigenvector, with no LDS
//    staging and no transpose;
//  * k_proj_mfma_wide (default): the same products on 64-copy x 256-mode
//    workgroup tiles, the state tile shared by the wavefronts through LDS;
//  * k_proj_energy: the FP64 vector-ALU form (64x64 tile staged through LDS),
//    kept as the autotuner's alternative;
//  * k_proj_bf16x3_wide: BF16 matrix cores with three-term operand splits
//    (mode energies within about 1e-8 of FP64; never chosen automatically).
#include <cstring>

#include "nlc_internal.hpp"

namespace nlc {
namespace {

constexpr int kTc = 64, kTj = 64, kTi = 16;

__global__ void __launch_bounds__(256)
k_proj_energy(const double2 *__restrict__ state, uint32_t n, uint32_t copies, const double *__restrict__ V,
              const double *__restrict__ omega, double *__restrict__ partial) {
	__shared__ double2 As[kTi][kTc + 1];
	__shared__ double Bs[kTi][kTj + 1];
	__shared__ double red[16][kTj + 1];
	const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
	const uint32_t j0 = blockIdx.x * kTj, c0 = blockIdx.y * kTc;
	double aq[4][4], ap[4][4];
#pragma unroll
	for (int r = 0; r < 4; ++r)
#pragma unroll
		for (int s = 0; s < 4; ++s) aq[r][s] = ap[r][s] = 0.0;
	for (uint32_t i0 = 0; i0 < n; i0 += kTi) {
		for (int e = threadIdx.x; e < kTi * kTc; e += 256) {
			const int ii = e % kTi, cc = e / kTi;
			const uint32_t c = c0 + cc, i = i0 + ii;
			As[ii][cc] = (c < copies && i < n) ? state[size_t(c) * n + i] : make_double2(0.0, 0.0);
		}
		for (int e = threadIdx.x; e < kTi * kTj; e += 256) {
			const int ii = e % kTi, jj = e / kTi;
			const uint32_t j = j0 + jj, i = i0 + ii;
			Bs[ii][jj] = (j < n && i < n) ? V[size_t(j) * n + i] : 0.0;
		}
		__syncthreads();
#pragma unroll 4
		for (int ii = 0; ii < kTi; ++ii) {
			double2 a[4];
			double b[4];
#pragma unroll
			for (int r = 0; r < 4; ++r) a[r] = As[ii][ty * 4 + r];
#pragma unroll
			for (int s = 0; s < 4; ++s) b[s] = Bs[ii][tx * 4 + s];
#pragma unroll
			for (int r = 0; r < 4; ++r)
#pragma unroll
				for (int s = 0; s < 4; ++s) {
					aq[r][s] = fma(a[r].x, b[s], aq[r][s]);
					ap[r][s] = fma(a[r].y, b[s], ap[r][s]);
				}
		}
		__syncthreads();
	}
	// epilogue: e(c, j) = (omega_j q.V_j)^2 + (p.V_j)^2, summed over this tile's copies
#pragma unroll
	for (int s = 0; s < 4; ++s) {
		const uint32_t j = j0 + tx * 4 + s;
		const double w = j < n ? omega[j] : 0.0;
		double sum = 0.0;
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			const double wq = w * aq[r][s];
			sum += fma(wq, wq, ap[r][s] * ap[r][s]);
		}
		red[ty][tx * 4 + s] = sum;
	}
	__syncthreads();
	if (threadIdx.x < kTj) {
		double sum = 0.0;
		for (int t = 0; t < 16; ++t) sum += red[t][threadIdx.x];
		const uint32_t j = j0 + threadIdx.x;
		if (j < n) partial[size_t(blockIdx.y) * n + j] = sum;
	}
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kMfRows = 4;               // 16-copy row blocks per wavefront (64 copies)
constexpr int kMfCols = 2;               // 16-mode column blocks per wavefront (32 modes)
constexpr int kMfWaves = 4;              // wavefronts per workgroup, side by side over modes
constexpr int kMfModes = kMfWaves * 16 * kMfCols; // 128 modes per workgroup

// Workgroup b covers copies [64 y, 64 y + 64) and modes [128 x, 128 x + 128),
// wave w of the group modes 128 x + 32 w + [0, 32); output partial[y][j].
// XCD-aware numbering: workgroups go to the 8 XCDs round-robin (b mod 8), so
// b = 8 (mode_blocks u + x) + c with y = 8 u + c puts every mode block of copy
// block y on the same XCD at about the same time, and the copies' (q, p) rows
// come from that XCD's L2 after the first read instead of once per mode block
// from HBM.
// Operand layout of v_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16]
// in accumulator element r.
__global__ void __launch_bounds__(256, 2)
k_proj_mfma(const double2 *__restrict__ state, uint32_t n, uint32_t copies, const double *__restrict__ V,
            const double *__restrict__ omega, double *__restrict__ partial) {
	const uint32_t mode_blocks = (n + kMfModes - 1) / kMfModes;
	const uint32_t xcd = blockIdx.x & 7, rest = blockIdx.x >> 3;
	const uint32_t bx = rest % mode_blocks, by = (rest / mode_blocks) * 8u + xcd;
	if (by * 64u >= copies) return; // padding workgroup of the last group of 8 copy blocks
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int r16 = lane & 15, g = lane >> 4;
	const uint32_t c_base = by * 64u + r16;
	const uint32_t j_base = bx * uint32_t(kMfModes) + uint32_t(wave) * 32u + r16;

	f64x4 acc_q[kMfRows][kMfCols], acc_p[kMfRows][kMfCols];
#pragma unroll
	for (int rb = 0; rb < kMfRows; ++rb)
#pragma unroll
		for (int cb = 0; cb < kMfCols; ++cb) acc_q[rb][cb] = acc_p[rb][cb] = f64x4{0.0, 0.0, 0.0, 0.0};

	const bool full_rows = by * 64u + 64u <= copies;
	const bool full_cols = bx * uint32_t(kMfModes) + uint32_t(kMfModes) <= n;
	for (uint32_t k0 = 0; k0 < n; k0 += 16) {
		const uint32_t site = k0 + 4u * uint32_t(g);
		const bool full_k = k0 + 16u <= n;
		double2 a[kMfRows][4];
		double b[kMfCols][4];
#pragma unroll
		for (int rb = 0; rb < kMfRows; ++rb) {
			const uint32_t c = c_base + 16u * uint32_t(rb);
			const double2 *src = state + size_t(c) * n + site;
			if (full_k && full_rows) {
#pragma unroll
				for (int t = 0; t < 4; ++t) a[rb][t] = src[t];
			} else {
#pragma unroll
				for (int t = 0; t < 4; ++t)
					a[rb][t] = (c < copies && site + t < n) ? src[t] : make_double2(0.0, 0.0);
			}
		}
#pragma unroll
		for (int cb = 0; cb < kMfCols; ++cb) {
			const uint32_t j = j_base + 16u * uint32_t(cb);
			const double *src = V + size_t(j) * n + site;
			if (full_k && full_cols) {
#pragma unroll
				for (int t = 0; t < 4; ++t) b[cb][t] = src[t];
			} else {
#pragma unroll
				for (int t = 0; t < 4; ++t) b[cb][t] = (j < n && site + t < n) ? src[t] : 0.0;
			}
		}
#pragma unroll
		for (int t = 0; t < 4; ++t)
#pragma unroll
			for (int rb = 0; rb < kMfRows; ++rb)
#pragma unroll
				for (int cb = 0; cb < kMfCols; ++cb) {
					acc_q[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb][t].x, b[cb][t], acc_q[rb][cb], 0, 0, 0);
					acc_p[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb][t].y, b[cb][t], acc_p[rb][cb], 0, 0, 0);
				}
	}

	// epilogue: lane holds modes j_base + 16 cb and copies 16 rb + 4 g + r of the tile
#pragma unroll
	for (int cb = 0; cb < kMfCols; ++cb) {
		const uint32_t j = j_base + 16u * uint32_t(cb);
		const double w = j < n ? omega[j] : 0.0;
		double sum = 0.0;
#pragma unroll
		for (int rb = 0; rb < kMfRows; ++rb)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const double wq = w * acc_q[rb][cb][r], pp = acc_p[rb][cb][r];
				sum += fma(wq, wq, pp * pp);
			}
		// add the four lane groups (copy rows 4g..4g+3 of each block); a + b == b + a
		// exactly, so every lane of a group ends with the same value
		sum += __shfl_xor(sum, 16, 64);
		sum += __shfl_xor(sum, 32, 64);
		if (g == 0 && j < n) partial[size_t(by) * n + j] = sum;
	}
}

// ------------------------------------------------------- FP64 MFMA, wide tile
//
// k_proj_mfma_wide: a workgroup of 8 wavefronts owns 64 copies x 256 modes;
// wavefront w takes copy half w / 4 (32 copies, two 16-row blocks) and mode
// group w % 4 (64 modes, four 16-column blocks).  The state tile of each 16
// sites (64 copies x 16 sites of (q, p), 16 KB) is loaded once per workgroup
// into LDS (double-buffered; rows padded to 17 entries, so the 16 copy rows
// that one read instruction touches fall in distinct banks) and read from
// there by the four mode groups, instead of being loaded by each wavefront;
// each wavefront loads its eigenvector rows itself, one step ahead.  Global
// loads per product fall to half of k_proj_mfma's, and with 256 modes per
// workgroup a copy block is read by n/256 workgroups (two at 512 sites, on
// the same XCD at about the same time).  Same K order and operand layout.
constexpr int kWRows = 2, kWCols = 4;   // per wavefront: 32 copies x 64 modes
constexpr int kWModes = 256;            // modes per workgroup
constexpr int kWPad = 17;               // LDS row stride of the state tile, in (q, p) pairs

__device__ __forceinline__ void lds_fence_barrier() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool FULL>
__device__ __forceinline__ void wide_loop(const double2 *__restrict__ state, uint32_t n, uint32_t copies,
                                          const double *__restrict__ V, uint32_t c0, uint32_t j_base, int half,
                                          int r16, int g, double2 (&As)[2][64 * kWPad],
                                          f64x4 (&acc_q)[kWRows][kWCols], f64x4 (&acc_p)[kWRows][kWCols]) {
	const int tid = threadIdx.x;
	// staging: thread tid carries copy tid / 8, sites 2 (tid % 8) and + 1 of each 16-site step
	const uint32_t sc = uint32_t(tid >> 3), ss = 2u * uint32_t(tid & 7);
	const double2 *srow = state + size_t(c0 + sc) * n + ss;
	const bool s_ok = FULL || c0 + sc < copies;
	auto stage = [&](uint32_t k0, double2 &x0, double2 &x1) {
		if (FULL) {
			x0 = srow[k0], x1 = srow[k0 + 1];
		} else {
			x0 = (s_ok && k0 + ss < n) ? srow[k0] : make_double2(0.0, 0.0);
			x1 = (s_ok && k0 + ss + 1 < n) ? srow[k0 + 1] : make_double2(0.0, 0.0);
		}
	};
	auto load_b = [&](uint32_t k0, double (&b)[kWCols][4]) {
#pragma unroll
		for (int cb = 0; cb < kWCols; ++cb) {
			const uint32_t j = j_base + 16u * uint32_t(cb);
			const double *src = V + size_t(j) * n + k0 + 4u * uint32_t(g);
#pragma unroll
			for (int t = 0; t < 4; ++t)
				b[cb][t] = (FULL || (j < n && k0 + 4u * uint32_t(g) + t < n)) ? src[t] : 0.0;
		}
	};
	double2 x0, x1;
	stage(0, x0, x1);
	As[0][sc * kWPad + ss] = x0;
	As[0][sc * kWPad + ss + 1] = x1;
	double b[kWCols][4];
	load_b(0, b);
	// wait for the first operands here: otherwise the wait for them, placed at
	// the top of the loop, also holds every later step until that step's own
	// prefetch arrives (the loop body's operand registers are written by
	// loads in the first step and by copies after)
#pragma unroll
	for (int cb = 0; cb < kWCols; ++cb)
#pragma unroll
		for (int t = 0; t < 4; ++t) asm volatile("" : "+v"(b[cb][t]));
	lds_fence_barrier();
	const double2 *arow = &As[0][(32u * uint32_t(half) + uint32_t(r16)) * kWPad + 4u * uint32_t(g)];
	uint32_t buf = 0;
	for (uint32_t k0 = 0; k0 < n; k0 += 16) {
		const bool more = k0 + 16u < n;
		double bn[kWCols][4];
		if (more) {
			stage(k0 + 16u, x0, x1);
			load_b(k0 + 16u, bn);
		}
		const double2 *ar = arow + buf * (64 * kWPad);
#pragma unroll
		for (int t = 0; t < 4; ++t) {
			double2 a[kWRows];
#pragma unroll
			for (int rb = 0; rb < kWRows; ++rb) a[rb] = ar[16 * kWPad * rb + t];
#pragma unroll
			for (int rb = 0; rb < kWRows; ++rb)
#pragma unroll
				for (int cb = 0; cb < kWCols; ++cb) {
					acc_q[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb].x, b[cb][t], acc_q[rb][cb], 0, 0, 0);
					acc_p[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb].y, b[cb][t], acc_p[rb][cb], 0, 0, 0);
				}
		}
		if (more) {
			// the other buffer was last read before the previous barrier
			As[buf ^ 1u][sc * kWPad + ss] = x0;
			As[buf ^ 1u][sc * kWPad + ss + 1] = x1;
#pragma unroll
			for (int cb = 0; cb < kWCols; ++cb)
#pragma unroll
				for (int t = 0; t < 4; ++t) b[cb][t] = bn[cb][t];
		}
		lds_fence_barrier();
		buf ^= 1u;
	}
}

__global__ void __launch_bounds__(512, 1)
k_proj_mfma_wide(const double2 *__restrict__ state, uint32_t n, uint32_t copies, const double *__restrict__ V,
                 const double *__restrict__ omega, double *__restrict__ partial) {
	__shared__ double2 As[2][64 * kWPad];
	__shared__ double red[2][kWModes];
	const uint32_t mode_blocks = (n + kWModes - 1) / kWModes;
	const uint32_t xcd = blockIdx.x & 7, rest = blockIdx.x >> 3;
	const uint32_t bx = rest % mode_blocks, by = (rest / mode_blocks) * 8u + xcd;
	if (by * 64u >= copies) return; // padding workgroup of the last group of 8 copy blocks (uniform)
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int half = wave >> 2, mg = wave & 3, r16 = lane & 15, g = lane >> 4;
	const uint32_t c0 = by * 64u;
	const uint32_t j_base = bx * uint32_t(kWModes) + uint32_t(mg) * 64u + uint32_t(r16);

	f64x4 acc_q[kWRows][kWCols], acc_p[kWRows][kWCols];
#pragma unroll
	for (int rb = 0; rb < kWRows; ++rb)
#pragma unroll
		for (int cb = 0; cb < kWCols; ++cb) acc_q[rb][cb] = acc_p[rb][cb] = f64x4{0.0, 0.0, 0.0, 0.0};
	const bool full = c0 + 64u <= copies && bx * uint32_t(kWModes) + uint32_t(kWModes) <= n && n % 16u == 0;
	if (full) wide_loop<true>(state, n, copies, V, c0, j_base, half, r16, g, As, acc_q, acc_p);
	else wide_loop<false>(state, n, copies, V, c0, j_base, half, r16, g, As, acc_q, acc_p);

	// epilogue: lane holds modes j_base + 16 cb of copies 32 half + 16 rb + 4 g + r
#pragma unroll
	for (int cb = 0; cb < kWCols; ++cb) {
		const uint32_t j = j_base + 16u * uint32_t(cb);
		const double w = j < n ? omega[j] : 0.0;
		double sum = 0.0;
#pragma unroll
		for (int rb = 0; rb < kWRows; ++rb)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const double wq = w * acc_q[rb][cb][r], pp = acc_p[rb][cb][r];
				sum += fma(wq, wq, pp * pp);
			}
		sum += __shfl_xor(sum, 16, 64);
		sum += __shfl_xor(sum, 32, 64);
		if (g == 0) red[half][mg * 64 + 16 * cb + r16] = sum;
	}
	__syncthreads();
	if (threadIdx.x < kWModes) {
		const uint32_t j = bx * uint32_t(kWModes) + threadIdx.x;
		if (j < n) partial[size_t(by) * n + j] = red[0][threadIdx.x] + red[1][threadIdx.x];
	}
}

// ------------------------------------------------------- split BF16 form
//
// Each FP64 operand x is split into three BF16 terms, x = h + m + l (each the
// nearest BF16 of the remainder, through FP32), which carries x to about
// 2^-24 relative.  The product keeps the six significant cross terms,
// hh + (hm + mh + hl + lh + mm), on v_mfma_f32_16x16x32_bf16 with the leading
// term and the corrections in separate FP32 accumulators (the corrections are
// 2^-8 smaller, so their rounding is negligible), summed in FP64 at the end.
// The eigenvectors are split once on the host into planes [h m l][mode][site];
// the state is split inside the kernel.  Requires n % 32 == 0.  Operand layout
// of the instruction: lane l supplies A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16], e = 0..7, and holds D[4 (l / 16) + r][l % 16] in
// accumulator element r.
//
// k_proj_bf16x3_wide: a 64-copy x 256-mode workgroup tile, eight wavefronts of
// 32 copies x 64 modes, with both operands staged through LDS one 32-site step
// ahead.  The state is read in FP64 and split in registers while the matrix
// cores run the step before, so there is no split pass and no split planes in
// HBM; the eigenvector splits are read once per workgroup.  LDS per buffer:
// state splits [qh qm ql ph pm pl][64 copies][32 sites] (24 KB) and
// eigenvector splits [h m l][256 modes][32 sites] (48 KB); two buffers and the
// copy reduction take 148 KB, one workgroup per CU.  Each fragment read is 64
// lanes x 16 bytes (row r16 at a 64-byte stride, lane group g at 16 bytes).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split3(double x, __bf16 &h, __bf16 &m, __bf16 &l) {
	h = __bf16(float(x));
	const double r1 = x - double(float(h));
	m = __bf16(float(r1));
	l = __bf16(float(r1 - double(float(m))));
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

constexpr int kBModes = 256;                                                 // modes per workgroup
constexpr uint32_t kBfA = 6u * 64u * 64u, kBfB = 3u * uint32_t(kBModes) * 64u; // bytes per buffer
constexpr uint32_t kBfLds = 2u * (kBfA + kBfB) + 2u * uint32_t(kBModes) * uint32_t(sizeof(double));

// 16-byte piece c of LDS row r sits at piece c ^ swz(r): a fragment read puts
// 16 rows at a 64-byte stride in one access, which would otherwise hit each
// bank four times
__device__ __forceinline__ uint32_t swz(uint32_t row) { return (row >> 2) & 3u; }

template <bool FULL>
__device__ __forceinline__ void bf16_wide_loop(const double2 *__restrict__ state, uint32_t n, uint32_t copies,
                                               const __bf16 *__restrict__ vplanes, uint32_t c0, uint32_t jw0,
                                               int half, int mg, int r16, int g, unsigned char *lds,
                                               f32x4 (&hq)[2][4], f32x4 (&cq)[2][4], f32x4 (&hp)[2][4],
                                               f32x4 (&cp)[2][4]) {
	const int tid = threadIdx.x;
	const size_t vsz = size_t(n) * n;
	// staging: state copy tid / 8, sites 4 (tid % 8) + [0, 4); eigenvector row
	// tid / 2, sites 16 (tid % 2) + [0, 16), two 16-byte pieces per plane
	const uint32_t sc = uint32_t(tid >> 3), ss = 4u * uint32_t(tid & 7);
	const uint32_t vr = uint32_t(tid >> 1), vs = 16u * uint32_t(tid & 1);
	const bool s_ok = FULL || c0 + sc < copies;
	const bool v_ok = FULL || jw0 + vr < n;
	const double2 *srow = state + size_t(s_ok ? c0 + sc : 0u) * n + ss;
	const __bf16 *vrow = vplanes + size_t(v_ok ? jw0 + vr : 0u) * n + vs;
	auto load = [&](uint32_t k0, double2 (&x)[4], uint4 (&v)[3][2]) {
#pragma unroll
		for (int t = 0; t < 4; ++t) x[t] = s_ok ? srow[k0 + t] : make_double2(0.0, 0.0);
#pragma unroll
		for (int pl = 0; pl < 3; ++pl)
#pragma unroll
			for (int u = 0; u < 2; ++u)
				v[pl][u] = v_ok ? *(const uint4 *) (vrow + pl * vsz + k0 + 8u * uint32_t(u)) : make_uint4(0, 0, 0, 0);
	};
	auto store = [&](unsigned char *buf, const double2 (&x)[4], const uint4 (&v)[3][2]) {
		bf16x4 h[2], m[2], l[2];
#pragma unroll
		for (int t = 0; t < 4; ++t) {
			__bf16 a, b, c;
			split3(x[t].x, a, b, c);
			h[0][t] = a, m[0][t] = b, l[0][t] = c;
			split3(x[t].y, a, b, c);
			h[1][t] = a, m[1][t] = b, l[1][t] = c;
		}
		unsigned char *srow_l = buf + sc * 64u + 16u * ((ss >> 3) ^ swz(sc)) + 2u * (ss & 7u);
#pragma unroll
		for (int part = 0; part < 2; ++part) {
			*(bf16x4 *) (srow_l + (3 * part + 0) * 4096) = h[part];
			*(bf16x4 *) (srow_l + (3 * part + 1) * 4096) = m[part];
			*(bf16x4 *) (srow_l + (3 * part + 2) * 4096) = l[part];
		}
		unsigned char *vrow_l = buf + kBfA + vr * 64u;
#pragma unroll
		for (int pl = 0; pl < 3; ++pl)
#pragma unroll
			for (int u = 0; u < 2; ++u)
				*(uint4 *) (vrow_l + pl * (kBModes * 64) + 16u * (((vs >> 3) + uint32_t(u)) ^ swz(vr))) = v[pl][u];
	};
	// rows 16 apart have the same swizzle, so one offset serves every rb and cb
	const uint32_t a_off = (32u * uint32_t(half) + uint32_t(r16)) * 64u + 16u * (uint32_t(g) ^ swz(uint32_t(r16)));
	const uint32_t b_off = kBfA + (64u * uint32_t(mg) + uint32_t(r16)) * 64u + 16u * (uint32_t(g) ^ swz(uint32_t(r16)));
	auto compute = [&](const unsigned char *buf) {
		bf16x8 a[2][6];
#pragma unroll
		for (int rb = 0; rb < 2; ++rb)
#pragma unroll
			for (int pl = 0; pl < 6; ++pl) a[rb][pl] = *(const bf16x8 *) (buf + a_off + rb * 1024 + pl * 4096);
#pragma unroll
		for (int cb = 0; cb < 4; ++cb) {
			bf16x8 v[3];
#pragma unroll
			for (int pl = 0; pl < 3; ++pl) v[pl] = *(const bf16x8 *) (buf + b_off + cb * 1024 + pl * (kBModes * 64));
#pragma unroll
			for (int rb = 0; rb < 2; ++rb) {
				const bf16x8 *q = a[rb], *p = a[rb] + 3;
				hq[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q[0], v[0], hq[rb][cb], 0, 0, 0);
				hp[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[0], v[0], hp[rb][cb], 0, 0, 0);
				f32x4 x = cq[rb][cb], y = cp[rb][cb];
				x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q[0], v[1], x, 0, 0, 0);
				y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[0], v[1], y, 0, 0, 0);
				x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q[1], v[0], x, 0, 0, 0);
				y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[1], v[0], y, 0, 0, 0);
				x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q[0], v[2], x, 0, 0, 0);
				y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[0], v[2], y, 0, 0, 0);
				x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q[2], v[0], x, 0, 0, 0);
				y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[2], v[0], y, 0, 0, 0);
				x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q[1], v[1], x, 0, 0, 0);
				y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[1], v[1], y, 0, 0, 0);
				cq[rb][cb] = x, cp[rb][cb] = y;
			}
		}
	};
	double2 x[4];
	uint4 v[3][2];
	load(0, x, v);
	store(lds, x, v);
	lds_fence_barrier();
	uint32_t buf = 0;
	for (uint32_t k0 = 0; k0 < n; k0 += 32) {
		const bool more = k0 + 32u < n;
		if (more) load(k0 + 32u, x, v);
		compute(lds + buf * (kBfA + kBfB));
		// the other buffer was last read before the previous barrier
		if (more) store(lds + (buf ^ 1u) * (kBfA + kBfB), x, v);
		lds_fence_barrier();
		buf ^= 1u;
	}
}

__global__ void __launch_bounds__(512, 1)
k_proj_bf16x3_wide(const double2 *__restrict__ state, uint32_t n, uint32_t copies,
                   const __bf16 *__restrict__ vplanes, const double *__restrict__ omega, double *__restrict__ partial) {
	extern __shared__ __align__(16) unsigned char lds_bf[]; // kBfLds bytes
	double(*red)[kBModes] = (double(*)[kBModes])(lds_bf + 2u * (kBfA + kBfB));
	const uint32_t mode_blocks = (n + kBModes - 1) / kBModes;
	const uint32_t xcd = blockIdx.x & 7, rest = blockIdx.x >> 3;
	const uint32_t bx = rest % mode_blocks, by = (rest / mode_blocks) * 8u + xcd;
	if (by * 64u >= copies) return; // padding workgroup of the last group of 8 copy blocks (uniform)
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int half = wave >> 2, mg = wave & 3, r16 = lane & 15, g = lane >> 4;
	const uint32_t c0 = by * 64u, jw0 = bx * uint32_t(kBModes);

	f32x4 hq[2][4], cq[2][4], hp[2][4], cp[2][4];
#pragma unroll
	for (int rb = 0; rb < 2; ++rb)
#pragma unroll
		for (int cb = 0; cb < 4; ++cb)
			hq[rb][cb] = cq[rb][cb] = hp[rb][cb] = cp[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
	const bool full = c0 + 64u <= copies && jw0 + uint32_t(kBModes) <= n;
	if (full) bf16_wide_loop<true>(state, n, copies, vplanes, c0, jw0, half, mg, r16, g, lds_bf, hq, cq, hp, cp);
	else bf16_wide_loop<false>(state, n, copies, vplanes, c0, jw0, half, mg, r16, g, lds_bf, hq, cq, hp, cp);

	// epilogue: lane holds modes jw0 + 64 mg + 16 cb + r16 of copies
	// c0 + 32 half + 16 rb + 4 g + r
#pragma unroll
	for (int cb = 0; cb < 4; ++cb) {
		const uint32_t j = jw0 + 64u * uint32_t(mg) + 16u * uint32_t(cb) + uint32_t(r16);
		const double w = j < n ? omega[j] : 0.0;
		double sum = 0.0;
#pragma unroll
		for (int rb = 0; rb < 2; ++rb)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const double wq = w * (double(hq[rb][cb][r]) + double(cq[rb][cb][r]));
				const double pp = double(hp[rb][cb][r]) + double(cp[rb][cb][r]);
				sum += fma(wq, wq, pp * pp);
			}
		sum += __shfl_xor(sum, 16, 64);
		sum += __shfl_xor(sum, 32, 64);
		if (g == 0) red[half][64 * mg + 16 * cb + r16] = sum;
	}
	__syncthreads();
	if (threadIdx.x < kBModes) {
		const uint32_t j = jw0 + threadIdx.x;
		if (j < n) partial[size_t(by) * n + j] = red[0][threadIdx.x] + red[1][threadIdx.x];
	}
}

} // namespace

uint32_t ddnkg_partial_rows(uint32_t copies) { return (copies + kTc - 1) / kTc; }

bool ddnkg_bf16_supported(uint32_t n) { return n % 32 == 0; }

void ddnkg_split_host(const double *x, size_t count, uint16_t *planes) {
	// the device split3 on the host: nearest BF16 of the FP32 nearest, remainders in FP64
	auto bf16 = [](float f) -> uint16_t {
		uint32_t u;
		std::memcpy(&u, &f, 4);
		u += 0x7FFFu + ((u >> 16) & 1u);
		return uint16_t(u >> 16);
	};
	auto back = [](uint16_t h) -> double {
		const uint32_t u = uint32_t(h) << 16;
		float f;
		std::memcpy(&f, &u, 4);
		return double(f);
	};
	for (size_t i = 0; i < count; ++i) {
		const uint16_t h = bf16(float(x[i]));
		const double r1 = x[i] - back(h);
		const uint16_t m = bf16(float(r1));
		const uint16_t l = bf16(float(r1 - back(m)));
		planes[i] = h, planes[count + i] = m, planes[2 * count + i] = l;
	}
}

hipError_t launch_energy_ddnkg(hipStream_t st, const double2 *state, uint32_t n, uint32_t copies, const double *V,
                               const double *omega, double *partial, int variant, const uint16_t *vplanes) {
	if (variant == kProjValu) {
		const dim3 grid((n + kTj - 1) / kTj, ddnkg_partial_rows(copies));
		hipLaunchKernelGGL(k_proj_energy, grid, dim3(256), 0, st, state, n, copies, V, omega, partial);
	} else if (variant == kProjBf16) {
		if (!ddnkg_bf16_supported(n) || !vplanes) return hipErrorInvalidValue;
		hipError_t e = hipFuncSetAttribute((const void *) k_proj_bf16x3_wide,
		                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(kBfLds));
		if (e != hipSuccess) return e;
		const uint32_t mode_blocks = (n + kBModes - 1) / kBModes;
		const uint32_t groups = (ddnkg_partial_rows(copies) + 7) / 8; // of 8 copy blocks, one per XCD
		hipLaunchKernelGGL(k_proj_bf16x3_wide, dim3(groups * 8 * mode_blocks), dim3(512), kBfLds, st, state, n, copies,
		                   (const __bf16 *) vplanes, omega, partial);
	} else if (variant == kProjMfmaWide) {
		const uint32_t mode_blocks = (n + kWModes - 1) / kWModes;
		const uint32_t groups = (ddnkg_partial_rows(copies) + 7) / 8; // of 8 copy blocks, one per XCD
		hipLaunchKernelGGL(k_proj_mfma_wide, dim3(groups * 8 * mode_blocks), dim3(512), 0, st, state, n, copies, V,
		                   omega, partial);
	} else {
		const uint32_t mode_blocks = (n + kMfModes - 1) / kMfModes;
		const uint32_t groups = (ddnkg_partial_rows(copies) + 7) / 8; // of 8 copy blocks, one per XCD
		hipLaunchKernelGGL(k_proj_mfma, dim3(groups * 8 * mode_blocks), dim3(256), 0, st, state, n, copies, V, omega,
		                   partial);
	}
	return hipGetLastError();
}

} // namespace nlc
