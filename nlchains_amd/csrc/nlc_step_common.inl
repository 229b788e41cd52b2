// Shared pieces of the register-resident step kernels (wave and workgroup
// forms): the Toda exp table, DPP / ds_bpermute lane moves and the FPUT / Toda
// bond terms.  Included inside namespace nlc { namespace { ... } } by
// nlc_step_wave.inl and nlc_step_group.inl, once per translation unit.
__constant__ double kExp2Tab[kExpTabSize] = NLC_EXP2_TABLE;

// Toda kernels: copy the exp table into LDS (every thread of the workgroup
// reaches the barrier).
template <int MODEL>
__device__ __forceinline__ void load_exp_table(double *lds) {
	if constexpr (MODEL == TODA) {
		for (uint32_t i = threadIdx.x; i < uint32_t(kExpTabSize); i += blockDim.x) lds[i] = kExp2Tab[i];
		__syncthreads();
	}
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
	const int lo = __double2loint(v), hi = __double2hiint(v);
	// every control used here reads a valid lane, so bound_ctrl never applies
	const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
	const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
	return __hiloint2double(rhi, rlo);
}

__device__ __forceinline__ double bperm_f64(double v, int src_lane) {
	const int addr = src_lane << 2;
	const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
	const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
	return __hiloint2double(hi, lo);
}

// DPP controls (GFX9 encoding): wave_rol:1 = 0x134, wave_ror:1 = 0x13C,
// row_ror:n = 0x120+n, quad_perm = 8-bit lane selector.
constexpr int kWaveRol1 = 0x134;
constexpr int kWaveRor1 = 0x13C;
constexpr int kRowRor1 = 0x121;
constexpr int kRowRor15 = 0x12F;
constexpr int kQuadLeft = (3 << 0) | (0 << 2) | (1 << 4) | (2 << 6);  // lane i <- lane (i-1)&3
constexpr int kQuadRight = (1 << 0) | (2 << 2) | (3 << 4) | (0 << 6); // lane i <- lane (i+1)&3
constexpr int kQuadSwap = (1 << 0) | (0 << 2) | (3 << 4) | (2 << 6);  // lane i <- lane i^1

// Value held by the lane on the left (lane-1 within the chain's L-lane ring).
template <int L>
__device__ __forceinline__ double from_left(double v, int lane) {
	if constexpr (L == 1) return v;
	else if constexpr (L == 2) return dpp_f64<kQuadSwap>(v);
	else if constexpr (L == 4) return dpp_f64<kQuadLeft>(v);
	else if constexpr (L == 16) return dpp_f64<kRowRor1>(v);
	else if constexpr (L == 64) return dpp_f64<kWaveRor1>(v);
	else return bperm_f64(v, (lane & ~(L - 1)) | ((lane - 1) & (L - 1)));
}

// Value held by the lane on the right (lane+1 within the ring).
template <int L>
__device__ __forceinline__ double from_right(double v, int lane) {
	if constexpr (L == 1) return v;
	else if constexpr (L == 2) return dpp_f64<kQuadSwap>(v);
	else if constexpr (L == 4) return dpp_f64<kQuadRight>(v);
	else if constexpr (L == 16) return dpp_f64<kRowRor15>(v);
	else if constexpr (L == 64) return dpp_f64<kWaveRol1>(v);
	else return bperm_f64(v, (lane & ~(L - 1)) | ((lane + 1) & (L - 1)));
}

// The same moves for FP32 values (the FP32-compensated kernels).
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
	return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

template <int L>
__device__ __forceinline__ float from_left(float v, int lane) {
	if constexpr (L == 1) return v;
	else if constexpr (L == 2) return dpp_f32<kQuadSwap>(v);
	else if constexpr (L == 4) return dpp_f32<kQuadLeft>(v);
	else if constexpr (L == 16) return dpp_f32<kRowRor1>(v);
	else if constexpr (L == 64) return dpp_f32<kWaveRor1>(v);
	else
		return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane & ~(L - 1)) | ((lane - 1) & (L - 1))) << 2,
		                                                   __float_as_int(v)));
}

template <int L>
__device__ __forceinline__ float from_right(float v, int lane) {
	if constexpr (L == 1) return v;
	else if constexpr (L == 2) return dpp_f32<kQuadSwap>(v);
	else if constexpr (L == 4) return dpp_f32<kQuadRight>(v);
	else if constexpr (L == 16) return dpp_f32<kRowRor15>(v);
	else if constexpr (L == 64) return dpp_f32<kWaveRol1>(v);
	else
		return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane & ~(L - 1)) | ((lane + 1) & (L - 1))) << 2,
		                                                   __float_as_int(v)));
}

// Force of DNKG / dDNKG on a site with neighbours l, r and mass = m_x + 2:
// l + r - q (m_x + 2 + beta q^2) (DNKG_FPUT_Toda.cu:24-31, dDNKG.cu:17-19).
// kFormCubic works on sigma q with sigma = sqrt(beta), where the cubic
// coefficient is 1: one FMA for m + 2 + q^2 instead of a product and an FMA.
template <int FORM>
__device__ __forceinline__ double site_force(double l, double q, double r, double mass, const StepArgs &a) {
	const double t = (FORM == kFormCubic) ? fma(q, q, mass) : fma(a.beta, q * q, mass);
	return fma(-q, t, l + r);
}

// Bond force term of FPUT / Toda for the bond x = q_{i} - q_{i-1}; the force on
// site i is bond(x_{i+1}) - bond(x_i).
//   FPUT: V'(x) = x + alpha x^2 + beta x^3  (Horner), the same force as
//         (l+r-2c)(1+alpha(r-l)) + beta(d_r^3-d_l^3) of DNKG_FPUT_Toda.cu:33-50;
//         on sigma q, kFormCubic (alpha = 0, sigma = sqrt(beta)) is x + x^3 and
//         kFormQuad (beta = 0, sigma = alpha) is x + x^2;
//   Toda: e^{2 alpha x}, the force of DNKG_FPUT_Toda.cu:52-66 without its factor
//         1/(2 alpha), which the kick folds into its weight (the constant
//         -1/(2 alpha) of V' cancels in the difference), evaluated as
//         2^(x'/128) on the bond x' = v_i - v_{i-1} of v = (256 alpha/ln2) q,
//         table driven (fewest instructions), or for S >= 25 sites per lane,
//         where the state takes 100+ VGPRs and hoisted table loads would
//         leave one wave per SIMD, the table-free form, 2^x' on the bond of
//         v = (2 alpha/ln2) q (toda_unit).
template <int MODEL, int S, int FORM>
__device__ __forceinline__ double bond_force(double x, const StepArgs &a, const double *etab) {
	if constexpr (MODEL == FPUT) {
		if constexpr (FORM == kFormCubic) return fma(x * x, x, x);
		else if constexpr (FORM == kFormQuad) return fma(x, x, x);
		else return x * fma(x, fma(a.beta, x, a.alpha), 1.0);
	} else if constexpr (S >= 25) return exp2_poly(x); // Toda: x = v_i - v_{i-1}
	else return exp2_128_tab(x, etab);
}

// Unit of the Toda state in a kernel with S sites per lane: 1 for the table
// form's v = (256 alpha/ln2) q (bond exponent in 1/128 octaves), 2^-7 for the
// table-free form's (2 alpha/ln2) q (octaves).  Scaling by a power of two is
// exact, so both hold the same trajectory.
template <int MODEL, int S>
__device__ __forceinline__ constexpr double toda_unit() {
	return (MODEL == TODA && S >= 25) ? 0x1p-7 : 1.0;
}

// The state as the kernels hold it: sigma q and sigma p for the scaled forms
// (then sigma F(q) = F~(sigma q), so drift and kick weights stay as they
// are), v = sigma q with unscaled p for Toda (sigma = 256 alpha/ln2 times
// toda_unit), else q and p.  The way back multiplies by 1/sigma: a division
// would put its 10-instruction sequence between the stores of a lane, and the
// L2 then writes partly filled lines back (a third more HBM write traffic
// measured).
template <int MODEL, int FORM, int S>
__device__ __forceinline__ double2 to_registers(double2 v, const StepArgs &a) {
	if constexpr (MODEL == TODA) return make_double2((a.scale * toda_unit<MODEL, S>()) * v.x, v.y);
	else if constexpr (FORM != kFormPlain) return make_double2(a.scale * v.x, a.scale * v.y);
	else return v;
}

template <int MODEL, int FORM, int S>
__device__ __forceinline__ double2 from_registers(double q, double p, const StepArgs &a) {
	if constexpr (MODEL == TODA) return make_double2(q * (a.inv_scale / toda_unit<MODEL, S>()), p);
	else if constexpr (FORM != kFormPlain) return make_double2(q * a.inv_scale, p * a.inv_scale);
	else return make_double2(q, p);
}

