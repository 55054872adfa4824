// device.hpp -- device-side building blocks shared by every kernel file.
//
// All kernels are compiled with -ffp-contract=off: each + and * rounds on its
// own, in the order the reference's C++ expressions evaluate them, so the FP64
// form reproduces the reference's bits (the reference's x86-64 build has no
// FMA). The only fused operations are the explicit fma() calls of cube_cr.
#pragma once

#include <hip/hip_runtime.h>

#include "core.hpp"

namespace mmspx {

// check_boundary (MMSP.utility.cpp:42-57, serial branch) for local limits
// [0, n): maps a coordinate one or two cells outside the grid back inside.
__device__ __forceinline__ int bc_map(int x, int n, int b0, int b1)
{
	if (x < 0) {
		if (b0 == MMSPX_NEUMANN || b0 == MMSPX_DIRICHLET) return 0;
		if (b0 == MMSPX_PERIODIC) return n + x;
		if (b0 == MMSPX_MIRROR) return -x;
	} else if (x >= n) {
		if (b1 == MMSPX_NEUMANN || b1 == MMSPX_DIRICHLET) return n - 1;
		if (b1 == MMSPX_PERIODIC) return x - n;
		if (b1 == MMSPX_MIRROR) return 2 * (n - 1) - x;
	}
	return x;
}

// Correctly rounded x^3: x*x = p + e and p*x = hi + lo exactly (Dekker's
// two-product through fma), then one final rounding of hi + (lo + e*x). The
// reference calls libm pow(x, 3.0); glibc's result agrees with this except for
// about 1e-3 of inputs, where it is off by one ulp.
__device__ __forceinline__ double cube_cr(double x)
{
	double p = x * x;
	double e = __builtin_fma(x, x, -p);
	double hi = p * x;
	double lo = __builtin_fma(p, x, -hi);
	lo = lo + e * x;
	return hi + lo;
}

// ---- lane shifts across one 64-lane wave (DPP wave_shr:1 / wave_shl:1) -----
// from_prev: lane i receives lane i-1's value; from_next: lane i+1's. The lane
// with no source (lane 0, resp. 63) keeps `edge`. Callers keep every lane of
// the wave active at the call (the builtin is convergent, so the compiler does
// not move it under a branch).
__device__ __forceinline__ int dpp_from_prev(int v, int edge)
{
	return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_from_next(int v, int edge)
{
	return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float from_prev(float v, float edge)
{
	return __builtin_bit_cast(float, dpp_from_prev(__builtin_bit_cast(int, v), __builtin_bit_cast(int, edge)));
}
__device__ __forceinline__ float from_next(float v, float edge)
{
	return __builtin_bit_cast(float, dpp_from_next(__builtin_bit_cast(int, v), __builtin_bit_cast(int, edge)));
}
__device__ __forceinline__ double from_prev(double v, double edge)
{
	const int2 s = __builtin_bit_cast(int2, v), e = __builtin_bit_cast(int2, edge);
	return __builtin_bit_cast(double, make_int2(dpp_from_prev(s.x, e.x), dpp_from_prev(s.y, e.y)));
}
__device__ __forceinline__ double from_next(double v, double edge)
{
	const int2 s = __builtin_bit_cast(int2, v), e = __builtin_bit_cast(int2, edge);
	return __builtin_bit_cast(double, make_int2(dpp_from_next(s.x, e.x), dpp_from_next(s.y, e.y)));
}

// lane k's value of v, the same in every lane (k a wave-uniform lane index)
__device__ __forceinline__ float lane_value(float v, int k)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}
__device__ __forceinline__ double lane_value(double v, int k)
{
	const int2 s = __builtin_bit_cast(int2, v);
	return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(s.x, k), __builtin_amdgcn_readlane(s.y, k)));
}

// ---- noise of the stochastic forms -------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds, each two 32 x 32 -> 64-bit products
// of words 0 and 2 with the round multipliers, the high halves xored with
// words 1 and 3 and the round key; the key advances by the Weyl constants
// between rounds. c is the counter in and the four output words out.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
	for (int r = 0; r < 10; r++) {
		const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
		// one 32 x 32 -> 64-bit product gives both halves (v_mad_u64_u32)
		const uint64_t p0 = (uint64_t)m0 * c[0], p1 = (uint64_t)m1 * c[2];
		const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
		const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
		const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
		c[0] = n0;
		c[1] = lo1;
		c[2] = n2;
		c[3] = lo0;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
}

// natural log of u in (0, 1], a normal float, from +, -, *, / only (so the
// host restatement gives the same bits): u = m 2^e with m in (1/sqrt 2,
// sqrt 2], ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716, by its series to
// s^9 (float).
__device__ __forceinline__ float ln_unit(float u)
{
	const uint32_t bits = __float_as_uint(u);
	int e = (int)(bits >> 23) - 127;
	float m = __uint_as_float((bits & 0x7FFFFFu) | 0x3F800000u);
	if (m > 1.41421356f) {
		m = m * 0.5f;
		e += 1;
	}
	const float s = (m - 1.0f) / (m + 1.0f), z = s * s;
	float q = 1.0f / 9.0f;
	q = 1.0f / 7.0f + z * q;
	q = 1.0f / 5.0f + z * q;
	q = 1.0f / 3.0f + z * q;
	q = 1.0f + z * q;
	return (float)e * 0.693147182f + 2.0f * s * q;
}

// cos(2 pi t) and sin(2 pi t) for t in [0, 1) a multiple of 2^-24: the
// quadrant q and f = 4t - q are exact; with theta = f pi/2 folded into
// [0, pi/4] (cos(f pi/2) = sin((1-f) pi/2)), Taylor series to theta^10 (cos)
// and theta^9 (sin), below float rounding there, give A = cos(f pi/2) and
// B = sin(f pi/2); the quadrant then permutes and signs them.
struct CosSin {
	float c, s;
};
__device__ __forceinline__ CosSin cos_sin_2pi(float t)
{
	const float v = t * 4.0f;
	const int q = (int)v;
	const float f = v - (float)q;
	const bool hi = f > 0.5f;
	const float h = hi ? 1.0f - f : f;
	const float th = h * 1.57079633f, z = th * th;
	const float cs = 1.0f + z * (-1.0f / 2.0f + z * (1.0f / 24.0f + z * (-1.0f / 720.0f +
	                 z * (1.0f / 40320.0f + z * (-1.0f / 3628800.0f)))));
	const float sn = th * (1.0f + z * (-1.0f / 6.0f + z * (1.0f / 120.0f + z * (-1.0f / 5040.0f +
	                 z * (1.0f / 362880.0f)))));
	const float A = hi ? sn : cs, B = hi ? cs : sn;
	CosSin r;
	r.c = q == 0 ? A : q == 1 ? -B : q == 2 ? -A : B;
	r.s = q == 0 ? B : q == 1 ? A : q == 2 ? -B : -A;
	return r;
}

// The pair of standard normal deviates of one Philox4x32-10 block, key {seed
// low, seed high}, counter {r, xh, step low, step high}: Box-Muller on its
// first two words, u1 = (w0 + 1/2) 2^-32 in (0, 1], u2 = (w1 >> 8) 2^-24 in
// [0, 1), R = sqrt(-2 ln u1), the pair {R cos(2 pi u2), R sin(2 pi u2)}, in
// float. A node with axis-0 index x and index r within its axis-0 plane draws
// the first of pair (r, x >> 1) when x is even and the second when x is odd,
// so a kernel marching along axis 0 forms one pair per two planes. (The
// reference's gaussian(), model_A.cpp:13-39, keeps the second value of its
// polar pair for the next call in the same way; its rejection loop would
// leave the lanes of a wave waiting on each other, so Box-Muller here.)
__device__ __forceinline__ CosSin normal_pair(uint64_t seed, uint64_t step, uint32_t xh, uint32_t r)
{
	uint32_t c[4] = {r, xh, (uint32_t)step, (uint32_t)(step >> 32)};
	philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
	const float u1 = ((float)c[0] + 0.5f) * 0x1p-32f;
	const float u2 = (float)(c[1] >> 8) * 0x1p-24f;
	const float R = sqrtf(-2.0f * ln_unit(u1));
	const CosSin cs = cos_sin_2pi(u2);
	CosSin g;
	g.c = R * cs.c;
	g.s = R * cs.s;
	return g;
}

__device__ __forceinline__ float normal_deviate(uint64_t seed, uint64_t step, int x, uint32_t r)
{
	const CosSin g = normal_pair(seed, step, (uint32_t)x >> 1, r);
	return (x & 1) ? g.s : g.c;
}

// The same deviate for a kernel that visits the planes of a column in order:
// the partner plane's value (x ^ 1) is kept in `keep`, valid for plane `kx`.
__device__ __forceinline__ float normal_deviate_cached(uint64_t seed, uint64_t step, int x, uint32_t r, int kx,
                                                       float& keep)
{
	if (x == kx) return keep;
	const CosSin g = normal_pair(seed, step, (uint32_t)x >> 1, r);
	keep = (x & 1) ? g.c : g.s;
	return (x & 1) ? g.s : g.c;
}

// ---- precision forms -----------------------------------------------------
template <int PREC> struct Form;

// double storage, double arithmetic (grid<dim,double>)
template <> struct Form<MMSPX_F64> {
	using S = double;  // storage
	using L = double;  // laplacian accumulator
	__device__ static inline L lap_zero() { return 0.0; }
	__device__ static inline void lap_add(L& lap, double w, float, S yh, S y, S yl)
	{
		lap += w * (yh - 2.0 * y + yl);
	}
	__device__ static inline S ac(S y, L lap, const ModelParams& p)
	{
		return y - p.dtM * (p.nr * y + p.u * cube_cr(y) - p.K * lap);
	}
	__device__ static inline S mu_poly(S c, L lap, const ModelParams& p)
	{
		return p.nr * c + p.u * cube_cr(c) - p.K * lap;
	}
	__device__ static inline S mu_dw(S c, L lap, const ModelParams& p)
	{
		return 2.0 * c * (1.0 - c) * (1.0 - 2.0 * c) - p.kappa * lap;
	}
	__device__ static inline S ch(S c, L lapmu, const ModelParams& p) { return c + p.dtM * lapmu; }
	// unit weight (dx = 1) and unit gradient coefficient: w*x and K*x with w = K
	// = 1.0 are exactly x, so these drop the multiplications and keep the bits
	__device__ static inline void lap_add1(L& lap, S yh, S y, S yl) { lap += (yh - 2.0 * y + yl); }
	__device__ static inline S mu_poly1(S c, L lap, const ModelParams& p) { return p.nr * c + p.u * cube_cr(c) - lap; }
	__device__ static inline S mu_dw1(S c, L lap) { return 2.0 * c * (1.0 - c) * (1.0 - 2.0 * c) - lap; }
	using Sum = double;
	__device__ static inline S gg(S y, Sum sum, L lap, const ModelParams& p)
	{
		return y - p.dt * (-y - cube_cr(y) + 2.0 * (y * sum - lap));
	}
	// model_A.cpp:99-100 and model_B.cpp:100-101: noise = gaussian(0, sigma) = 0 + g*sigma
	__device__ static inline S ac_noise(S y, L lap, float gd, const ModelParams& p)
	{
		const double xi = 0.0 + (double)gd * p.sigma;
		return y - p.dtM * (p.nr * y + p.u * cube_cr(y) - p.K * lap + xi);
	}
	__device__ static inline S mu_noise(S c, L lap, float gd, const ModelParams& p)
	{
		const double xi = 0.0 + (double)gd * p.sigma;
		return p.nr * c + p.u * cube_cr(c) - p.K * lap + xi;
	}
	// Ostwald ripening, ostwald.cpp:97-99 (w), :108 (C) and :117-118 (order parameters);
	// pow(x, 2) as x*x, pow(x, 3) as the correctly rounded cube
	using OSum = double;
	__device__ static inline S ow_w(S C, OSum sum, L lap, const ModelParams& p)
	{
		const double* o = p.ow;
		return -o[OW_A] * (C - o[OW_CM]) + o[OW_B] * cube_cr(C - o[OW_CM]) + o[OW_DA] * cube_cr(C - o[OW_CA]) +
		       o[OW_DB] * cube_cr(C - o[OW_CB]) - o[OW_G] * (C - o[OW_CA]) * sum - o[OW_K] * lap;
	}
	__device__ static inline S ow_phi(S v, S C, OSum sum, L lap, const ModelParams& p)
	{
		const double* o = p.ow;
		const double d = C - o[OW_CA];
		return v - o[OW_DTL] * (-o[OW_G] * (d * d) + o[OW_DELTA] * cube_cr(v) + o[OW_EPS] * v * (sum - v * v) -
		                        o[OW_K] * lap);
	}
	__device__ static inline S ow_c(S C, L lapw, const ModelParams& p) { return C + p.ow[OW_DTD] * lapw; }
};

// float storage, float arithmetic in the same order (the fast FP32 path)
template <> struct Form<MMSPX_F32> {
	using S = float;
	using L = float;
	__device__ static inline L lap_zero() { return 0.0f; }
	__device__ static inline void lap_add(L& lap, double, float wf, S yh, S y, S yl)
	{
		lap += wf * (yh - 2.0f * y + yl);
	}
	__device__ static inline S ac(S y, L lap, const ModelParams& p)
	{
		float c3 = y * y * y;
		return y - p.dtMf * (p.nrf * y + p.uf * c3 - p.Kf * lap);
	}
	__device__ static inline S mu_poly(S c, L lap, const ModelParams& p)
	{
		return p.nrf * c + p.uf * (c * c * c) - p.Kf * lap;
	}
	__device__ static inline S mu_dw(S c, L lap, const ModelParams& p)
	{
		return 2.0f * c * (1.0f - c) * (1.0f - 2.0f * c) - p.kappaf * lap;
	}
	__device__ static inline S ch(S c, L lapmu, const ModelParams& p) { return c + p.dtMf * lapmu; }
	__device__ static inline void lap_add1(L& lap, S yh, S y, S yl) { lap += (yh - 2.0f * y + yl); }
	__device__ static inline S mu_poly1(S c, L lap, const ModelParams& p)
	{
		return p.nrf * c + p.uf * (c * c * c) - lap;
	}
	__device__ static inline S mu_dw1(S c, L lap) { return 2.0f * c * (1.0f - c) * (1.0f - 2.0f * c) - lap; }
	using Sum = float;
	__device__ static inline S gg(S y, Sum sum, L lap, const ModelParams& p)
	{
		return y - p.dtf * (-y - y * y * y + 2.0f * (y * sum - lap));
	}
	__device__ static inline S ac_noise(S y, L lap, float gd, const ModelParams& p)
	{
		const float xi = gd * p.sigmaf;
		float c3 = y * y * y;
		return y - p.dtMf * (p.nrf * y + p.uf * c3 - p.Kf * lap + xi);
	}
	__device__ static inline S mu_noise(S c, L lap, float gd, const ModelParams& p)
	{
		const float xi = gd * p.sigmaf;
		return p.nrf * c + p.uf * (c * c * c) - p.Kf * lap + xi;
	}
	using OSum = float;
	__device__ static inline S ow_w(S C, OSum sum, L lap, const ModelParams& p)
	{
		const float* o = p.owf;
		const float a = C - o[OW_CM], b = C - o[OW_CA], c = C - o[OW_CB];
		return -o[OW_A] * a + o[OW_B] * (a * a * a) + o[OW_DA] * (b * b * b) + o[OW_DB] * (c * c * c) -
		       o[OW_G] * b * sum - o[OW_K] * lap;
	}
	__device__ static inline S ow_phi(S v, S C, OSum sum, L lap, const ModelParams& p)
	{
		const float* o = p.owf;
		const float d = C - o[OW_CA];
		return v - o[OW_DTL] * (-o[OW_G] * (d * d) + o[OW_DELTA] * (v * v * v) + o[OW_EPS] * v * (sum - v * v) -
		                        o[OW_K] * lap);
	}
	__device__ static inline S ow_c(S C, L lapw, const ModelParams& p) { return C + p.owf[OW_DTD] * lapw; }
};

// float storage with the reference's T=float semantics: the laplacian adds a
// double term into a float per axis (MMSP.grid.cpp:2063, 2085); the update
// expression is double because dt, 2.0 and pow promote it, and is rounded to
// float on assignment (allen-cahn.cpp:70, graingrowth.cpp:105).
template <> struct Form<MMSPX_F32REF> {
	using S = float;
	using L = float;
	__device__ static inline L lap_zero() { return 0.0f; }
	__device__ static inline void lap_add(L& lap, double w, float, S yh, S y, S yl)
	{
		lap = (float)((double)lap + w * ((double)yh - 2.0 * (double)y + (double)yl));
	}
	__device__ static inline S ac(S y, L lap, const ModelParams& p)
	{
		double d = y;
		return (float)(d - p.dtM * (p.nr * d + p.u * cube_cr(d) - p.K * (double)lap));
	}
	__device__ static inline S mu_poly(S c, L lap, const ModelParams& p)
	{
		double d = c;
		return (float)(p.nr * d + p.u * cube_cr(d) - p.K * (double)lap);
	}
	__device__ static inline S mu_dw(S c, L lap, const ModelParams& p)
	{
		double d = c;
		return (float)(2.0 * d * (1.0 - d) * (1.0 - 2.0 * d) - p.kappa * (double)lap);
	}
	__device__ static inline S ch(S c, L lapmu, const ModelParams& p)
	{
		return (float)((double)c + p.dtM * (double)lapmu);
	}
	__device__ static inline void lap_add1(L& lap, S yh, S y, S yl)
	{
		lap = (float)((double)lap + ((double)yh - 2.0 * (double)y + (double)yl));
	}
	__device__ static inline S mu_poly1(S c, L lap, const ModelParams& p)
	{
		double d = c;
		return (float)(p.nr * d + p.u * cube_cr(d) - (double)lap);
	}
	__device__ static inline S mu_dw1(S c, L lap)
	{
		double d = c;
		return (float)(2.0 * d * (1.0 - d) * (1.0 - 2.0 * d) - (double)lap);
	}
	using Sum = float;
	__device__ static inline S gg(S y, Sum sum, L lap, const ModelParams& p)
	{
		double d = y;
		float inner = y * sum - lap;
		return (float)(d - p.dt * (-d - cube_cr(d) + 2.0 * (double)inner));
	}
	// T noise = gaussian(...): the double deviate rounded to float, then
	// promoted again in the double expression
	__device__ static inline S ac_noise(S y, L lap, float gd, const ModelParams& p)
	{
		double d = y;
		const float xi = (float)(0.0 + (double)gd * p.sigma);
		return (float)(d - p.dtM * (p.nr * d + p.u * cube_cr(d) - p.K * (double)lap + (double)xi));
	}
	__device__ static inline S mu_noise(S c, L lap, float gd, const ModelParams& p)
	{
		double d = c;
		const float xi = (float)(0.0 + (double)gd * p.sigma);
		return (float)(p.nr * d + p.u * cube_cr(d) - p.K * (double)lap + (double)xi);
	}
};

// mu of the Cahn-Hilliard family: MU 0 the reference's polynomial, 1 the
// double well, 2 the polynomial plus noise (gd the node's deviate)
template <int PREC, int MU>
__device__ __forceinline__ typename Form<PREC>::S mu_of(typename Form<PREC>::S c, typename Form<PREC>::L lap,
                                                        const ModelParams& p, float gd)
{
	using F = Form<PREC>;
	if constexpr (MU == 2) {
		return F::mu_noise(c, lap, gd, p);
	} else if constexpr (MU == 1) {
		(void)gd;
		return F::mu_dw(c, lap, p);
	} else {
		(void)gd;
		return F::mu_poly(c, lap, p);
	}
}

// Pairwise variational derivatives (grain_growth/anisotropic/phase_field/
// graingrowth.cpp:141-157) and rates (:160-167) over the field slots set in
// `act`, in the reference's loop and operation order (ascending slots); dp
// gets the rates, zero in the other slots. Up to 8 slots the loops unroll and
// the arrays live in registers.
template <int PREC, int QMAX>
__device__ __forceinline__ void gga_rates(const typename Form<PREC>::S (&phi)[QMAX],
                                          const typename Form<PREC>::S (&lap)[QMAX], uint32_t act,
                                          const ModelParams& p, typename Form<PREC>::S (&dp)[QMAX])
{
	using S = typename Form<PREC>::S;
	constexpr bool dbl = PREC == MMSPX_F64;
	constexpr int UN = QMAX <= 8 ? QMAX : 1;
	const S c2 = dbl ? (S)p.an[AN_C2] : (S)p.anf[AN_C2];
	const S w = dbl ? (S)p.an[AN_W] : (S)p.anf[AN_W];
	const S mu = dbl ? (S)p.an[AN_MU] : (S)p.anf[AN_MU];
	S dF[QMAX];
#pragma unroll UN
	for (int h = 0; h < QMAX; h++) {
		dF[h] = 0;
		dp[h] = 0;
	}
#pragma unroll UN
	for (int h = 0; h < QMAX; h++) {
		if (!(act >> h & 1u)) continue;
#pragma unroll UN
		for (int j = h + 1; j < QMAX; j++) {
			if (!(act >> j & 1u)) continue;
			dF[h] += c2 * lap[j] + w * phi[j];
			dF[j] += c2 * lap[h] + w * phi[h];
#pragma unroll UN
			for (int k = j + 1; k < QMAX; k++) {
				if (!(act >> k & 1u)) continue;
				dF[h] += S(3) * phi[j] * phi[k];
				dF[j] += S(3) * phi[k] * phi[h];
				dF[k] += S(3) * phi[h] * phi[j];
			}
		}
	}
#pragma unroll UN
	for (int h = 0; h < QMAX; h++) {
		if (!(act >> h & 1u)) continue;
#pragma unroll UN
		for (int j = h + 1; j < QMAX; j++) {
			if (!(act >> j & 1u)) continue;
			dp[h] -= mu * (dF[h] - dF[j]);
			dp[j] -= mu * (dF[j] - dF[h]);
		}
	}
}

// Multi-phase grain growth at one node with at least two fields present
// (graingrowth.cpp:141-185): `act` holds the present fields (cnt of them) and
// lap their laplacians. gga_rates, then the clamped update and the projection
// onto the simplex (:170-185); phi holds the node's values on entry and the
// new ones on return. Fields at or past Q are skipped. For 5 to 8 field slots,
// a node with at most 4 fields present packs them, in ascending order, into 4
// slots first: the same additions in the same order as the full loops (which
// skip the absent fields), on a quarter of the field triples.
template <int PREC, int QMAX>
__device__ __forceinline__ void gga_node(typename Form<PREC>::S (&phi)[QMAX], const typename Form<PREC>::S (&lap)[QMAX],
                                         uint32_t act, typename Form<PREC>::S cnt, int Q, const ModelParams& p)
{
	using S = typename Form<PREC>::S;
	constexpr bool dbl = PREC == MMSPX_F64;
	constexpr int UN = QMAX <= 8 ? QMAX : 1;
	constexpr int K = 4;
	const S dt = dbl ? (S)p.an[AN_DT] : (S)p.anf[AN_DT];
	S dp[QMAX];
	if constexpr (QMAX > K && QMAX <= 8) {
		if (cnt <= S(K)) {
			S cphi[K], clap[K], cdp[K];
#pragma unroll
			for (int s = 0; s < K; s++) cphi[s] = clap[s] = 0;
#pragma unroll
			for (int f = 0; f < QMAX; f++) {
				const bool on = act >> f & 1u;
				const int r = __popc(act & ((1u << f) - 1u));  // the field's rank among the present ones
#pragma unroll
				for (int s = 0; s < K; s++) {
					if (on && r == s) {
						cphi[s] = phi[f];
						clap[s] = lap[f];
					}
				}
			}
			gga_rates<PREC, K>(cphi, clap, (1u << (int)cnt) - 1u, p, cdp);
#pragma unroll
			for (int f = 0; f < QMAX; f++) {
				const bool on = act >> f & 1u;
				const int r = __popc(act & ((1u << f) - 1u));
				S v = 0;
#pragma unroll
				for (int s = 0; s < K; s++)
					if (on && r == s) v = cdp[s];
				dp[f] = v;
			}
		} else {
			gga_rates<PREC, QMAX>(phi, lap, act, p, dp);
		}
	} else {
		gga_rates<PREC, QMAX>(phi, lap, act, p, dp);
	}
	// clamped update and projection onto the simplex (:170-185)
	S sum = 0;
	const S rate = dt * (S(2) / cnt);
#pragma unroll UN
	for (int h = 0; h < QMAX; h++) {
		if (h >= Q) continue;
		S v = phi[h] + rate * dp[h];
		if (v > S(1)) v = S(1);
		if (v < S(0)) v = S(0);
		phi[h] = v;
		sum += v;
	}
	const S rsum = fabs(sum) > S(0) ? S(1) / sum : S(0);
#pragma unroll UN
	for (int h = 0; h < QMAX; h++)
		if (h < Q) phi[h] = phi[h] * rsum;
}

}  // namespace mmspx
