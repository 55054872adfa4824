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
	using Sum = double;
	__device__ static inline S gg(S y, Sum sum, L lap, const ModelParams& p)
	{
		return y - p.dt * (-y - cube_cr(y) + 2.0 * (y * sum - lap));
	}
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
	using Sum = float;
	__device__ static inline S gg(S y, Sum sum, L lap, const ModelParams& p)
	{
		return y - p.dtf * (-y - y * y * y + 2.0f * (y * sum - lap));
	}
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
	using Sum = float;
	__device__ static inline S gg(S y, Sum sum, L lap, const ModelParams& p)
	{
		double d = y;
		float inner = y * sum - lap;
		return (float)(d - p.dt * (-d - cube_cr(d) + 2.0 * (double)inner));
	}
};

}  // namespace mmspx
