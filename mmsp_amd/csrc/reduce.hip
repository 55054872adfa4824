// reduce.hip -- FP64 compensated reductions: per-field mass and free energy.
//
// The reference sums energies and masses naively in double on the host
// (convex_splitting/cahn-hilliard.cpp:179-213, graingrowth.cpp:131-136). Here
// every thread keeps a Neumaier (sum, compensation) pair over a grid-stride
// range, pairs are combined across the 64-lane wavefront with shuffles and
// across the block through LDS, and each block writes one pair. The host adds
// the block pairs in long double. The block count is fixed, so the result does
// not depend on scheduling.
//
// Free energy densities (forward differences, neighbours through
// check_boundary; dx folded in as (u(x+dx)-u(x))/dx):
//   AC, CH   : -r/2 c^2 + u/4 c^4 + K/2 |grad c|^2
//   CH_DW    : c^2 (1-c)^2 + kappa/2 |grad c|^2
//   GG       : sum_j(-phi_j^2/2 + phi_j^4/4) + sum_{j<k} phi_j^2 phi_k^2 + 1 * sum_j |grad phi_j|^2
//              (kappa = 2, the coefficient of lap in graingrowth.cpp:105)
#include "device.hpp"

namespace mmspx {

struct NSum {
	double s, c;
};

__device__ __forceinline__ void nadd(NSum& a, double v)
{
	double t = a.s + v;
	if (fabs(a.s) >= fabs(v))
		a.c += (a.s - t) + v;
	else
		a.c += (v - t) + a.s;
	a.s = t;
}

__device__ __forceinline__ void ncombine(NSum& a, const NSum& b)
{
	nadd(a, b.s);
	a.c += b.c;
}

__device__ void block_reduce_store(NSum acc, double* out)
{
	__shared__ double ls[kRedThreads / 64], lc[kRedThreads / 64];
	for (int off = 32; off > 0; off >>= 1) {
		NSum o;
		o.s = __shfl_down(acc.s, off, 64);
		o.c = __shfl_down(acc.c, off, 64);
		ncombine(acc, o);
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) {
		ls[wave] = acc.s;
		lc[wave] = acc.c;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		NSum t = {ls[0], lc[0]};
		for (int w = 1; w < kRedThreads / 64; w++) ncombine(t, NSum{ls[w], lc[w]});
		out[2 * blockIdx.x] = t.s;
		out[2 * blockIdx.x + 1] = t.c;
	}
}

template <typename S>
__global__ void __launch_bounds__(kRedThreads) mass_kernel(const S* __restrict__ a, int nodes, double* part)
{
	NSum acc = {0.0, 0.0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += gridDim.x * blockDim.x) nadd(acc, (double)a[i]);
	block_reduce_store(acc, part);
}

struct EParams {
	int kind;  // 0 poly, 1 double well, 2 grain growth
	double r, u, kap;
};

template <typename S, int DIM>
__global__ void __launch_bounds__(kRedThreads) energy_kernel(const S* __restrict__ a, Geom g, EParams ep, double* part)
{
	NSum acc = {0.0, 0.0};
	const int64_t N = g.nodes;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3] = {0, 0, 0};
		int r = i;
		for (int d = 0; d < DIM; d++) {
			x[d] = r / g.s[d];
			r -= x[d] * g.s[d];
		}
		double e;
		if (ep.kind == 0) {
			double c = a[i], c2 = c * c;
			e = -0.5 * ep.r * c2 + 0.25 * ep.u * (c2 * c2);
		} else if (ep.kind == 1) {
			double c = a[i], m = 1.0 - c;
			e = (c * c) * (m * m);
		} else {
			double s2 = 0.0;
			e = 0.0;
			for (int j = 0; j < g.fields; j++) {
				double v = a[j * N + i];
				double q = v * v;
				e += -0.5 * q + 0.25 * (q * q);
				e += q * s2;
				s2 += q;
			}
		}
		double gsum = 0.0;
		for (int d = 0; d < DIM; d++) {
			int xn = x[d] + 1;
			if (xn >= g.n[d]) xn = bc_map(xn, g.n[d], g.b0[d], g.b1[d]);
			const int ih = i + (xn - x[d]) * g.s[d];
			for (int j = 0; j < g.fields; j++) {
				double df = ((double)a[j * N + ih] - (double)a[j * N + i]) / g.dx[d];
				gsum += df * df;
			}
		}
		e += 0.5 * ep.kap * gsum;
		nadd(acc, e);
	}
	block_reduce_store(acc, part);
}

template <typename S>
static int launch_energy(const Geom& g, const EParams& ep, const S* a, double* part, hipStream_t st)
{
	if (g.dim == 1)
		hipLaunchKernelGGL((energy_kernel<S, 1>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, a, g, ep, part);
	else if (g.dim == 2)
		hipLaunchKernelGGL((energy_kernel<S, 2>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, a, g, ep, part);
	else
		hipLaunchKernelGGL((energy_kernel<S, 3>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, a, g, ep, part);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

// MASS: part holds fields * kRedBlocks pairs; ENERGY: kRedBlocks pairs.
int launch_reduce(const Geom& g, int prec, int what, const ModelParams& p, const void* a, double* part,
                  hipStream_t st)
{
	const bool f64 = prec == MMSPX_F64;
	if (what == MMSPX_MASS) {
		for (int j = 0; j < g.fields; j++) {
			double* pj = part + (size_t)j * 2 * kRedBlocks;
			if (f64)
				hipLaunchKernelGGL(mass_kernel<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, st,
				                   (const double*)a + (size_t)j * g.nodes, g.nodes, pj);
			else
				hipLaunchKernelGGL(mass_kernel<float>, dim3(kRedBlocks), dim3(kRedThreads), 0, st,
				                   (const float*)a + (size_t)j * g.nodes, g.nodes, pj);
			MMSPX_HIP(hipGetLastError());
		}
		return 0;
	}
	EParams ep;
	if (p.model == MMSPX_AC || p.model == MMSPX_CH || noisy_model(p.model)) {
		// the noisy forms: the free energy of their deterministic part
		ep.kind = 0;
		ep.r = -p.nr;
		ep.u = p.u;
		ep.kap = p.K;
	} else if (p.model == MMSPX_CH_DW) {
		ep.kind = 1;
		ep.r = ep.u = 0.0;
		ep.kap = p.kappa;
	} else if (p.model == MMSPX_GG || p.model == MMSPX_ZENER) {
		// Zener pinning: the grain-growth functional (the particle field included)
		ep.kind = 2;
		ep.r = ep.u = 0.0;
		ep.kap = 2.0;
	} else {
		return fail(MMSPX_EUNSUPPORTED, "no free energy for this model");
	}
	if (f64) return launch_energy<double>(g, ep, (const double*)a, part, st);
	return launch_energy<float>(g, ep, (const float*)a, part, st);
}

double sum_partials(const double* part, int count)
{
	long double s = 0.0L, c = 0.0L;
	for (int k = 0; k < 2 * count; k++) {
		long double v = part[k];
		long double t = s + v;
		if (fabsl(s) >= fabsl(v))
			c += (s - t) + v;
		else
			c += (v - t) + s;
		s = t;
	}
	return (double)(s + c);
}

}  // namespace mmspx
