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
#include <algorithm>
#include <thread>
#include <vector>

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

// ---- state digest ---------------------------------------------------------------
// A 64-bit digest of a state in the host's node-major order: the sum, modulo
// 2^64, over every element of mix(bits ^ mix(index * golden + c)), index =
// node * fields + field, bits = the element's bit pattern (a float's 32 bits
// zero-extended). A sum is independent of the order of its terms, so the
// device (field-major) and the host (node-major, any number of threads) get
// the same value; mix is the SplitMix64 finalizer.
__host__ __device__ __forceinline__ uint64_t digest_mix(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t digest_term(uint64_t index, uint64_t bits)
{
	return digest_mix(bits ^ digest_mix(index * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull));
}

template <typename S, typename U>
__global__ void __launch_bounds__(kRedThreads) digest_kernel(const S* __restrict__ a, int nodes, int fields,
                                                             uint64_t* part)
{
	uint64_t acc = 0;
	const int64_t total = (int64_t)nodes * fields;
	for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
		const int64_t j = k / nodes, n = k - j * nodes;  // field-major device index
		U u;
		const S v = a[k];
		memcpy(&u, &v, sizeof(U));
		acc += digest_term((uint64_t)(n * fields + j), (uint64_t)u);
	}
	for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
	__shared__ uint64_t lw[kRedThreads / 64];
	if ((threadIdx.x & 63) == 0) lw[threadIdx.x >> 6] = acc;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint64_t t = 0;
		for (int w = 0; w < kRedThreads / 64; w++) t += lw[w];
		part[blockIdx.x] = t;
	}
}

int launch_digest(const Geom& g, int prec, const void* a, uint64_t* part, hipStream_t st)
{
	if (prec == MMSPX_F64)
		hipLaunchKernelGGL((digest_kernel<double, uint64_t>), dim3(kRedBlocks), dim3(kRedThreads), 0, st,
		                   (const double*)a, g.nodes, g.fields, part);
	else
		hipLaunchKernelGGL((digest_kernel<float, uint32_t>), dim3(kRedBlocks), dim3(kRedThreads), 0, st,
		                   (const float*)a, g.nodes, g.fields, part);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

uint64_t host_digest(const void* data, int64_t count, int elem_bytes)
{
	const int64_t min_per_thread = (int64_t)1 << 20;
	int threads = (int)std::min<int64_t>(16, std::max<int64_t>(1, count / min_per_thread));
	threads = std::min<int>(threads, std::max(1u, std::thread::hardware_concurrency()));
	std::vector<uint64_t> sums((size_t)threads, 0);
	auto work = [&](int t) {
		const int64_t b = count * t / threads, e = count * (t + 1) / threads;
		uint64_t acc = 0;
		if (elem_bytes == 8) {
			const uint64_t* w = (const uint64_t*)data;
			for (int64_t i = b; i < e; i++) acc += digest_term((uint64_t)i, w[i]);
		} else {
			const uint32_t* w = (const uint32_t*)data;
			for (int64_t i = b; i < e; i++) acc += digest_term((uint64_t)i, (uint64_t)w[i]);
		}
		sums[(size_t)t] = acc;
	};
	std::vector<std::thread> pool;
	for (int t = 1; t < threads; t++) pool.emplace_back(work, t);
	work(0);
	for (auto& th : pool) th.join();
	uint64_t total = 0;
	for (uint64_t v : sums) total += v;
	return total;
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
