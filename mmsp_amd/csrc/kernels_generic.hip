// kernels_generic.hip -- one thread per node, any dimension, any boundary
// condition, every precision form.
//
// These kernels are the general replacement for the per-node loops of the
// reference's update() bodies (allen-cahn.cpp:68-71, cahn-hilliard.cpp:69-76,
// graingrowth.cpp:91-107). Neighbours outside the grid are mapped the way
// check_boundary maps them (MMSP.utility.cpp:42-57), so no ghost layer is
// stored: the periodic "halo fill" happens in the index arithmetic. Cahn-
// Hilliard runs its two passes (mu, then the update) exactly as the reference
// does, so non-periodic boundaries see mu through the same mapping.
//
// The fused/tiled kernels (kernels_fused.hip) take over the hot shapes; these
// serve every other case and are the in-library baseline the tuner compares
// against.
#include "device.hpp"

namespace mmspx {

template <int DIM>
__device__ __forceinline__ void coords(const Geom& g, int i, int x[3])
{
	if (DIM == 1) {
		x[0] = i;
	} else if (DIM == 2) {
		x[0] = i / g.n[1];
		x[1] = i - x[0] * g.n[1];
	} else {
		x[0] = i / g.s[0];
		int r = i - x[0] * g.s[0];
		x[1] = r / g.n[2];
		x[2] = r - x[1] * g.n[2];
	}
}

// node index of the neighbour at offset `off` along axis d
__device__ __forceinline__ int nbr(const Geom& g, int i, const int x[3], int d, int off)
{
	int xn = x[d] + off;
	if (xn < 0 || xn >= g.n[d]) xn = bc_map(xn, g.n[d], g.b0[d], g.b1[d]);
	return i + (xn - x[d]) * g.s[d];
}

template <int DIM, int PREC>
__device__ __forceinline__ typename Form<PREC>::L lap_at(const Geom& g, const typename Form<PREC>::S* __restrict__ a,
                                                        int i, const int x[3])
{
	using F = Form<PREC>;
	typename F::L lap = F::lap_zero();
	typename F::S y = a[i];
#pragma unroll
	for (int d = 0; d < DIM; d++) {
		typename F::S yh = a[nbr(g, i, x, d, +1)];
		typename F::S yl = a[nbr(g, i, x, d, -1)];
		F::lap_add(lap, g.w[d], g.wf[d], yh, y, yl);
	}
	return lap;
}

// NOISE: model A (model_A.cpp:97-101), the deviate drawn for (step, node i)
template <int DIM, int PREC, bool NOISE>
__global__ void __launch_bounds__(256) ac_generic(const typename Form<PREC>::S* __restrict__ a,
                                                  typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p)
{
	using F = Form<PREC>;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3];
		coords<DIM>(g, i, x);
		if constexpr (NOISE)
			b[i] = F::ac_noise(a[i], lap_at<DIM, PREC>(g, a, i, x),
			                   normal_deviate(p.seed, p.step, x[0], (uint32_t)(i - x[0] * g.s[0])), p);
		else
			b[i] = F::ac(a[i], lap_at<DIM, PREC>(g, a, i, x), p);
	}
}

// MU as mu_of: 0 the reference's polynomial, 1 double well, 2 polynomial plus noise
template <int DIM, int PREC, int MU>
__global__ void __launch_bounds__(256) ch_mu_generic(const typename Form<PREC>::S* __restrict__ a,
                                                     typename Form<PREC>::S* __restrict__ mu, Geom g, ModelParams p)
{
	using F = Form<PREC>;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3];
		coords<DIM>(g, i, x);
		typename F::L lap = lap_at<DIM, PREC>(g, a, i, x);
		float gd = 0.0f;
		if constexpr (MU == 2) gd = normal_deviate(p.seed, p.step, x[0], (uint32_t)(i - x[0] * g.s[0]));
		mu[i] = mu_of<PREC, MU>(a[i], lap, p, gd);
	}
}

template <int DIM, int PREC>
__global__ void __launch_bounds__(256) ch_update_generic(const typename Form<PREC>::S* __restrict__ a,
                                                         const typename Form<PREC>::S* __restrict__ mu,
                                                         typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p)
{
	using F = Form<PREC>;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3];
		coords<DIM>(g, i, x);
		b[i] = F::ch(a[i], lap_at<DIM, PREC>(g, mu, i, x), p);
	}
}

// Grain growth on field-major storage (plane j at offset j*nodes): the
// reference's per-node vector<T> becomes Q coalesced planes. ZEN: Zener
// pinning (zener.cpp:117-135), field 0 -- the particles -- keeps its value.
template <int DIM, int PREC, bool ZEN>
__global__ void __launch_bounds__(256) gg_generic(const typename Form<PREC>::S* __restrict__ a,
                                                  typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p)
{
	using F = Form<PREC>;
	using S = typename F::S;
	const int64_t N = g.nodes;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3];
		coords<DIM>(g, i, x);
		int ih[3], il[3];
#pragma unroll
		for (int d = 0; d < DIM; d++) {
			ih[d] = nbr(g, i, x, d, +1);
			il[d] = nbr(g, i, x, d, -1);
		}
		typename F::Sum sum = 0;
		for (int j = 0; j < g.fields; j++) {
			S v = a[j * N + i];
			sum += v * v;
		}
		for (int j = 0; j < g.fields; j++) {
			const S* __restrict__ pl = a + j * N;
			S y = pl[i];
			typename F::L lap = F::lap_zero();
#pragma unroll
			for (int d = 0; d < DIM; d++) F::lap_add(lap, g.w[d], g.wf[d], pl[ih[d]], y, pl[il[d]]);
			b[j * N + i] = (ZEN && j == 0) ? y : F::gg(y, sum, lap, p);
		}
	}
}

// Ostwald ripening (ostwald.cpp:87-122), field 0 the concentration C. Pass 1:
// w (the reference's first node loop) and the order parameters' update (the
// j-loop of its second), both from the old state; pass 2: C' = C + dt*D*lap(w).
template <int DIM, int PREC>
__global__ void __launch_bounds__(256) ow_pass1(const typename Form<PREC>::S* __restrict__ a,
                                                typename Form<PREC>::S* __restrict__ w,
                                                typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p)
{
	using F = Form<PREC>;
	using S = typename F::S;
	const int64_t N = g.nodes;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3];
		coords<DIM>(g, i, x);
		int ih[3], il[3];
#pragma unroll
		for (int d = 0; d < DIM; d++) {
			ih[d] = nbr(g, i, x, d, +1);
			il[d] = nbr(g, i, x, d, -1);
		}
		typename F::OSum sum = 0;
		for (int j = 1; j < g.fields; j++) {
			const S v = a[j * N + i];
			sum += v * v;
		}
		const S C = a[i];
		typename F::L lap0 = F::lap_zero();
#pragma unroll
		for (int d = 0; d < DIM; d++) F::lap_add(lap0, g.w[d], g.wf[d], a[ih[d]], C, a[il[d]]);
		w[i] = F::ow_w(C, sum, lap0, p);
		for (int j = 1; j < g.fields; j++) {
			const S* __restrict__ pl = a + j * N;
			const S v = pl[i];
			typename F::L lap = F::lap_zero();
#pragma unroll
			for (int d = 0; d < DIM; d++) F::lap_add(lap, g.w[d], g.wf[d], pl[ih[d]], v, pl[il[d]]);
			b[j * N + i] = F::ow_phi(v, C, sum, lap, p);
		}
	}
}

// Multi-phase (anisotropic) grain growth, grain_growth/anisotropic/phase_field/
// graingrowth.cpp:108-188, one thread per node, at most kAnisoMaxFields
// fields. The fields present in the node's 2*dim+1 neighbourhood (mapped as
// check_boundary maps) form the active set; below two the node is copied.
// Otherwise gga_node (device.hpp) updates it.
// QMAX: the field count the per-node arrays are compiled for; every loop over
// fields runs to QMAX with the fields past g.fields skipped, so for QMAX up to
// 8 the loops unroll, the arrays are indexed by constants and live in
// registers (the reference's example has 4 fields).
template <int DIM, int PREC, int QMAX>
__global__ void __launch_bounds__(256) gga_generic(const typename Form<PREC>::S* __restrict__ a,
                                                   typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p)
{
	using F = Form<PREC>;
	using S = typename F::S;
	const int64_t N = g.nodes;
	const int Q = g.fields;
	// full unrolling (register arrays) up to 8 fields; more keep their loops
	constexpr int UN = QMAX <= 8 ? QMAX : 1;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3];
		coords<DIM>(g, i, x);
		int ih[3], il[3];
#pragma unroll
		for (int d = 0; d < DIM; d++) {
			ih[d] = nbr(g, i, x, d, +1);
			il[d] = nbr(g, i, x, d, -1);
		}
		// the fields present in the neighbourhood (graingrowth.cpp:116-131)
		uint32_t act = 0;
		S cnt = 0;
		S phi[QMAX];
#pragma unroll UN
		for (int h = 0; h < QMAX; h++) {
			phi[h] = 0;
			if (h < Q) {
				const S* __restrict__ pl = a + h * N;
				phi[h] = pl[i];
				bool on = phi[h] > S(0);
#pragma unroll
				for (int d = 0; d < DIM; d++) on = on || pl[ih[d]] > S(0) || pl[il[d]] > S(0);
				if (on) {
					act |= 1u << h;
					cnt += S(1);
				}
			}
		}
		if (cnt >= S(2)) {
			S lap[QMAX];
#pragma unroll UN
			for (int h = 0; h < QMAX; h++) {
				typename F::L l = F::lap_zero();
				if (act >> h & 1u) {
					const S* __restrict__ pl = a + h * N;
#pragma unroll
					for (int d = 0; d < DIM; d++) F::lap_add(l, g.w[d], g.wf[d], pl[ih[d]], phi[h], pl[il[d]]);
				}
				lap[h] = l;
			}
			gga_node<PREC, QMAX>(phi, lap, act, cnt, Q, p);
		}
#pragma unroll UN
		for (int h = 0; h < QMAX; h++)
			if (h < Q) b[h * N + i] = phi[h];
	}
}

template <int DIM, int PREC>
__global__ void __launch_bounds__(256) ow_pass2(const typename Form<PREC>::S* __restrict__ a,
                                                const typename Form<PREC>::S* __restrict__ w,
                                                typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p)
{
	using F = Form<PREC>;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nodes; i += gridDim.x * blockDim.x) {
		int x[3];
		coords<DIM>(g, i, x);
		b[i] = F::ow_c(a[i], lap_at<DIM, PREC>(g, w, i, x), p);
	}
}

static int grid_blocks(int64_t n)
{
	int64_t b = (n + 255) / 256;
	if (b > 16384) b = 16384;
	if (b < 1) b = 1;
	return (int)b;
}

template <int DIM, int PREC>
static int launch_dim_prec(const Geom& g, const ModelParams& p, const void* av, void* bv, void* auxv, hipStream_t st,
                           int64_t* launches)
{
	using S = typename Form<PREC>::S;
	const S* a = (const S*)av;
	S* b = (S*)bv;
	S* aux = (S*)auxv;
	const int blocks = grid_blocks(g.nodes);
	switch (p.model) {
	case MMSPX_AC:
		hipLaunchKernelGGL((ac_generic<DIM, PREC, false>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
		*launches += 1;
		break;
	case MMSPX_AC_NOISE:
		hipLaunchKernelGGL((ac_generic<DIM, PREC, true>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
		*launches += 1;
		break;
	case MMSPX_CH:
		hipLaunchKernelGGL((ch_mu_generic<DIM, PREC, 0>), dim3(blocks), dim3(256), 0, st, a, aux, g, p);
		hipLaunchKernelGGL((ch_update_generic<DIM, PREC>), dim3(blocks), dim3(256), 0, st, a, aux, b, g, p);
		*launches += 2;
		break;
	case MMSPX_CH_NOISE:
		hipLaunchKernelGGL((ch_mu_generic<DIM, PREC, 2>), dim3(blocks), dim3(256), 0, st, a, aux, g, p);
		hipLaunchKernelGGL((ch_update_generic<DIM, PREC>), dim3(blocks), dim3(256), 0, st, a, aux, b, g, p);
		*launches += 2;
		break;
	case MMSPX_CH_DW:
		hipLaunchKernelGGL((ch_mu_generic<DIM, PREC, 1>), dim3(blocks), dim3(256), 0, st, a, aux, g, p);
		hipLaunchKernelGGL((ch_update_generic<DIM, PREC>), dim3(blocks), dim3(256), 0, st, a, aux, b, g, p);
		*launches += 2;
		break;
	case MMSPX_GG:
		hipLaunchKernelGGL((gg_generic<DIM, PREC, false>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
		*launches += 1;
		break;
	case MMSPX_ZENER:
		hipLaunchKernelGGL((gg_generic<DIM, PREC, true>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
		*launches += 1;
		break;
	case MMSPX_GG_ANISO:
		if constexpr (PREC == MMSPX_F32REF) {
			return fail(MMSPX_EUNSUPPORTED, "anisotropic grain growth runs the MMSPX_F64 and MMSPX_F32 forms");
		} else {
			if (g.fields <= 4)
				hipLaunchKernelGGL((gga_generic<DIM, PREC, 4>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
			else if (g.fields <= 8)
				hipLaunchKernelGGL((gga_generic<DIM, PREC, 8>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
			else if (g.fields <= 16)
				hipLaunchKernelGGL((gga_generic<DIM, PREC, 16>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
			else
				hipLaunchKernelGGL((gga_generic<DIM, PREC, kAnisoMaxFields>), dim3(blocks), dim3(256), 0, st, a, b, g, p);
			*launches += 1;
		}
		break;
	case MMSPX_OSTWALD:
		if constexpr (PREC == MMSPX_F32REF) {
			return fail(MMSPX_EUNSUPPORTED, "Ostwald ripening runs the MMSPX_F64 and MMSPX_F32 forms");
		} else {
			hipLaunchKernelGGL((ow_pass1<DIM, PREC>), dim3(blocks), dim3(256), 0, st, a, aux, b, g, p);
			hipLaunchKernelGGL((ow_pass2<DIM, PREC>), dim3(blocks), dim3(256), 0, st, a, aux, b, g, p);
			*launches += 2;
		}
		break;
	default:
		return fail(MMSPX_EINVAL, "unknown model");
	}
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int DIM>
static int launch_dim(const Geom& g, int prec, const ModelParams& p, const void* a, void* b, void* aux, hipStream_t st,
                      int64_t* launches)
{
	switch (prec) {
	case MMSPX_F64: return launch_dim_prec<DIM, MMSPX_F64>(g, p, a, b, aux, st, launches);
	case MMSPX_F32: return launch_dim_prec<DIM, MMSPX_F32>(g, p, a, b, aux, st, launches);
	case MMSPX_F32REF: return launch_dim_prec<DIM, MMSPX_F32REF>(g, p, a, b, aux, st, launches);
	}
	return fail(MMSPX_EINVAL, "unknown precision");
}

int launch_generic_step(const Geom& g, int prec, const ModelParams& p, const void* a, void* b, void* aux,
                        hipStream_t st, int64_t* launches)
{
	switch (g.dim) {
	case 1: return launch_dim<1>(g, prec, p, a, b, aux, st, launches);
	case 2: return launch_dim<2>(g, prec, p, a, b, aux, st, launches);
	case 3: return launch_dim<3>(g, prec, p, a, b, aux, st, launches);
	}
	return fail(MMSPX_EINVAL, "dimension must be 1, 2 or 3");
}

// ---- layout and type conversion --------------------------------------------------
template <typename T>
__global__ void transpose_kernel(const T* __restrict__ src, T* __restrict__ dst, int nodes, int fields, bool to_fm)
{
	const int64_t total = (int64_t)nodes * fields;
	for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
		// k indexes the destination
		if (to_fm) {
			int64_t j = k / nodes, i = k - j * nodes;
			dst[k] = src[i * fields + j];
		} else {
			int64_t i = k / fields, j = k - i * fields;
			dst[k] = src[j * nodes + i];
		}
	}
}

int launch_transpose(const void* src, void* dst, int nodes, int fields, size_t elem, bool to_fm, hipStream_t st)
{
	const int blocks = grid_blocks((int64_t)nodes * fields);
	if (elem == 8)
		hipLaunchKernelGGL(transpose_kernel<double>, dim3(blocks), dim3(256), 0, st, (const double*)src, (double*)dst,
		                   nodes, fields, to_fm);
	else
		hipLaunchKernelGGL(transpose_kernel<float>, dim3(blocks), dim3(256), 0, st, (const float*)src, (float*)dst,
		                   nodes, fields, to_fm);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <typename A, typename B>
__global__ void convert_kernel(const A* __restrict__ src, B* __restrict__ dst, int64_t n)
{
	for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
		dst[k] = (B)src[k];
}

int launch_convert(const void* src, int src_f64, void* dst, int dst_f64, int64_t n, hipStream_t st)
{
	const int blocks = grid_blocks(n);
	if (src_f64 && !dst_f64)
		hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks), dim3(256), 0, st, (const double*)src,
		                   (float*)dst, n);
	else if (!src_f64 && dst_f64)
		hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks), dim3(256), 0, st, (const float*)src,
		                   (double*)dst, n);
	else if (src_f64)
		hipLaunchKernelGGL((convert_kernel<double, double>), dim3(blocks), dim3(256), 0, st, (const double*)src,
		                   (double*)dst, n);
	else
		hipLaunchKernelGGL((convert_kernel<float, float>), dim3(blocks), dim3(256), 0, st, (const float*)src,
		                   (float*)dst, n);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

// Plain streaming copy: the HBM calibration kernel. Each thread moves four
// 16-byte vectors per iteration (loads issued before stores) over a
// grid-stride range; n4 counts groups of four vectors.
__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n4)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n4; k += stride) {
		uint4 v0 = src[k], v1 = src[k + n4], v2 = src[k + 2 * n4], v3 = src[k + 3 * n4];
		dst[k] = v0;
		dst[k + n4] = v1;
		dst[k + 2 * n4] = v2;
		dst[k + 3 * n4] = v3;
	}
}

int launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st)
{
	const int64_t n4 = (int64_t)(bytes / (4 * sizeof(uint4)));
	int cus = 256, dev = 0;
	if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	hipLaunchKernelGGL(copy_kernel, dim3(cus * 8), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n4);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

}  // namespace mmspx
