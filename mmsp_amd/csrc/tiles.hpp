// tiles.hpp -- per-tile device bodies shared by the single-grid kernels
// (kernels_fused.hip) and the batched/persistent kernels (batch.hip).
//
// Every body reproduces the operation order of the generic kernels (and so of
// the reference), so a grid gives the same bits whichever kernel advances it.
#pragma once

#include "device.hpp"

namespace mmspx {

__device__ __forceinline__ int wrap(int v, int n)
{
	v %= n;
	return v < 0 ? v + n : v;
}

// The node a tile position v (any integer) reads along an axis of n nodes:
// periodic sides wrap; on the other sides one or two cells outside map as
// check_boundary maps them (MMSP.utility.cpp:42-57), clamped into the grid;
// positions further out are never read and take node 0.
__device__ __forceinline__ int tile_map(int v, int n, int lo, int hi)
{
	if (v >= 0 && v < n) return v;
	if (v < 0 ? lo == MMSPX_PERIODIC : hi == MMSPX_PERIODIC) return wrap(v, n);
	if (v < -2 || v > n + 1) return 0;
	const int m = bc_map(v, n, lo, hi);
	return m < 0 ? 0 : (m >= n ? n - 1 : m);
}

// Fused Cahn-Hilliard tiles on a non-periodic axis compute mu on the whole
// tile, but a boundary node's c' reads mu at the node check_boundary maps its
// outer neighbour to, as the reference's second pass does: the node itself
// (Neumann, Dirichlet) or its inner neighbour (mirror).
template <typename S>
__device__ __forceinline__ void edge_mu(S& ml, S& mh, S m, int v, int n, int lo, int hi)
{
	const S l = ml, h = mh;
	if (v == 0 && lo != MMSPX_PERIODIC) ml = lo == MMSPX_MIRROR ? h : m;
	if (v == n - 1 && hi != MMSPX_PERIODIC) mh = hi == MMSPX_MIRROR ? l : m;
}

// ---- Cahn-Hilliard, one (TY x TZ) tile marching over planes [xb, xe) ---------
// Periodic grids only; DIM 2 uses a single row (TY = 1, no axis-1 halo).
// Two bodies with identical arithmetic (and so identical bits):
//
// ch_tile (LDS rings): phi planes go global -> registers (one plane ahead) ->
// a 4-slot LDS ring with a 2-cell halo; mu goes to a 3-slot LDS ring with a
// 1-cell halo; c' is written one plane behind. Every neighbour, the axis-0
// ones included, is read from LDS.
//
// ch_tile_rq (register queues): each thread owns fixed points of the mu
// region and keeps phi at planes P-1, P, P+1 and mu at P-2, P-1, P of those
// points in registers, so only the in-plane neighbours go through LDS (one
// phi plane and one mu plane, each double-buffered): fewer LDS operations per
// cell and half the LDS footprint, hence more workgroups per CU.
template <int PREC, int MU, int DIM, int TY, int TZ, int NT>
__device__ __forceinline__ void ch_tile(const typename Form<PREC>::S* __restrict__ a,
                                        typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                        int y0, int z0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	constexpr int HY = DIM == 3 ? 2 : 0;
	constexpr int PY = TY + 2 * HY, PZ = TZ + 4;
	constexpr int MY = TY + HY, MZ = TZ + 2;
	constexpr int PN = PY * PZ, MN = MY * MZ, ON = TY * TZ;
	constexpr int PREG = (PN + NT - 1) / NT;
	constexpr int MREG = (MN + NT - 1) / NT;
	constexpr int OREG = (ON + NT - 1) / NT;
	__shared__ S sphi[4][PN];
	__shared__ S smu[3][MN];

	const int nx = g.n[0];
	const int ny = DIM == 3 ? g.n[1] : 1;
	const int nz = DIM == 3 ? g.n[2] : g.n[1];
	const int64_t plane = (int64_t)ny * nz;

	int poff[PREG];
#pragma unroll
	for (int k = 0; k < PREG; k++) {
		const int e = threadIdx.x + k * NT;
		poff[k] = 0;
		if (e < PN) {
			const int r = e / PZ, c = e - (e / PZ) * PZ;
			const int y = DIM == 3 ? wrap(y0 - HY + r, ny) : 0;
			const int z = wrap(z0 - 2 + c, nz);
			poff[k] = y * nz + z;
		}
	}
	auto load_plane = [&](int x, S* regs) {
		const S* __restrict__ src = a + (int64_t)wrap(x, nx) * plane;
#pragma unroll
		for (int k = 0; k < PREG; k++) {
			const int e = threadIdx.x + k * NT;
			if (e < PN) regs[k] = src[poff[k]];
		}
	};
	auto store_plane = [&](int slot, const S* regs) {
#pragma unroll
		for (int k = 0; k < PREG; k++) {
			const int e = threadIdx.x + k * NT;
			if (e < PN) sphi[slot][e] = regs[k];
		}
	};

	S pre[PREG];
	// the previous tile of a persistent loop may still read the rings
	__syncthreads();
	load_plane(xb - 2, pre);
	store_plane((xb - 2) & 3, pre);
	load_plane(xb - 1, pre);
	store_plane((xb - 1) & 3, pre);
	load_plane(xb, pre);

	for (int P = xb - 1; P <= xe; P++) {
		// phi(P+1) is in registers: publish it, then start loading phi(P+2)
		store_plane((P + 1) & 3, pre);
		if (P + 2 <= xe + 1) load_plane(P + 2, pre);
		__syncthreads();
		{
			const S* __restrict__ fm = sphi[(P - 1) & 3];
			const S* __restrict__ fc = sphi[P & 3];
			const S* __restrict__ fp = sphi[(P + 1) & 3];
			S* __restrict__ mu = smu[wrap(P, 3)];
#pragma unroll
			for (int k = 0; k < MREG; k++) {
				const int e = threadIdx.x + k * NT;
				if (e < MN) {
					const int r = e / MZ, c = e - (e / MZ) * MZ;
					const int q = (r + HY / 2) * PZ + (c + 1);
					const S y = fc[q];
					typename F::L lap = F::lap_zero();
					F::lap_add(lap, g.w[0], g.wf[0], fp[q], y, fm[q]);
					if (DIM == 3) F::lap_add(lap, g.w[1], g.wf[1], fc[q + PZ], y, fc[q - PZ]);
					F::lap_add(lap, g.w[DIM - 1], g.wf[DIM - 1], fc[q + 1], y, fc[q - 1]);
					float gd = 0.0f;
					if constexpr (MU == 2)
						gd = normal_deviate(p.seed, p.step, wrap(P, nx),
						                    (uint32_t)((DIM == 3 ? wrap(y0 - HY / 2 + r, ny) * nz : 0) +
						                               wrap(z0 - 1 + c, nz)));
					mu[e] = mu_of<PREC, MU>(y, lap, p, gd);
				}
			}
		}
		__syncthreads();
		if (P - 1 >= xb) {
			const S* __restrict__ mm = smu[wrap(P - 2, 3)];
			const S* __restrict__ mc = smu[wrap(P - 1, 3)];
			const S* __restrict__ mp = smu[wrap(P, 3)];
			const S* __restrict__ fc = sphi[(P - 1) & 3];
			S* __restrict__ dst = b + (int64_t)(P - 1) * plane;
#pragma unroll
			for (int k = 0; k < OREG; k++) {
				const int e = threadIdx.x + k * NT;
				if (e < ON) {
					const int r = e / TZ, c = e - (e / TZ) * TZ;
					const int y = y0 + r, z = z0 + c;
					const int q = (r + HY / 2) * MZ + (c + 1);
					const S m = mc[q];
					typename F::L lap = F::lap_zero();
					F::lap_add(lap, g.w[0], g.wf[0], mp[q], m, mm[q]);
					if (DIM == 3) F::lap_add(lap, g.w[1], g.wf[1], mc[q + MZ], m, mc[q - MZ]);
					F::lap_add(lap, g.w[DIM - 1], g.wf[DIM - 1], mc[q + 1], m, mc[q - 1]);
					const S c0 = fc[(r + HY) * PZ + (c + 2)];
					if (y < ny && z < nz) dst[(int64_t)y * nz + z] = F::ch(c0, lap, p);
				}
			}
		}
	}
}

template <int PREC, int MU, int DIM, int TY, int TZ, int NT>
__device__ __forceinline__ void ch_tile_rq(const typename Form<PREC>::S* __restrict__ a,
                                           typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                           int y0, int z0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	constexpr int HY = DIM == 3 ? 2 : 0;
	constexpr int PY = TY + 2 * HY, PZ = TZ + 4;   // phi plane tile (2-cell halo)
	constexpr int MY = TY + HY, MZ = TZ + 2;       // mu region (1-cell halo): the threads' own points
	constexpr int PN = PY * PZ, MN = MY * MZ;
	constexpr int RN = PN - MN;                    // phi tile points outside the mu region
	constexpr int MREG = (MN + NT - 1) / NT;
	constexpr int RREG = (RN + NT - 1) / NT;
	__shared__ S sphi[2][PN];
	__shared__ S smu[2][MN];

	const int nx = g.n[0];
	const int ny = DIM == 3 ? g.n[1] : 1;
	const int nz = DIM == 3 ? g.n[2] : g.n[1];
	const int64_t plane = (int64_t)ny * nz;

	// own points: mu-region index e = tid + k*NT; (r, c) in the mu region is
	// phi tile (r + HY/2, c + 1), i.e. node (y0 - HY/2 + r, z0 - 1 + c)
	constexpr int AY = 1, AZ = DIM - 1;  // the BC slots of the in-plane axes
	int ooff[MREG], oq[MREG];
	bool oout[MREG];
	int ozy[MREG];  // node (y, z) of an in-grid own point, for the boundary mapping of mu
#pragma unroll
	for (int k = 0; k < MREG; k++) {
		const int e = threadIdx.x + k * NT;
		ooff[k] = 0;
		oq[k] = 0;
		oout[k] = false;
		ozy[k] = 0;
		if (e < MN) {
			const int r = e / MZ, c = e - (e / MZ) * MZ;
			const int y = DIM == 3 ? y0 - HY / 2 + r : 0;
			const int z = z0 - 1 + c;
			ooff[k] = (DIM == 3 ? tile_map(y, ny, g.b0[AY], g.b1[AY]) : 0) * nz + tile_map(z, nz, g.b0[AZ], g.b1[AZ]);
			oq[k] = (r + HY / 2) * PZ + (c + 1);
			const bool inner = (DIM == 3 ? (r >= 1 && r <= TY) : true) && c >= 1 && c <= TZ;
			oout[k] = inner && (DIM == 3 ? y < ny : true) && z < nz;
			ozy[k] = (y << 16) | z;
		}
	}
	// ring points: the phi tile's rows 0 and PY-1 and columns 0 and PZ-1
	int roff[RREG], rq[RREG];
#pragma unroll
	for (int j = 0; j < RREG; j++) {
		const int t = threadIdx.x + j * NT;
		roff[j] = 0;
		rq[j] = 0;
		if (t < RN) {
			int pr, pc;
			if (DIM == 3) {
				if (t < PZ) {
					pr = 0;
					pc = t;
				} else if (t < 2 * PZ) {
					pr = PY - 1;
					pc = t - PZ;
				} else {
					const int u = t - 2 * PZ;
					pr = 1 + u / 2;
					pc = (u & 1) ? PZ - 1 : 0;
				}
			} else {
				pr = 0;
				pc = t ? PZ - 1 : 0;
			}
			const int y = DIM == 3 ? tile_map(y0 - HY + pr, ny, g.b0[AY], g.b1[AY]) : 0;
			roff[j] = y * nz + tile_map(z0 - 2 + pc, nz, g.b0[AZ], g.b1[AZ]);
			rq[j] = pr * PZ + pc;
		}
	}
	// Loads are unconditional: lanes past the region load node 0 of the plane
	// (offset 0) and planes are mapped, so every address is inside the grid. With a
	// fixed count of loads per plane and no register copies of a plane in
	// flight, the compiler waits for a plane's values only where they are
	// first used, one plane after the load.
	auto load_own = [&](int x, S* v) {
		const S* __restrict__ src = a + (int64_t)tile_map(x, nx, g.b0[0], g.b1[0]) * plane;
#pragma unroll
		for (int k = 0; k < MREG; k++) v[k] = src[ooff[k]];
	};
	auto load_ring = [&](int x, S* v) {
		const S* __restrict__ src = a + (int64_t)tile_map(x, nx, g.b0[0], g.b1[0]) * plane;
#pragma unroll
		for (int j = 0; j < RREG; j++) v[j] = src[roff[j]];
	};

	// some axis is not periodic: boundary nodes read mu through the mapping
	bool edges = false;
	for (int d = 0; d < DIM; d++) edges |= g.b0[d] != MMSPX_PERIODIC || g.b1[d] != MMSPX_PERIODIC;

	// phi queue: four register planes whose roles (P-1, P, P+1, P+2 in flight)
	// rotate with the unrolled steps; ring phi: two planes (P, P+1 in flight)
	S f0[MREG], f1[MREG], f2[MREG], f3[MREG];
	S mm[MREG], mc[MREG], mp[MREG];  // mu at P-2, P-1, P
	S r0[RREG], r1[RREG];
#pragma unroll
	for (int k = 0; k < MREG; k++) f3[k] = mm[k] = mc[k] = mp[k] = 0;
#pragma unroll
	for (int j = 0; j < RREG; j++) r1[j] = 0;
	// the previous tile of a persistent loop may still read the planes
	__syncthreads();
	load_own(xb - 2, f0);
	load_own(xb - 1, f1);
	load_own(xb, f2);
	load_ring(xb - 1, r0);
	// noise (MU 2): the own points' deviates of the partner plane nkx
	int nkx = -1;
	float nkeep[MU == 2 ? MREG : 1];
	(void)nkeep;

	// one plane: fm, fc, fp hold phi at P-1, P, P+1; fn receives P+2; rc holds
	// the ring at P and rn receives P+1
	auto step = [&](int P, const S* fm, const S* fc, const S* fp, S* fn, const S* rc, S* rn) {
		// loads for the next planes go out first (past the chunk they fetch
		// wrapped planes that are not used)
		load_own(P + 2, fn);
		load_ring(P + 1, rn);
		S* __restrict__ lphi = sphi[P & 1];
#pragma unroll
		for (int k = 0; k < MREG; k++)
			if (threadIdx.x + k * NT < MN) lphi[oq[k]] = fc[k];
#pragma unroll
		for (int j = 0; j < RREG; j++)
			if (threadIdx.x + j * NT < RN) lphi[rq[j]] = rc[j];
		__syncthreads();
		// mu(P) at the own points: axis-0 neighbours from the queue, the rest from LDS
		S* __restrict__ lmu = smu[P & 1];
		const int nx_w = wrap(P, nx);
		const bool nhit = nx_w == nkx;
		(void)nhit;
#pragma unroll
		for (int k = 0; k < MREG; k++) {
			if (threadIdx.x + k * NT < MN) {
				const int q = oq[k];
				const S y = fc[k];
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], fp[k], y, fm[k]);
				if (DIM == 3) F::lap_add(lap, g.w[1], g.wf[1], lphi[q + PZ], y, lphi[q - PZ]);
				F::lap_add(lap, g.w[DIM - 1], g.wf[DIM - 1], lphi[q + 1], y, lphi[q - 1]);
				float gd = 0.0f;
				if constexpr (MU == 2) gd = normal_deviate_cached(p.seed, p.step, nx_w, (uint32_t)ooff[k], nkx, nkeep[k]);
				mp[k] = mu_of<PREC, MU>(y, lap, p, gd);
				lmu[threadIdx.x + k * NT] = mp[k];
			}
		}
		if constexpr (MU == 2) nkx = nhit ? -1 : (nx_w ^ 1);
		__syncthreads();
		// c'(P-1) at the own interior points
		if (P - 1 >= xb) {
			const S* __restrict__ pmu = smu[(P - 1) & 1];
			S* __restrict__ dst = b + (int64_t)(P - 1) * plane;
#pragma unroll
			for (int k = 0; k < MREG; k++) {
				if (oout[k]) {
					const int e = threadIdx.x + k * NT;
					const S m = mc[k];
					S xh = mp[k], xl = mm[k];
					S yh = DIM == 3 ? pmu[e + MZ] : m, yl = DIM == 3 ? pmu[e - MZ] : m;
					S zh = pmu[e + 1], zl = pmu[e - 1];
					if (edges) {
						edge_mu(xl, xh, m, P - 1, nx, g.b0[0], g.b1[0]);
						if (DIM == 3) edge_mu(yl, yh, m, ozy[k] >> 16, ny, g.b0[AY], g.b1[AY]);
						edge_mu(zl, zh, m, ozy[k] & 0xffff, nz, g.b0[AZ], g.b1[AZ]);
					}
					typename F::L lap = F::lap_zero();
					F::lap_add(lap, g.w[0], g.wf[0], xh, m, xl);
					if (DIM == 3) F::lap_add(lap, g.w[1], g.wf[1], yh, m, yl);
					F::lap_add(lap, g.w[DIM - 1], g.wf[DIM - 1], zh, m, zl);
					dst[ooff[k]] = F::ch(fm[k], lap, p);
				}
			}
		}
#pragma unroll
		for (int k = 0; k < MREG; k++) {
			mm[k] = mc[k];
			mc[k] = mp[k];
		}
	};
	for (int P = xb - 1; P <= xe; P += 4) {
		step(P, f0, f1, f2, f3, r0, r1);
		if (P + 1 > xe) break;
		step(P + 1, f1, f2, f3, f0, r1, r0);
		if (P + 2 > xe) break;
		step(P + 2, f2, f3, f0, f1, r0, r1);
		if (P + 3 > xe) break;
		step(P + 3, f3, f0, f1, f2, r1, r0);
	}
}

// ---- 3-D Cahn-Hilliard, two steps per pass (temporal blocks along the march) --
// Periodic grids. One pass over a (TY x TZ) tile and planes [xb, xe) advances
// two explicit steps: c^{n+2} = c^{n+1} + dt*M*lap(mu^{n+1}), c^{n+1} = c^n +
// dt*M*lap(mu^n), mu^k = mu(c^k) - kappa*lap(c^k) (cahn-hilliard.cpp:69-76,
// twice). c^{n+2} on the tile needs mu^{n+1} on a 1-cell in-plane halo, c^{n+1}
// on 2, mu^n on 3 and c^n on 4, each on the axis-0 planes either side. Each
// thread owns fixed points of the mu^n region (the tile and a 3-cell halo); one
// iteration of the march, at plane P, computes
//   mu^n(P)          at the own points                (region R3)
//   c^{n+1}(P-1)     at the own points in R2 (2-cell halo)
//   mu^{n+1}(P-2)    at the own points in R1 (1-cell halo)
//   c^{n+2}(P-3)     at the own points of the tile, written to HBM.
// Axial neighbours come from per-point register queues (c^n at P-1, P, P+1 and
// P+2 in flight; mu^n, c^{n+1}, mu^{n+1} at their three planes); in-plane
// neighbours through LDS planes. The planes of mu^n, c^{n+1} and mu^{n+1} are
// double-buffered: an iteration reads the planes the previous one wrote, so
// one barrier per iteration orders everything; the c^n plane is read in the
// iteration that writes it, so it is double-buffered too (PHI2) or takes a
// second barrier. Every value is the one the one-step kernels compute for that
// node and step (the halo recomputes what the neighbouring tile computes), so
// the result is bitwise two single steps; c^{n+1} never goes to HBM.
// UNIT: every grid weight 1/dx^2 and the gradient coefficient are 1, so the
// multiplications by them are left out (they are identities: same bits).
template <int PREC, int MU, int TY, int TZ, int NT, bool PHI2, bool UNIT>
__device__ __forceinline__ void ch_tile_rq2(const typename Form<PREC>::S* __restrict__ a,
                                            typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                            int y0, int z0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	static_assert(MU == 0 || MU == 1, "the two-step pass takes the deterministic chemical potentials");
	constexpr int P4Y = TY + 8, P4Z = TZ + 8;  // c^n plane tile (4-cell halo)
	constexpr int R3Y = TY + 6, R3Z = TZ + 6;  // mu^n: the threads' own points
	constexpr int R2Y = TY + 4, R2Z = TZ + 4;  // c^{n+1}
	constexpr int R1Y = TY + 2, R1Z = TZ + 2;  // mu^{n+1}
	constexpr int PN = P4Y * P4Z, N3 = R3Y * R3Z, N2 = R2Y * R2Z, N1 = R1Y * R1Z;
	constexpr int RN = PN - N3;  // c^n tile points outside R3 (a one-cell ring)
	constexpr int M = (N3 + NT - 1) / NT;
	constexpr int MR = (RN + NT - 1) / NT;
	constexpr int NPHI = PHI2 ? 2 : 1;
	__shared__ S lphi[NPHI][PN];
	__shared__ S lmu0[2][N3];
	__shared__ S lc1[2][N2];
	__shared__ S lmu1[2][N1];

	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const int64_t plane = (int64_t)ny * nz;

	// own point k: R3 index e = tid + k*NT at (r, c) of R3, node (y0-3+r, z0-3+c);
	// q4 its index in the c^n tile, q2 / q1 in R2 / R1 (valid where the flags say)
	int ooff[M], q4[M], q2[M], q1[M], fl[M];  // fl: 1 in R3, 2 in R2, 4 in R1, 8 written
#pragma unroll
	for (int k = 0; k < M; k++) {
		const int e = threadIdx.x + k * NT;
		ooff[k] = q4[k] = q2[k] = q1[k] = fl[k] = 0;
		if (e < N3) {
			const int r = e / R3Z, c = e - (e / R3Z) * R3Z;
			const int y = y0 - 3 + r, z = z0 - 3 + c;
			ooff[k] = wrap(y, ny) * nz + wrap(z, nz);
			q4[k] = (r + 1) * P4Z + (c + 1);
			int f = 1;
			if (r >= 1 && r < R2Y + 1 && c >= 1 && c < R2Z + 1) {
				f |= 2;
				q2[k] = (r - 1) * R2Z + (c - 1);
			}
			if (r >= 2 && r < R1Y + 2 && c >= 2 && c < R1Z + 2) {
				f |= 4;
				q1[k] = (r - 2) * R1Z + (c - 2);
			}
			if (r >= 3 && r < TY + 3 && c >= 3 && c < TZ + 3 && y < ny && z < nz) f |= 8;
			fl[k] = f;
		}
	}
	// ring points: rows 0 and P4Y-1 and columns 0 and P4Z-1 of the c^n tile
	int roff[MR], rq[MR];
#pragma unroll
	for (int j = 0; j < MR; j++) {
		const int t = threadIdx.x + j * NT;
		roff[j] = rq[j] = 0;
		if (t < RN) {
			int pr, pc;
			if (t < P4Z) {
				pr = 0;
				pc = t;
			} else if (t < 2 * P4Z) {
				pr = P4Y - 1;
				pc = t - P4Z;
			} else {
				const int u = t - 2 * P4Z;
				pr = 1 + u / 2;
				pc = (u & 1) ? P4Z - 1 : 0;
			}
			roff[j] = wrap(y0 - 4 + pr, ny) * nz + wrap(z0 - 4 + pc, nz);
			rq[j] = pr * P4Z + pc;
		}
	}
	auto lapadd = [&](typename F::L& lap, int ax, S yh, S y, S yl) {
		if constexpr (UNIT)
			F::lap_add1(lap, yh, y, yl);
		else
			F::lap_add(lap, g.w[ax], g.wf[ax], yh, y, yl);
	};
	auto mu = [&](S c, typename F::L lap) -> S {
		if constexpr (UNIT && MU == 1) return F::mu_dw1(c, lap);
		else if constexpr (UNIT) return F::mu_poly1(c, lap, p);
		else return mu_of<PREC, MU>(c, lap, p, 0.0f);
	};
	// unconditional loads (lanes past the region load offset 0; planes wrap)
	auto load_own = [&](int x, S* v) {
		const S* __restrict__ src = a + (int64_t)wrap(x, nx) * plane;
#pragma unroll
		for (int k = 0; k < M; k++) v[k] = src[ooff[k]];
	};
	auto load_ring = [&](int x, S* v) {
		const S* __restrict__ src = a + (int64_t)wrap(x, nx) * plane;
#pragma unroll
		for (int j = 0; j < MR; j++) v[j] = src[roff[j]];
	};

	S f0[M], f1[M], f2[M], f3[M];   // c^n queue: roles rotate over the four unrolled iterations
	S m0m[M], m0c[M], m0p[M];        // mu^n at P-2, P-1, P
	S c1m[M], c1c[M], c1p[M];        // c^{n+1} at P-3, P-2, P-1
	S m1m[M], m1c[M], m1p[M];        // mu^{n+1} at P-4, P-3, P-2
	S r0[MR], r1[MR];
#pragma unroll
	for (int k = 0; k < M; k++) {
		f3[k] = m0m[k] = m0c[k] = m0p[k] = 0;
		c1m[k] = c1c[k] = c1p[k] = m1m[k] = m1c[k] = m1p[k] = 0;
	}
#pragma unroll
	for (int j = 0; j < MR; j++) r1[j] = 0;
	// the previous tile of a persistent loop may still read the planes
	__syncthreads();
	const int x0 = xb - 3;  // the first iteration: mu^n at xb-3 needs c^n from xb-4
	load_own(x0 - 1, f0);
	load_own(x0, f1);
	load_own(x0 + 1, f2);
	load_ring(x0, r0);

	auto iter = [&](int P, const S* fm, const S* fc, const S* fp, S* fn, const S* rc, S* rn) {
		load_own(P + 2, fn);
		load_ring(P + 1, rn);
		S* __restrict__ lp = lphi[PHI2 ? (P & 1) : 0];
#pragma unroll
		for (int k = 0; k < M; k++)
			if (fl[k]) lp[q4[k]] = fc[k];
#pragma unroll
		for (int j = 0; j < MR; j++)
			if (threadIdx.x + j * NT < RN) lp[rq[j]] = rc[j];
		__syncthreads();
		// mu^n(P) at every own point
		S* __restrict__ w0 = lmu0[P & 1];
#pragma unroll
		for (int k = 0; k < M; k++) {
			if (fl[k]) {
				const int q = q4[k];
				const S y = fc[k];
				typename F::L lap = F::lap_zero();
				lapadd(lap, 0, fp[k], y, fm[k]);
				lapadd(lap, 1, lp[q + P4Z], y, lp[q - P4Z]);
				lapadd(lap, 2, lp[q + 1], y, lp[q - 1]);
				m0p[k] = mu(y, lap);
				w0[threadIdx.x + k * NT] = m0p[k];
			}
		}
		if (!PHI2) __syncthreads();  // the c^n plane is rewritten by the next iteration
		// c^{n+1}(P-1) at the own points in R2: in-plane mu^n(P-1) from the
		// plane the previous iteration wrote
		if (P - 1 >= xb - 2) {
			const S* __restrict__ r0p = lmu0[(P - 1) & 1];
			S* __restrict__ w1 = lc1[(P - 1) & 1];
#pragma unroll
			for (int k = 0; k < M; k++) {
				if (fl[k] & 2) {
					const int e = threadIdx.x + k * NT;
					const S m = m0c[k];
					typename F::L lap = F::lap_zero();
					lapadd(lap, 0, m0p[k], m, m0m[k]);
					lapadd(lap, 1, r0p[e + R3Z], m, r0p[e - R3Z]);
					lapadd(lap, 2, r0p[e + 1], m, r0p[e - 1]);
					c1p[k] = F::ch(fm[k], lap, p);
					w1[q2[k]] = c1p[k];
				}
			}
		}
		// mu^{n+1}(P-2) at the own points in R1: in-plane c^{n+1}(P-2)
		if (P - 2 >= xb - 1) {
			const S* __restrict__ r1p = lc1[(P - 2) & 1];
			S* __restrict__ w2 = lmu1[(P - 2) & 1];
#pragma unroll
			for (int k = 0; k < M; k++) {
				if (fl[k] & 4) {
					const int q = q2[k];
					const S y = c1c[k];
					typename F::L lap = F::lap_zero();
					lapadd(lap, 0, c1p[k], y, c1m[k]);
					lapadd(lap, 1, r1p[q + R2Z], y, r1p[q - R2Z]);
					lapadd(lap, 2, r1p[q + 1], y, r1p[q - 1]);
					m1p[k] = mu(y, lap);
					w2[q1[k]] = m1p[k];
				}
			}
		}
		// c^{n+2}(P-3) at the own points of the tile: in-plane mu^{n+1}(P-3)
		if (P - 3 >= xb) {
			const S* __restrict__ r2p = lmu1[(P - 3) & 1];
			S* __restrict__ dst = b + (int64_t)(P - 3) * plane;
#pragma unroll
			for (int k = 0; k < M; k++) {
				if (fl[k] & 8) {
					const int q = q1[k];
					const S m = m1c[k];
					typename F::L lap = F::lap_zero();
					lapadd(lap, 0, m1p[k], m, m1m[k]);
					lapadd(lap, 1, r2p[q + R1Z], m, r2p[q - R1Z]);
					lapadd(lap, 2, r2p[q + 1], m, r2p[q - 1]);
					dst[ooff[k]] = F::ch(c1m[k], lap, p);
				}
			}
		}
#pragma unroll
		for (int k = 0; k < M; k++) {
			m0m[k] = m0c[k];
			m0c[k] = m0p[k];
			c1m[k] = c1c[k];
			c1c[k] = c1p[k];
			m1m[k] = m1c[k];
			m1c[k] = m1p[k];
		}
	};
	const int xlast = xe + 2;  // c^{n+2} at xe-1 is written in the iteration at xe+2
	for (int P = x0; P <= xlast; P += 4) {
		iter(P, f0, f1, f2, f3, r0, r1);
		if (P + 1 > xlast) break;
		iter(P + 1, f1, f2, f3, f0, r1, r0);
		if (P + 2 > xlast) break;
		iter(P + 2, f2, f3, f0, f1, r0, r1);
		if (P + 3 > xlast) break;
		iter(P + 3, f3, f0, f1, f2, r1, r0);
	}
}

// ---- 3-D Allen-Cahn, model A, grain growth, Zener pinning: plane march -------
// Any boundary condition. One (TY x TZ) tile of the (axis 1, axis 2) plane
// marches over the axis-0 planes [xb, xe). Each thread owns fixed nodes of the
// tile and keeps their Q fields at planes P-1, P and P+1 (P+2 in flight) in
// registers; the in-plane neighbours come from plane P's tile in LDS, which
// has a one-cell halo filled through check_boundary's mapping
// (MMSP.utility.cpp:42-57): the halo cell at y = -1 holds the node the
// reference's laplacian reads there, and the axis-0 planes are mapped the same
// way. The LDS tile is double-buffered, so one barrier per plane orders its
// writes and reads. The per-node arithmetic is the generic kernels' (the same
// bits). KIND 0 Allen-Cahn (allen-cahn.cpp:68-71), 1 grain growth
// (graingrowth.cpp:91-107), 2 model A (model_A.cpp:97-101), 3 Zener pinning
// (zener.cpp:117-135: field 0 held), 4 multi-phase grain growth
// (grain_growth/anisotropic/phase_field/graingrowth.cpp:108-188: the fields
// present among the node and its six neighbours, then gga_node; fewer than
// two present, the node is copied). Ostwald ripening (ostwald.cpp:87-122) in
// two marches, the generic kernels' two passes: KIND 5 forms w of field 0
// into `wout` (one scalar plane per node) and writes the order parameters
// (fields 1..Q-1) to b; KIND 6 marches over w (a = w, Q = 1) and writes
// field 0 of b, C + dt*D*lap(w), reading C at the node from `cin`.
template <int PREC, int Q, int KIND, int TY, int TZ, int NT>
__device__ __forceinline__ void march3d(const typename Form<PREC>::S* __restrict__ a,
                                        typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                        int y0, int z0, int xb, int xe,
                                        const typename Form<PREC>::S* __restrict__ cin = nullptr,
                                        typename Form<PREC>::S* __restrict__ wout = nullptr)
{
	using F = Form<PREC>;
	using S = typename F::S;
	constexpr int PY = TY + 2, PZ = TZ + 2;
	constexpr int PN = PY * PZ, ON = TY * TZ, RN = PN - ON;
	constexpr int M = (ON + NT - 1) / NT;
	constexpr int MR = (RN + NT - 1) / NT;
	__shared__ S lt[2][Q][PN];

	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const int64_t plane = (int64_t)ny * nz;
	const int64_t N = (int64_t)nx * plane;
	// a coordinate one cell outside the grid maps as check_boundary maps it;
	// tile positions further out are never read as neighbours (offset 0)
	auto map1 = [&](int v, int n, int lo, int hi) { return (v >= -1 && v <= n) ? bc_map(v, n, lo, hi) : 0; };

	int ooff[M], oq[M];
	bool oout[M];
#pragma unroll
	for (int k = 0; k < M; k++) {
		const int e = threadIdx.x + k * NT;
		ooff[k] = oq[k] = 0;
		oout[k] = false;
		if (e < ON) {
			const int r = e / TZ, c = e - (e / TZ) * TZ;
			const int y = y0 + r, z = z0 + c;
			ooff[k] = map1(y, ny, g.b0[1], g.b1[1]) * nz + map1(z, nz, g.b0[2], g.b1[2]);
			oq[k] = (r + 1) * PZ + (c + 1);
			oout[k] = y < ny && z < nz;
		}
	}
	int roff[MR], rq[MR];
#pragma unroll
	for (int j = 0; j < MR; j++) {
		const int t = threadIdx.x + j * NT;
		roff[j] = rq[j] = 0;
		if (t < RN) {
			int pr, pc;
			if (t < PZ) {
				pr = 0;
				pc = t;
			} else if (t < 2 * PZ) {
				pr = PY - 1;
				pc = t - PZ;
			} else {
				const int u = t - 2 * PZ;
				pr = 1 + u / 2;
				pc = (u & 1) ? PZ - 1 : 0;
			}
			roff[j] = map1(y0 - 1 + pr, ny, g.b0[1], g.b1[1]) * nz + map1(z0 - 1 + pc, nz, g.b0[2], g.b1[2]);
			rq[j] = pr * PZ + pc;
		}
	}
	auto plane_of = [&](int x) { return (int64_t)map1(x, nx, g.b0[0], g.b1[0]) * plane; };
	auto load_own = [&](int x, S (*v)[M]) {
		const S* __restrict__ src = a + plane_of(x);
#pragma unroll
		for (int q = 0; q < Q; q++)
#pragma unroll
			for (int k = 0; k < M; k++) v[q][k] = src[q * N + ooff[k]];
	};
	auto load_ring = [&](int x, S (*v)[MR]) {
		const S* __restrict__ src = a + plane_of(x);
#pragma unroll
		for (int q = 0; q < Q; q++)
#pragma unroll
			for (int j = 0; j < MR; j++) v[q][j] = src[q * N + roff[j]];
	};

	S f0[Q][M], f1[Q][M], f2[Q][M], f3[Q][M];
	S r0[Q][MR], r1[Q][MR];
	int nkx = -1;
	float nkeep[KIND == 2 ? M : 1];
	(void)nkx;
	(void)nkeep;
	// the previous tile of a persistent loop may still read the planes
	__syncthreads();
	load_own(xb - 1, f0);
	load_own(xb, f1);
	load_own(xb + 1, f2);
	load_ring(xb, r0);

	auto step = [&](int P, S (*fm)[M], S (*fc)[M], S (*fp)[M], S (*fn)[M], S (*rc)[MR], S (*rn)[MR]) {
		load_own(P + 2, fn);
		load_ring(P + 1, rn);
		S(*lp)[PN] = lt[P & 1];
#pragma unroll
		for (int q = 0; q < Q; q++) {
#pragma unroll
			for (int k = 0; k < M; k++)
				if (threadIdx.x + k * NT < ON) lp[q][oq[k]] = fc[q][k];
#pragma unroll
			for (int j = 0; j < MR; j++)
				if (threadIdx.x + j * NT < RN) lp[q][rq[j]] = rc[q][j];
		}
		__syncthreads();
		S* __restrict__ dst = b + (int64_t)P * plane;
		const bool nhit = P == nkx;
		(void)nhit;
#pragma unroll
		for (int k = 0; k < M; k++) {
			if (!oout[k]) continue;
			const int qq = oq[k];
			typename F::Sum sum = 0;
			if constexpr (KIND == 1 || KIND == 3) {
#pragma unroll
				for (int q = 0; q < Q; q++) sum += fc[q][k] * fc[q][k];
			}
			float gd = 0.0f;
			if constexpr (KIND == 2) gd = normal_deviate_cached(p.seed, p.step, P, (uint32_t)ooff[k], nkx, nkeep[k]);
			if constexpr (KIND == 5) {
				// Ostwald pass 1: sum_j phi_j^2 over the order parameters, w
				// from lap C, then each order parameter's local update
				typename F::OSum osum = 0;
#pragma unroll
				for (int q = 1; q < Q; q++) osum += fc[q][k] * fc[q][k];
				const S C = fc[0][k];
#pragma unroll
				for (int q = 0; q < Q; q++) {
					const S y = fc[q][k];
					typename F::L lap = F::lap_zero();
					F::lap_add(lap, g.w[0], g.wf[0], fp[q][k], y, fm[q][k]);
					F::lap_add(lap, g.w[1], g.wf[1], lp[q][qq + PZ], y, lp[q][qq - PZ]);
					F::lap_add(lap, g.w[2], g.wf[2], lp[q][qq + 1], y, lp[q][qq - 1]);
					if (q == 0) wout[(int64_t)P * plane + ooff[k]] = F::ow_w(C, osum, lap, p);
					else dst[q * N + ooff[k]] = F::ow_phi(y, C, osum, lap, p);
				}
				continue;
			}
			if constexpr (KIND == 6) {
				// Ostwald pass 2: C' = C + dt*D*lap(w)
				const S y = fc[0][k];
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], fp[0][k], y, fm[0][k]);
				F::lap_add(lap, g.w[1], g.wf[1], lp[0][qq + PZ], y, lp[0][qq - PZ]);
				F::lap_add(lap, g.w[2], g.wf[2], lp[0][qq + 1], y, lp[0][qq - 1]);
				dst[ooff[k]] = F::ow_c(cin[(int64_t)P * plane + ooff[k]], lap, p);
				continue;
			}
			if constexpr (KIND == 4) {
				// the active set and the laplacians of its fields, axis by axis
				// as the generic kernel forms them
				S phi[Q], lap[Q];
				uint32_t act = 0;
				S cnt = 0;
#pragma unroll
				for (int q = 0; q < Q; q++) {
					const S y = fc[q][k];
					const S y1h = lp[q][qq + PZ], y1l = lp[q][qq - PZ], y2h = lp[q][qq + 1], y2l = lp[q][qq - 1];
					phi[q] = y;
					const bool on = y > S(0) || fp[q][k] > S(0) || fm[q][k] > S(0) || y1h > S(0) || y1l > S(0) ||
					                y2h > S(0) || y2l > S(0);
					typename F::L l = F::lap_zero();
					if (on) {
						act |= 1u << q;
						cnt += S(1);
						F::lap_add(l, g.w[0], g.wf[0], fp[q][k], y, fm[q][k]);
						F::lap_add(l, g.w[1], g.wf[1], y1h, y, y1l);
						F::lap_add(l, g.w[2], g.wf[2], y2h, y, y2l);
					}
					lap[q] = l;
				}
				if (cnt >= S(2)) gga_node<PREC, Q>(phi, lap, act, cnt, Q, p);
#pragma unroll
				for (int q = 0; q < Q; q++) dst[q * N + ooff[k]] = phi[q];
				continue;
			}
#pragma unroll
			for (int q = 0; q < Q; q++) {
				const S y = fc[q][k];
				if (KIND == 3 && q == 0) {
					dst[ooff[k]] = y;  // the particles keep their values
					continue;
				}
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], fp[q][k], y, fm[q][k]);
				F::lap_add(lap, g.w[1], g.wf[1], lp[q][qq + PZ], y, lp[q][qq - PZ]);
				F::lap_add(lap, g.w[2], g.wf[2], lp[q][qq + 1], y, lp[q][qq - 1]);
				S out;
				if constexpr (KIND == 0) out = F::ac(y, lap, p);
				else if constexpr (KIND == 2) out = F::ac_noise(y, lap, gd, p);
				else out = F::gg(y, sum, lap, p);
				dst[q * N + ooff[k]] = out;
			}
		}
		if constexpr (KIND == 2) nkx = nhit ? -1 : (P ^ 1);
	};
	for (int P = xb; P < xe; P += 4) {
		step(P, f0, f1, f2, f3, r0, r1);
		if (P + 1 >= xe) break;
		step(P + 1, f1, f2, f3, f0, r1, r0);
		if (P + 2 >= xe) break;
		step(P + 2, f2, f3, f0, f1, r0, r1);
		if (P + 3 >= xe) break;
		step(P + 3, f3, f0, f1, f2, r1, r0);
	}
}

// ---- 3-D Allen-Cahn, grain growth, Zener pinning: two steps per pass -------
// Periodic grids. The plane march of march3d advancing two steps per pass:
// each thread owns fixed points of the tile and a one-cell in-plane halo
// (region R1) and keeps, for each of them, the Q fields of phi^n at planes
// P-1, P, P+1 (P+2 in flight) and of phi^{n+1} at P-2, P-1, P in registers.
// The iteration at plane P publishes phi^n(P) on a plane tile with a two-cell
// halo and forms phi^{n+1}(P) on R1 (published too), then phi^{n+2}(P-1) on
// the tile from the phi^{n+1} plane the previous iteration published; both LDS
// planes are double-buffered, so one barrier per iteration orders everything.
// phi^{n+1} never goes to HBM. Every value is the one-step kernels' for that
// node and step (the halo recomputes what the neighbouring tile computes), so
// the result is bitwise two single steps. KIND 0 Allen-Cahn, 1 grain growth,
// 3 Zener pinning (field 0 held).
template <int PREC, int Q, int KIND, int TY, int TZ, int NT>
__device__ __forceinline__ void march3d2(const typename Form<PREC>::S* __restrict__ a,
                                         typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                         int y0, int z0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	static_assert(KIND == 0 || KIND == 1 || KIND == 3, "the two-step march takes the deterministic models");
	constexpr int P2Y = TY + 4, P2Z = TZ + 4;  // phi^n plane tile (two-cell halo)
	constexpr int R1Y = TY + 2, R1Z = TZ + 2;  // phi^{n+1}: the threads' own points
	constexpr int PN = P2Y * P2Z, N1 = R1Y * R1Z, RN = PN - N1;
	constexpr int M = (N1 + NT - 1) / NT;
	constexpr int MR = (RN + NT - 1) / NT;
	__shared__ S lphi[2][Q * PN];
	__shared__ S lq[2][Q * N1];

	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const int64_t plane = (int64_t)ny * nz;
	const int64_t N = (int64_t)nx * plane;

	// own point k: e = tid + k*NT names R1 point (r, c), node (y0-1+r,
	// z0-1+c): e < R1Y*TZ column e % TZ of row e / TZ, the next 2*R1Y the last
	// two columns (TZ, TZ+1) of each row, so a wave's points lie in whole row
	// segments (contiguous in LDS and in HBM); q2 the point's index in the
	// phi^n tile, i1 in the R1 plane; fl 1 in R1, 2 in the tile (written)
	int ooff[M], q2[M], i1[M], fl[M];
#pragma unroll
	for (int k = 0; k < M; k++) {
		const int e = threadIdx.x + k * NT;
		ooff[k] = q2[k] = i1[k] = fl[k] = 0;
		if (e < N1) {
			int r, c;
			if (e < R1Y * TZ) {
				r = e / TZ;
				c = e - r * TZ;
			} else {
				r = (e - R1Y * TZ) >> 1;
				c = TZ + ((e - R1Y * TZ) & 1);
			}
			const int y = y0 - 1 + r, z = z0 - 1 + c;
			ooff[k] = wrap(y, ny) * nz + wrap(z, nz);
			q2[k] = (r + 1) * P2Z + (c + 1);
			i1[k] = r * R1Z + c;
			fl[k] = 1 | ((r >= 1 && r < TY + 1 && c >= 1 && c < TZ + 1 && y < ny && z < nz) ? 2 : 0);
		}
	}
	// ring points: rows 0 and P2Y-1 and columns 0 and P2Z-1 of the phi^n tile
	int roff[MR], rq[MR];
#pragma unroll
	for (int j = 0; j < MR; j++) {
		const int t = threadIdx.x + j * NT;
		roff[j] = rq[j] = 0;
		if (t < RN) {
			int pr, pc;
			if (t < P2Z) {
				pr = 0;
				pc = t;
			} else if (t < 2 * P2Z) {
				pr = P2Y - 1;
				pc = t - P2Z;
			} else {
				const int u = t - 2 * P2Z;
				pr = 1 + u / 2;
				pc = (u & 1) ? P2Z - 1 : 0;
			}
			roff[j] = wrap(y0 - 2 + pr, ny) * nz + wrap(z0 - 2 + pc, nz);
			rq[j] = pr * P2Z + pc;
		}
	}
	auto load_own = [&](int x, S (*v)[M]) {
		const S* __restrict__ src = a + (int64_t)wrap(x, nx) * plane;
#pragma unroll
		for (int q = 0; q < Q; q++)
#pragma unroll
			for (int k = 0; k < M; k++) v[q][k] = src[q * N + ooff[k]];
	};
	auto load_ring = [&](int x, S (*v)[MR]) {
		const S* __restrict__ src = a + (int64_t)wrap(x, nx) * plane;
#pragma unroll
		for (int q = 0; q < Q; q++)
#pragma unroll
			for (int j = 0; j < MR; j++) v[q][j] = src[q * N + roff[j]];
	};
	// the update of point k from its values at the planes before (fm), at
	// (fc) and after (fp), its in-plane neighbours at index i of the LDS plane
	// l (field stride FS, row stride RS); field q's new value
	auto upd = [&](const S (*fm)[M], const S (*fc)[M], const S (*fp)[M], const S* l, int FS, int RS, int i, int k,
	               typename F::Sum sum, int q) -> S {
		const S y = fc[q][k];
		if (KIND == 3 && q == 0) return y;  // the particles keep their values
		typename F::L lap = F::lap_zero();
		F::lap_add(lap, g.w[0], g.wf[0], fp[q][k], y, fm[q][k]);
		F::lap_add(lap, g.w[1], g.wf[1], l[q * FS + i + RS], y, l[q * FS + i - RS]);
		F::lap_add(lap, g.w[2], g.wf[2], l[q * FS + i + 1], y, l[q * FS + i - 1]);
		if constexpr (KIND == 0) return F::ac(y, lap, p);
		else return F::gg(y, sum, lap, p);
	};
	auto sumsq = [&](const S (*v)[M], int k) {
		typename F::Sum sum = 0;
		if constexpr (KIND != 0) {
#pragma unroll
			for (int q = 0; q < Q; q++) sum += v[q][k] * v[q][k];
		}
		return sum;
	};

	S f0[Q][M], f1[Q][M], f2[Q][M], f3[Q][M];  // phi^n queue: roles rotate over four unrolled iterations
	S qm[Q][M], qc[Q][M], qp[Q][M];            // phi^{n+1} at P-2, P-1, P
	S r0[Q][MR], r1[Q][MR];
#pragma unroll
	for (int q = 0; q < Q; q++) {
#pragma unroll
		for (int k = 0; k < M; k++) f3[q][k] = qm[q][k] = qc[q][k] = qp[q][k] = 0;
#pragma unroll
		for (int j = 0; j < MR; j++) r1[q][j] = 0;
	}
	// the previous tile of a persistent loop may still read the planes
	__syncthreads();
	const int x0 = xb - 1;  // the first iteration forms phi^{n+1}(xb-1)
	load_own(x0 - 1, f0);
	load_own(x0, f1);
	load_own(x0 + 1, f2);
	load_ring(x0, r0);

	auto iter = [&](int P, const S (*fm)[M], const S (*fc)[M], const S (*fp)[M], S (*fn)[M], const S (*rc)[MR],
	                S (*rn)[MR]) {
		load_own(P + 2, fn);
		load_ring(P + 1, rn);
		S* __restrict__ lp = lphi[P & 1];
#pragma unroll
		for (int q = 0; q < Q; q++) {
#pragma unroll
			for (int k = 0; k < M; k++)
				if (fl[k]) lp[q * PN + q2[k]] = fc[q][k];
#pragma unroll
			for (int j = 0; j < MR; j++)
				if (threadIdx.x + j * NT < RN) lp[q * PN + rq[j]] = rc[q][j];
		}
		__syncthreads();
		// phi^{n+1}(P) at every own point, published for the next iteration
		S* __restrict__ lw = lq[P & 1];
#pragma unroll
		for (int k = 0; k < M; k++) {
			if (fl[k]) {
				const typename F::Sum sum = sumsq(fc, k);
#pragma unroll
				for (int q = 0; q < Q; q++) {
					qp[q][k] = upd(fm, fc, fp, lp, PN, P2Z, q2[k], k, sum, q);
					lw[q * N1 + i1[k]] = qp[q][k];
				}
			}
		}
		// phi^{n+2}(P-1) at the own points of the tile: in-plane phi^{n+1}(P-1)
		if (P - 1 >= xb) {
			const S* __restrict__ lr = lq[(P - 1) & 1];
			S* __restrict__ dst = b + (int64_t)(P - 1) * plane;
#pragma unroll
			for (int k = 0; k < M; k++) {
				if (fl[k] & 2) {
					const typename F::Sum sum = sumsq(qc, k);
#pragma unroll
					for (int q = 0; q < Q; q++)
						dst[q * N + ooff[k]] = upd(qm, qc, qp, lr, N1, R1Z, i1[k], k, sum, q);
				}
			}
		}
#pragma unroll
		for (int q = 0; q < Q; q++)
#pragma unroll
			for (int k = 0; k < M; k++) {
				qm[q][k] = qc[q][k];
				qc[q][k] = qp[q][k];
			}
	};
	const int xlast = xe;  // phi^{n+2} at xe-1 is written in the iteration at xe
	for (int P = x0; P <= xlast; P += 4) {
		iter(P, f0, f1, f2, f3, r0, r1);
		if (P + 1 > xlast) break;
		iter(P + 1, f1, f2, f3, f0, r1, r0);
		if (P + 2 > xlast) break;
		iter(P + 2, f2, f3, f0, f1, r0, r1);
		if (P + 3 > xlast) break;
		iter(P + 3, f3, f0, f1, f2, r1, r0);
	}
}

// ---- 3-D Ostwald ripening in one pass (periodic) ----------------------------
// The two passes of the generic kernels (ostwald.cpp:87-122: w and the order
// parameters, then C + dt*D*lap(w)) fused into one plane march. Each thread
// owns fixed points of the tile and its one-cell in-plane halo (region R1)
// with the Q fields at planes P-1, P, P+1 (P+2 in flight) in registers, and
// w at planes P-2, P-1, P. The iteration at plane P publishes the fields of
// plane P to LDS with a two-cell halo, forms w(P) on R1 and publishes it to a
// second LDS plane, writes the order parameters of plane P on the tile, then C
// of plane P-1 on the tile from w (planes P-2 and P from registers, in-plane
// from the plane the previous iteration published). Both planes are
// double-buffered: one barrier per iteration. w never goes to HBM; every value
// is the two passes' for that node (the halo recomputes the w the
// neighbouring tile computes), so the result is bitwise theirs.
template <int PREC, int Q, int TY, int TZ, int NT>
__device__ __forceinline__ void march3d_ow(const typename Form<PREC>::S* __restrict__ a,
                                           typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                           int y0, int z0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	constexpr int P2Y = TY + 4, P2Z = TZ + 4;  // the fields' plane tile (two-cell halo)
	constexpr int R1Y = TY + 2, R1Z = TZ + 2;  // w: the threads' own points
	constexpr int PN = P2Y * P2Z, N1 = R1Y * R1Z, RN = PN - N1;
	constexpr int M = (N1 + NT - 1) / NT;
	constexpr int MR = (RN + NT - 1) / NT;
	__shared__ S lphi[2][Q * PN];
	__shared__ S lw[2][N1];

	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const int64_t plane = (int64_t)ny * nz;
	const int64_t N = (int64_t)nx * plane;

	// own point k: e = tid + k*NT names R1 point (r, c), node (y0-1+r,
	// z0-1+c): e < R1Y*TZ column e % TZ of row e / TZ, the next 2*R1Y the last
	// two columns (TZ, TZ+1) of each row, so a wave's points lie in whole row
	// segments (contiguous in LDS and in HBM); q2 the point's index in the
	// fields' tile, i1 in the R1 plane; fl 1 in R1, 2 in the tile (written)
	int ooff[M], q2[M], i1[M], fl[M];
#pragma unroll
	for (int k = 0; k < M; k++) {
		const int e = threadIdx.x + k * NT;
		ooff[k] = q2[k] = i1[k] = fl[k] = 0;
		if (e < N1) {
			int r, c;
			if (e < R1Y * TZ) {
				r = e / TZ;
				c = e - r * TZ;
			} else {
				r = (e - R1Y * TZ) >> 1;
				c = TZ + ((e - R1Y * TZ) & 1);
			}
			const int y = y0 - 1 + r, z = z0 - 1 + c;
			ooff[k] = wrap(y, ny) * nz + wrap(z, nz);
			q2[k] = (r + 1) * P2Z + (c + 1);
			i1[k] = r * R1Z + c;
			fl[k] = 1 | ((r >= 1 && r < TY + 1 && c >= 1 && c < TZ + 1 && y < ny && z < nz) ? 2 : 0);
		}
	}
	// ring points: rows 0 and P2Y-1 and columns 0 and P2Z-1 of the fields' tile
	int roff[MR], rq[MR];
#pragma unroll
	for (int j = 0; j < MR; j++) {
		const int t = threadIdx.x + j * NT;
		roff[j] = rq[j] = 0;
		if (t < RN) {
			int pr, pc;
			if (t < P2Z) {
				pr = 0;
				pc = t;
			} else if (t < 2 * P2Z) {
				pr = P2Y - 1;
				pc = t - P2Z;
			} else {
				const int u = t - 2 * P2Z;
				pr = 1 + u / 2;
				pc = (u & 1) ? P2Z - 1 : 0;
			}
			roff[j] = wrap(y0 - 2 + pr, ny) * nz + wrap(z0 - 2 + pc, nz);
			rq[j] = pr * P2Z + pc;
		}
	}
	auto load_own = [&](int x, S (*v)[M]) {
		const S* __restrict__ src = a + (int64_t)wrap(x, nx) * plane;
#pragma unroll
		for (int q = 0; q < Q; q++)
#pragma unroll
			for (int k = 0; k < M; k++) v[q][k] = src[q * N + ooff[k]];
	};
	auto load_ring = [&](int x, S (*v)[MR]) {
		const S* __restrict__ src = a + (int64_t)wrap(x, nx) * plane;
#pragma unroll
		for (int q = 0; q < Q; q++)
#pragma unroll
			for (int j = 0; j < MR; j++) v[q][j] = src[q * N + roff[j]];
	};

	S f0[Q][M], f1[Q][M], f2[Q][M], f3[Q][M];  // the fields' queue: roles rotate over four unrolled iterations
	S wm[M], wc[M], wp[M];                     // w at P-2, P-1, P
	S r0[Q][MR], r1[Q][MR];
#pragma unroll
	for (int k = 0; k < M; k++) {
		wm[k] = wc[k] = wp[k] = 0;
#pragma unroll
		for (int q = 0; q < Q; q++) f3[q][k] = 0;
	}
#pragma unroll
	for (int q = 0; q < Q; q++)
#pragma unroll
		for (int j = 0; j < MR; j++) r1[q][j] = 0;
	// the previous tile of a persistent loop may still read the planes
	__syncthreads();
	const int x0 = xb - 1;  // the first iteration forms w(xb-1)
	load_own(x0 - 1, f0);
	load_own(x0, f1);
	load_own(x0 + 1, f2);
	load_ring(x0, r0);

	auto iter = [&](int P, const S (*fm)[M], const S (*fc)[M], const S (*fp)[M], S (*fn)[M], const S (*rc)[MR],
	                S (*rn)[MR]) {
		load_own(P + 2, fn);
		load_ring(P + 1, rn);
		S* __restrict__ lp = lphi[P & 1];
#pragma unroll
		for (int q = 0; q < Q; q++) {
#pragma unroll
			for (int k = 0; k < M; k++)
				if (fl[k]) lp[q * PN + q2[k]] = fc[q][k];
#pragma unroll
			for (int j = 0; j < MR; j++)
				if (threadIdx.x + j * NT < RN) lp[q * PN + rq[j]] = rc[q][j];
		}
		__syncthreads();
		// w(P) at every own point (published), the order parameters of plane P
		// on the tile
		S* __restrict__ lwp = lw[P & 1];
		const bool inside = P >= xb && P < xe;
		S* __restrict__ dst = b + (int64_t)P * plane;
#pragma unroll
		for (int k = 0; k < M; k++) {
			if (!fl[k]) continue;
			const int i = q2[k];
			typename F::OSum osum = 0;
#pragma unroll
			for (int q = 1; q < Q; q++) osum += fc[q][k] * fc[q][k];
			const S C = fc[0][k];
			{
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], fp[0][k], C, fm[0][k]);
				F::lap_add(lap, g.w[1], g.wf[1], lp[i + P2Z], C, lp[i - P2Z]);
				F::lap_add(lap, g.w[2], g.wf[2], lp[i + 1], C, lp[i - 1]);
				wp[k] = F::ow_w(C, osum, lap, p);
				lwp[i1[k]] = wp[k];
			}
			if (inside && (fl[k] & 2)) {
#pragma unroll
				for (int q = 1; q < Q; q++) {
					const S y = fc[q][k];
					typename F::L lap = F::lap_zero();
					F::lap_add(lap, g.w[0], g.wf[0], fp[q][k], y, fm[q][k]);
					F::lap_add(lap, g.w[1], g.wf[1], lp[q * PN + i + P2Z], y, lp[q * PN + i - P2Z]);
					F::lap_add(lap, g.w[2], g.wf[2], lp[q * PN + i + 1], y, lp[q * PN + i - 1]);
					dst[q * N + ooff[k]] = F::ow_phi(y, C, osum, lap, p);
				}
			}
		}
		// C of plane P-1 on the tile: C + dt*D*lap(w)
		if (P - 1 >= xb) {
			const S* __restrict__ lr = lw[(P - 1) & 1];
			S* __restrict__ dc = b + (int64_t)(P - 1) * plane;
#pragma unroll
			for (int k = 0; k < M; k++) {
				if (!(fl[k] & 2)) continue;
				const int e = i1[k];
				const S y = wc[k];
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], wp[k], y, wm[k]);
				F::lap_add(lap, g.w[1], g.wf[1], lr[e + R1Z], y, lr[e - R1Z]);
				F::lap_add(lap, g.w[2], g.wf[2], lr[e + 1], y, lr[e - 1]);
				dc[ooff[k]] = F::ow_c(fm[0][k], lap, p);
			}
		}
#pragma unroll
		for (int k = 0; k < M; k++) {
			wm[k] = wc[k];
			wc[k] = wp[k];
		}
	};
	for (int P = x0; P <= xe; P += 4) {
		iter(P, f0, f1, f2, f3, r0, r1);
		if (P + 1 > xe) break;
		iter(P + 1, f1, f2, f3, f0, r1, r0);
		if (P + 2 > xe) break;
		iter(P + 2, f2, f3, f0, f1, r0, r1);
		if (P + 3 > xe) break;
		iter(P + 3, f3, f0, f1, f2, r1, r0);
	}
}

// ---- 2-D Cahn-Hilliard, wave strips (periodic): one wave per (strip, chunk) --
// Lane l holds column c0 - 2 + l: mu is formed on lanes 1..62 and c' written
// on lanes 2..61 (kChWaveOut columns per strip); column neighbours come from
// the adjacent lanes by DPP, row neighbours from a register window: phi at
// rows x, x+d, x+2d (the next rows in flight) and mu at x-d, x, x+d while row
// x is written. The same per-node arithmetic as the two-pass form
// (cahn-hilliard.cpp:69-76), so the same bits; no LDS and no barriers. Odd
// chunks march downward (d = -1) so that two neighbouring chunks read their
// shared boundary rows at about the same time.
template <int PREC, int MU, bool DOWN>
__device__ __forceinline__ void ch_wave_dir(const typename Form<PREC>::S* __restrict__ a,
                                            typename Form<PREC>::S* __restrict__ b, const Geom& g,
                                            const ModelParams& p, int c0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	constexpr int d = DOWN ? -1 : 1;
	const int lane = threadIdx.x & 63;
	const int nx = g.n[0], ny = g.n[1];
	const int c = c0 - 2 + lane;
	const unsigned cm = (unsigned)tile_map(c, ny, g.b0[1], g.b1[1]);
	const bool out = lane >= 2 && lane < 2 + kChWaveOut && c < ny;
	const bool edges = g.b0[0] != MMSPX_PERIODIC || g.b1[0] != MMSPX_PERIODIC || g.b0[1] != MMSPX_PERIODIC ||
	                   g.b1[1] != MMSPX_PERIODIC;
	auto ld = [&](int x) { return a[(size_t)tile_map(x, nx, g.b0[0], g.b1[0]) * ny + cm]; };
	// mu at row r from phi at r-d (pb), r (pc), r+d (pa); rows come in
	// marching order, so noise (MU 2) keeps the partner row's deviate
	int nkx = -1;
	float nkeep = 0.0f;
	auto mu = [&](int r, S pb, S pc, S pa) {
		const S zl = from_prev(pc, pc), zh = from_next(pc, pc);
		typename F::L lap = F::lap_zero();
		F::lap_add(lap, g.w[0], g.wf[0], DOWN ? pb : pa, pc, DOWN ? pa : pb);
		F::lap_add(lap, g.w[1], g.wf[1], zh, pc, zl);
		float gd = 0.0f;
		if constexpr (MU == 2) {
			const int xw = wrap(r, nx);
			const bool hit = xw == nkx;
			gd = normal_deviate_cached(p.seed, p.step, xw, cm, nkx, nkeep);
			nkx = hit ? -1 : (xw ^ 1);
		}
		return mu_of<PREC, MU>(pc, lap, p, gd);
	};
	const int x0 = DOWN ? xe - 1 : xb, x1 = DOWN ? xb - 1 : xe;
	const int n = xe - xb;  // rows written
	// phi rows in a ring of eight registers: at the row x written at position
	// u of the unrolled march, slot (u + k) & 7 holds phi(x + k d) for k = 0..6
	// (rows 3..6 still in flight) and receives row x + 7d; the slots change
	// roles instead of being copied, so a row's loads are waited for only
	// where the row is first used. A row is loaded only if the chunk reads it.
	constexpr int RING = 8;
	S f[RING];
	const S fm2 = ld(x0 - 2 * d), fm1 = ld(x0 - d);
	f[0] = ld(x0);
	f[1] = ld(x0 + d);
	f[2] = ld(x0 + 2 * d);
#pragma unroll
	for (int k = 0; k < 4; k++) f[3 + k] = (k + 1 < n) ? ld(x0 + (3 + k) * d) : S(0);
	f[RING - 1] = 0;
	S mb = mu(x0 - d, fm2, fm1, f[0]);  // mu(x0 - d)
	S mc = mu(x0, fm1, f[0], f[1]);     // mu(x0)
	for (int x = x0, i = 0;;) {
#pragma unroll
		for (int u = 0; u < RING; u++, i++) {
			if (x == x1) return;
			// row x + 7d becomes phi(x'+2d) at row x' = x + 5d
			if (i + 5 < n) f[(u + RING - 1) % RING] = ld(x + (RING - 1) * d);
			const S f0 = f[u], f1 = f[(u + 1) % RING], f2 = f[(u + 2) % RING];
			const S ma = mu(x + d, f0, f1, f2);  // mu(x+d)
			S zl = from_prev(mc, mc), zh = from_next(mc, mc);
			S xh = DOWN ? mb : ma, xl = DOWN ? ma : mb;
			if (edges) {
				edge_mu(xl, xh, mc, x, nx, g.b0[0], g.b1[0]);
				edge_mu(zl, zh, mc, c, ny, g.b0[1], g.b1[1]);
			}
			typename F::L lap = F::lap_zero();
			F::lap_add(lap, g.w[0], g.wf[0], xh, mc, xl);
			F::lap_add(lap, g.w[1], g.wf[1], zh, mc, zl);
			if (out) b[(size_t)x * ny + c] = F::ch(f0, lap, p);
			mb = mc;
			mc = ma;
			x += d;
		}
	}
}

template <int PREC, int MU>
__device__ __forceinline__ void ch_wave(const typename Form<PREC>::S* __restrict__ a,
                                        typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                        int c0, int xb, int xe, bool down)
{
	if (down)
		ch_wave_dir<PREC, MU, true>(a, b, g, p, c0, xb, xe);
	else
		ch_wave_dir<PREC, MU, false>(a, b, g, p, c0, xb, xe);
}

// ---- 2-D Ostwald ripening, wave strips (periodic): one wave per (strip, chunk)
// Lane l holds column c0 - 2 + l, as the Cahn-Hilliard wave strips: w is
// formed on lanes 1..62 and the new values written on lanes 2..61. Rows: the
// concentration C at x .. x+2 (x+3 in flight) for w(x+1), the order
// parameters at x-1 .. x+1 (x+2 in flight), w at x-1 .. x+1, and the sums
// sum_j phi_j^2 of rows x and x+1. Row x's new C is C + dt*D*lap(w), its order
// parameters follow their local update: the arithmetic of ow_pass1/ow_pass2
// (ostwald.cpp:87-122), value by value, so the same bits.
template <int PREC, int Q>
__device__ __forceinline__ void ow_wave(const typename Form<PREC>::S* __restrict__ a,
                                        typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                        int c0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	constexpr int R = Q - 1;  // order parameters
	const int lane = threadIdx.x & 63;
	const int nx = g.n[0], ny = g.n[1];
	const size_t N = (size_t)nx * ny;
	const int c = c0 - 2 + lane;
	const unsigned cm = (unsigned)wrap(c, ny);
	const bool out = lane >= 2 && lane < 2 + kChWaveOut && c < ny;
	auto row = [&](int x) { return (size_t)wrap(x, nx) * ny + cm; };
	auto sum_of = [&](const S* ph) {
		typename F::OSum s = 0;
#pragma unroll
		for (int j = 0; j < R; j++) s += ph[j] * ph[j];
		return s;
	};
	// w at row r from C at r-1 (cb), r (cc), r+1 (ca) and the sum at r
	auto w_at = [&](S cb, S cc, S ca, typename F::OSum sum) {
		const S zl = from_prev(cc, cc), zh = from_next(cc, cc);
		typename F::L lap = F::lap_zero();
		F::lap_add(lap, g.w[0], g.wf[0], ca, cc, cb);
		F::lap_add(lap, g.w[1], g.wf[1], zh, cc, zl);
		return F::ow_w(cc, sum, lap, p);
	};
	S c0r = a[row(xb - 2)], c1r = a[row(xb - 1)];
	S cx = a[row(xb)], cx1 = a[row(xb + 1)], cx2 = a[row(xb + 2)];
	// five-slot rings for C (rows x .. x+2, x+3 in flight, x+4 loading) and
	// phi (rows x-1 .. x+1, x+2 in flight, x+3 loading) whose roles rotate over
	// a five-row unrolled march, so no row is copied while its loads are in
	// flight
	S C3 = (xb + 1 < xe) ? a[row(xb + 3)] : S(0), C4 = 0;
	S F0[R], F1[R], F2[R], F3[R], F4[R];
#pragma unroll
	for (int j = 0; j < R; j++) {
		const S* __restrict__ aj = a + (size_t)(j + 1) * N;
		F0[j] = aj[row(xb - 1)];
		F1[j] = aj[row(xb)];
		F2[j] = aj[row(xb + 1)];
		F3[j] = (xb + 1 < xe) ? aj[row(xb + 2)] : S(0);
		F4[j] = 0;
	}
	typename F::OSum sc = sum_of(F1);
	S wm = w_at(c0r, c1r, cx, sum_of(F0));  // w(x-1)
	S wc = w_at(c1r, cx, cx1, sc);          // w(x)
	// row x: C at x, x+1, x+2 (cx0, cx1, cx2), C(x+4) to cn; phi at x-1, x,
	// x+1 (fm, fc, fp), phi(x+3) to fn
	auto step = [&](int x, S cx0, S cx1_, S cx2_, S& cn, const S* fm, const S* fc, const S* fp, S* fn) {
		// loads for the next rows go out first
		if (x + 2 < xe) {
			cn = a[row(x + 4)];
#pragma unroll
			for (int j = 0; j < R; j++) fn[j] = (a + (size_t)(j + 1) * N)[row(x + 3)];
		}
		const typename F::OSum sp = sum_of(fp);
		const S wp = w_at(cx0, cx1_, cx2_, sp);  // w(x+1)
		// C'(x) = C + dt*D*lap(w)
		{
			const S zl = from_prev(wc, wc), zh = from_next(wc, wc);
			typename F::L lap = F::lap_zero();
			F::lap_add(lap, g.w[0], g.wf[0], wp, wc, wm);
			F::lap_add(lap, g.w[1], g.wf[1], zh, wc, zl);
			if (out) b[(size_t)x * ny + c] = F::ow_c(cx0, lap, p);
		}
#pragma unroll
		for (int j = 0; j < R; j++) {
			const S y = fc[j];
			const S yl = from_prev(y, y), yh = from_next(y, y);
			typename F::L lap = F::lap_zero();
			F::lap_add(lap, g.w[0], g.wf[0], fp[j], y, fm[j]);
			F::lap_add(lap, g.w[1], g.wf[1], yh, y, yl);
			if (out) b[(size_t)(j + 1) * N + (size_t)x * ny + c] = F::ow_phi(y, cx0, sc, lap, p);
		}
		wm = wc;
		wc = wp;
		sc = sp;
	};
	S C0 = cx, C1 = cx1, C2 = cx2;
	for (int x = xb; x < xe; x += 5) {
		step(x, C0, C1, C2, C4, F0, F1, F2, F4);
		if (x + 1 >= xe) break;
		step(x + 1, C1, C2, C3, C0, F1, F2, F3, F0);
		if (x + 2 >= xe) break;
		step(x + 2, C2, C3, C4, C1, F2, F3, F4, F1);
		if (x + 3 >= xe) break;
		step(x + 3, C3, C4, C0, C2, F3, F4, F0, F2);
		if (x + 4 >= xe) break;
		step(x + 4, C4, C0, C1, C3, F4, F0, F1, F3);
	}
}

// ---- 2-D Allen-Cahn, temporal blocks (periodic): K <= KMAX steps per pass ----
// A TX x kTbCols tile and its K-cell halo (at most 64 columns: one wave per
// row) are read once into LDS; the block then takes K steps in LDS, the valid
// region shrinking by one cell per step (cells near the rim are recomputed by
// the neighbouring tiles), and writes the tile. Every value at every step is
// computed from its four neighbours at the previous step in the reference's
// order (allen-cahn.cpp:68-71), so the result is bitwise that of K single
// steps, while HBM sees one read and one write per cell per K steps. Rows and
// columns past the grid's end hold wrapped nodes (consistent periodic images);
// only stores are masked.
template <int PREC, int TX, int KMAX, int NT>
__device__ __forceinline__ void ac_tblock(const typename Form<PREC>::S* __restrict__ a,
                                          typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                          int x0, int y0, int K)
{
	using F = Form<PREC>;
	using S = typename F::S;
	constexpr int TY = kTbCols;
	static_assert(TY + 2 * KMAX <= 64, "a region row must fit one wave");
	constexpr int NW = NT / 64;
	__shared__ S buf[2][TX + 2 * KMAX][64];
	const int nx = g.n[0], ny = g.n[1];
	const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
	const int RX = TX + 2 * K, RY = TY + 2 * K;
	const int yc = wrap(y0 - K + l, ny);
	// the previous tile of a persistent loop may still read the buffers
	__syncthreads();
	for (int r = w; r < RX; r += NW)
		if (l < RY) buf[0][r][l] = a[(int64_t)wrap(x0 - K + r, nx) * ny + yc];
	__syncthreads();
	for (int s = 1; s <= K; s++) {
		const S(*src)[64] = buf[(s - 1) & 1];
		S(*dst)[64] = buf[s & 1];
		const bool act = l >= s && l < RY - s;
		// every lane of a wave runs each row (the column neighbours come from
		// the adjacent lanes by DPP); lanes 0 and 63 are never in the valid
		// region, so no lane needs a neighbour from outside the wave
		for (int r = s + w; r < RX - s; r += NW) {
			const S y = src[r][l];
			const S yl = from_prev(y, y), yh = from_next(y, y);
			typename F::L lap = F::lap_zero();
			F::lap_add(lap, g.w[0], g.wf[0], src[r + 1][l], y, src[r - 1][l]);
			F::lap_add(lap, g.w[1], g.wf[1], yh, y, yl);
			if (act) dst[r][l] = F::ac(y, lap, p);
		}
		__syncthreads();
	}
	const S(*res)[64] = buf[K & 1];
	const int y = y0 + l - K;
	if (l >= K && l < K + TY && y < ny)
		for (int r = w; r < TX; r += NW) {
			const int x = x0 + r;
			if (x < nx) b[(int64_t)x * ny + y] = res[r + K][l];
		}
}

// ---- 2-D strips: TZ contiguous columns [y0, y0+TZ) marching rows [xb, xe) ----
// Each thread owns one column; the axis-0 neighbours stay in registers (a
// three-row window), the axis-1 neighbours come from a double-buffered LDS
// row. Any boundary condition: rows and halo columns go through check_boundary's
// mapping. One global load and one store per cell (per field).
template <int PREC, int TZ>
__device__ __forceinline__ void ac_strip(const typename Form<PREC>::S* __restrict__ a,
                                         typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                         int y0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	__shared__ S row[2][TZ + 2];
	const int nx = g.n[0], ny = g.n[1];
	const int t = threadIdx.x;
	const int nact = min(TZ, ny - y0);
	const bool act = t < nact;
	const int y = y0 + (act ? t : nact - 1);
	const int yl = bc_map(y0 - 1, ny, g.b0[1], g.b1[1]);
	const int yr = bc_map(y0 + nact, ny, g.b0[1], g.b1[1]);
	auto at = [&](int x) { return a + (int64_t)bc_map(x, nx, g.b0[0], g.b1[0]) * ny; };
	__syncthreads();
	// three-row window prev/cur/next plus row x+2 in flight; the halo
	// columns of the current row are loaded one row ahead as well
	S prev = at(xb - 1)[y];
	S cur = a[(int64_t)xb * ny + y];
	S next = at(xb + 1)[y];
	S hl = 0, hr = 0;
	if (t == 0) hl = a[(int64_t)xb * ny + yl];
	if (t == nact - 1) hr = a[(int64_t)xb * ny + yr];
	for (int x = xb; x < xe; x++) {
		// row x+2 is needed only if row x+1 is computed here; then x+2 <= nx,
		// at most one row past the edge, which check_boundary's mapping covers
		S nn = 0;
		if (x + 1 < xe) nn = at(x + 2)[y];
		S nhl = 0, nhr = 0;
		if (x + 1 < xe) {
			if (t == 0) nhl = a[(int64_t)(x + 1) * ny + yl];
			if (t == nact - 1) nhr = a[(int64_t)(x + 1) * ny + yr];
		}
		S* lr = row[x & 1];
		if (act) lr[t + 1] = cur;
		if (t == 0) lr[0] = hl;
		if (t == nact - 1) lr[nact + 1] = hr;
		__syncthreads();
		if (act) {
			typename F::L lap = F::lap_zero();
			F::lap_add(lap, g.w[0], g.wf[0], next, cur, prev);
			F::lap_add(lap, g.w[1], g.wf[1], lr[t + 2], cur, lr[t]);
			b[(int64_t)x * ny + y] = F::ac(cur, lap, p);
		}
		prev = cur;
		cur = next;
		next = nn;
		hl = nhl;
		hr = nhr;
	}
}

// Grain growth, Q fields stored field-major (plane j at offset j*nodes). Same
// window as ac_strip per field; the sum over fields stays in registers.
template <int PREC, int Q, int TZ>
__device__ __forceinline__ void gg_strip(const typename Form<PREC>::S* __restrict__ a,
                                         typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                         int y0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	__shared__ S row[2][Q][TZ + 2];
	const int nx = g.n[0], ny = g.n[1];
	const int64_t N = (int64_t)nx * ny;
	const int t = threadIdx.x;
	const int nact = min(TZ, ny - y0);
	const bool act = t < nact;
	const int y = y0 + (act ? t : nact - 1);
	const int yl = bc_map(y0 - 1, ny, g.b0[1], g.b1[1]);
	const int yr = bc_map(y0 + nact, ny, g.b0[1], g.b1[1]);
	auto rowoff = [&](int x) { return (int64_t)bc_map(x, nx, g.b0[0], g.b1[0]) * ny; };
	S prev[Q], cur[Q], next[Q], nn[Q];
	__syncthreads();
	{
		const int64_t rp = rowoff(xb - 1) + y, rc = (int64_t)xb * ny + y, rn = rowoff(xb + 1) + y;
#pragma unroll
		for (int j = 0; j < Q; j++) {
			prev[j] = a[j * N + rp];
			cur[j] = a[j * N + rc];
			next[j] = a[j * N + rn];
		}
	}
	for (int x = xb; x < xe; x++) {
		if (x + 1 < xe) {  // row x+2 only when it will be used (x+2 <= nx)
			const int64_t rn2 = rowoff(x + 2) + y;
#pragma unroll
			for (int j = 0; j < Q; j++) nn[j] = a[j * N + rn2];
		}
		const int64_t rc = (int64_t)x * ny;
		const int s = x & 1;
#pragma unroll
		for (int j = 0; j < Q; j++) {
			if (act) row[s][j][t + 1] = cur[j];
			if (t == 0) row[s][j][0] = a[j * N + rc + yl];
			if (t == nact - 1) row[s][j][nact + 1] = a[j * N + rc + yr];
		}
		__syncthreads();
		if (act) {
			typename F::Sum sum = 0;
#pragma unroll
			for (int j = 0; j < Q; j++) sum += cur[j] * cur[j];
#pragma unroll
			for (int j = 0; j < Q; j++) {
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], next[j], cur[j], prev[j]);
				F::lap_add(lap, g.w[1], g.wf[1], row[s][j][t + 2], cur[j], row[s][j][t]);
				b[j * N + rc + y] = F::gg(cur[j], sum, lap, p);
			}
		}
#pragma unroll
		for (int j = 0; j < Q; j++) {
			prev[j] = cur[j];
			cur[j] = next[j];
			next[j] = nn[j];
		}
	}
}

// ---- grain growth, 2-D tile of TX rows (axis 0) x TC columns (axis 1) ---------
// Each thread owns CPT cells. It loads the cells' Q field values once (all
// loads in flight together), forms sum_j phi_j^2 in the reference's order,
// then, field by field, publishes its values into a double-buffered LDS tile
// whose one-cell ring is loaded (one field ahead) by the threads in turn; the
// laplacian reads the tile, the update uses the registers. One global read
// and one write per field value (the ring reads hit in L2). Neighbours outside
// the grid come through check_boundary's mapping; positions more than one
// cell outside are never read by a cell of the grid and load a dummy in-range
// value.
__device__ __forceinline__ int edge_map(int x, int n, int b0, int b1)
{
	if (x >= -1 && x <= n) return bc_map(x, n, b0, b1);
	return x < 0 ? 0 : n - 1;
}

template <int PREC, int Q, int TX, int TC, int NT>
__device__ __forceinline__ void gg_tile(const typename Form<PREC>::S* __restrict__ a,
                                        typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                        int x0, int y0)
{
	using F = Form<PREC>;
	using S = typename F::S;
	static_assert((TX * TC) % NT == 0, "tile cells must divide evenly among threads");
	constexpr int CPT = TX * TC / NT;
	constexpr int PC = TC + 2, PN = (TX + 2) * PC;
	constexpr int RN = 2 * PC + 2 * TX;
	constexpr int RREG = (RN + NT - 1) / NT;
	__shared__ S tile[2][PN];
	const int nx = g.n[0], ny = g.n[1];
	const int64_t N = (int64_t)nx * ny;

	int node[CPT], q[CPT];
	bool own[CPT];
#pragma unroll
	for (int k = 0; k < CPT; k++) {
		const int e = threadIdx.x + k * NT;
		const int r = e / TC, c = e - (e / TC) * TC;
		const int x = x0 + r, y = y0 + c;
		own[k] = x < nx && y < ny;
		node[k] = edge_map(x, nx, g.b0[0], g.b1[0]) * ny + edge_map(y, ny, g.b0[1], g.b1[1]);
		q[k] = (r + 1) * PC + (c + 1);
	}
	int rnode[RREG], rq[RREG];
#pragma unroll
	for (int j = 0; j < RREG; j++) {
		const int t = threadIdx.x + j * NT;
		rnode[j] = 0;
		rq[j] = 0;
		if (t < RN) {
			int pr, pc;
			if (t < PC) {
				pr = 0;
				pc = t;
			} else if (t < 2 * PC) {
				pr = TX + 1;
				pc = t - PC;
			} else {
				const int u = t - 2 * PC;
				pr = 1 + u / 2;
				pc = (u & 1) ? PC - 1 : 0;
			}
			rnode[j] = edge_map(x0 - 1 + pr, nx, g.b0[0], g.b1[0]) * ny + edge_map(y0 - 1 + pc, ny, g.b0[1], g.b1[1]);
			rq[j] = pr * PC + pc;
		}
	}
	S v[CPT][Q];
#pragma unroll
	for (int f = 0; f < Q; f++)
#pragma unroll
		for (int k = 0; k < CPT; k++) v[k][f] = a[f * N + node[k]];
	typename F::Sum sum[CPT];
#pragma unroll
	for (int k = 0; k < CPT; k++) {
		sum[k] = 0;
#pragma unroll
		for (int f = 0; f < Q; f++) sum[k] += v[k][f] * v[k][f];
	}
	S ring[RREG], rnext[RREG];
#pragma unroll
	for (int j = 0; j < RREG; j++) ring[j] = threadIdx.x + j * NT < RN ? a[rnode[j]] : S(0);
#pragma unroll
	for (int f = 0; f < Q; f++) {
		if (f + 1 < Q) {
#pragma unroll
			for (int j = 0; j < RREG; j++)
				if (threadIdx.x + j * NT < RN) rnext[j] = a[(f + 1) * N + rnode[j]];
		}
		S* __restrict__ t = tile[f & 1];
#pragma unroll
		for (int k = 0; k < CPT; k++) t[q[k]] = v[k][f];
#pragma unroll
		for (int j = 0; j < RREG; j++)
			if (threadIdx.x + j * NT < RN) t[rq[j]] = ring[j];
		__syncthreads();
#pragma unroll
		for (int k = 0; k < CPT; k++) {
			if (own[k]) {
				const S y = v[k][f];
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], t[q[k] + PC], y, t[q[k] - PC]);
				F::lap_add(lap, g.w[1], g.wf[1], t[q[k] + 1], y, t[q[k] - 1]);
				b[f * N + node[k]] = F::gg(y, sum[k], lap, p);
			}
		}
#pragma unroll
		for (int j = 0; j < RREG; j++) ring[j] = rnext[j];
	}
}

// ---- 2-D wave strips: one wave per (strip, chunk) item, any boundary ----------
// Lane l holds column c0 + l of a 64-column strip (c0 a multiple of 64, so
// every row segment a wave reads or writes is two whole 128-byte lines).
// Axis-1 neighbours move between lanes by DPP wave shifts. The two columns
// beside the strip, for all Q fields of a row, fit one register: lane k < Q
// loads field k of the column left of the strip, lane Q + k field k of the
// column right of it, one row ahead like the rest; lanes 0 and 63 take them
// from there by readlane. Axis-0 neighbours stay in registers (rows x-1, x,
// x+1, with row x+2 in flight). Neighbours outside the grid come through
// check_boundary's mapping. No LDS and no barriers: each wave runs on its
// own. KIND 1, grain growth: Q fields stored field-major, sum_j phi_j^2
// formed in the reference's order before the field updates
// (graingrowth.cpp:91-107); KIND 0, Allen-Cahn (Q = 1, allen-cahn.cpp:68-71);
// KIND 2, model A (Allen-Cahn plus noise, model_A.cpp:97-101); KIND 3, Zener
// pinning (grain growth with field 0 held, zener.cpp:117-135); KIND 4,
// multi-phase grain growth (grain_growth/anisotropic/phase_field/
// graingrowth.cpp:108-188: the fields present among the node and its four
// neighbours, then gga_node; fewer than two present, the node is copied).
template <int PREC, int Q, int KIND>
__device__ __forceinline__ void strip_wave(const typename Form<PREC>::S* __restrict__ a,
                                           typename Form<PREC>::S* __restrict__ b, const Geom& g, const ModelParams& p,
                                           int c0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	static_assert(2 * Q <= 64, "the beside-the-strip columns of all fields must fit one register");
	const int lane = threadIdx.x & 63;
	const int nx = g.n[0], ny = g.n[1];
	const size_t N = (size_t)nx * ny;
	const int c = c0 + lane;
	const bool out = c < ny;
	// the lane's own column; lanes past the grid's end hold mapped columns:
	// the first of them, column ny, is the last column's right neighbour
	const unsigned cm = (unsigned)edge_map(c, ny, g.b0[1], g.b1[1]);
	// beside-the-strip loads: lane k < 2Q, field k mod Q, left side for k < Q
	const bool hl = lane < 2 * Q;
	const unsigned ch = (unsigned)edge_map(lane < Q ? c0 - 1 : c0 + 64, ny, g.b0[1], g.b1[1]);
	const S* __restrict__ ah = a + (size_t)(hl ? lane % Q : 0) * N;
	// offsets inside one field plane (fewer than 2^31 nodes)
	auto rrow = [&](int x) { return (unsigned)edge_map(x, nx, g.b0[0], g.b1[0]) * (unsigned)ny; };
	// four row slots (all Q fields of a row) whose roles (x-1, x, x+1, x+2 in
	// flight) rotate over a four-row unrolled march, so no row is copied
	// while its loads are in flight and each field's loads are waited for
	// only where that field is first used; the beside-the-strip register
	// likewise
	S wa[Q], wb[Q], wc[Q], wd[Q];
	S ha = 0, hb = 0, hc = 0, hd = 0;
	int nkx = -1;  // model A: the partner row of the kept deviate
	float nkeep = 0.0f;
	(void)nkx;
	(void)nkeep;
	// Row loads are unconditional: past the chunk's last needed row (xe) that
	// row is loaded again, and lanes at or past 2Q load field 0 beside the
	// strip, unused; every address is a mapped node of the grid.
	auto load_row = [&](int x, S* w, S& h) {
		const unsigned r = rrow(min(x, xe));
#pragma unroll
		for (int j = 0; j < Q; j++) w[j] = (a + j * N)[r + cm];
		h = ah[r + ch];
	};
#pragma unroll
	for (int j = 0; j < Q; j++) wd[j] = 0;
	load_row(xb - 1, wa, ha);
	load_row(xb, wb, hb);
	load_row(xb + 1, wc, hc);
	// row x from rows x-1 (w0), x (w1, beside-the-strip h1) and x+1 (w2);
	// row x+2 goes to w3 and h3
	auto row = [&](int x, const S* w0, const S* w1, const S* w2, S* w3, const S h1, S& h3) {
		// (the multi-phase march, with short chunks, skips the load it does not need)
		if (KIND != 4 || x + 1 < xe) load_row(x + 2, w3, h3);
		typename F::Sum sum = 0;
		if (KIND == 1 || KIND == 3) {
#pragma unroll
			for (int j = 0; j < Q; j++) sum += w1[j] * w1[j];
		}
		const unsigned o = (unsigned)x * (unsigned)ny + (unsigned)c;
		float gd = 0.0f;
		if constexpr (KIND == 2) {
			const bool hit = x == nkx;
			if (out) gd = normal_deviate_cached(p.seed, p.step, x, (uint32_t)c, nkx, nkeep);
			nkx = hit ? -1 : (x ^ 1);
		}
		if constexpr (KIND == 4) {
			S phi[Q], lap[Q];
			uint32_t act = 0;
			S cnt = 0;
#pragma unroll
			for (int j = 0; j < Q; j++) {
				const S y = w1[j];
				const S yl = from_prev(y, lane_value(h1, j)), yh = from_next(y, lane_value(h1, Q + j));
				phi[j] = y;
				const bool on = y > S(0) || w2[j] > S(0) || w0[j] > S(0) || yh > S(0) || yl > S(0);
				typename F::L l = F::lap_zero();
				if (on) {
					act |= 1u << j;
					cnt += S(1);
					F::lap_add(l, g.w[0], g.wf[0], w2[j], y, w0[j]);
					F::lap_add(l, g.w[1], g.wf[1], yh, y, yl);
				}
				lap[j] = l;
			}
			if (cnt >= S(2)) gga_node<PREC, Q>(phi, lap, act, cnt, Q, p);
			if (out) {
#pragma unroll
				for (int j = 0; j < Q; j++) (b + j * N)[o] = phi[j];
			}
		} else {
#pragma unroll
			for (int j = 0; j < Q; j++) {
				const S y = w1[j];
				const S yl = from_prev(y, lane_value(h1, j)), yh = from_next(y, lane_value(h1, Q + j));
				typename F::L lap = F::lap_zero();
				F::lap_add(lap, g.w[0], g.wf[0], w2[j], y, w0[j]);
				F::lap_add(lap, g.w[1], g.wf[1], yh, y, yl);
				if (out)
					(b + j * N)[o] = KIND == 1 ? F::gg(y, sum, lap, p)
					                 : KIND == 2 ? F::ac_noise(y, lap, gd, p)
					                 : KIND == 3 ? (j == 0 ? y : F::gg(y, sum, lap, p))
					                             : F::ac(y, lap, p);
			}
		}
	};
	if constexpr (KIND == 4) {
		// the multi-phase body is large: four copies of it overflow the
		// instruction cache, so one row per iteration and the slots advance
		// by copies
		for (int x = xb; x < xe; x++) {
			row(x, wa, wb, wc, wd, hb, hd);
#pragma unroll
			for (int j = 0; j < Q; j++) {
				wa[j] = wb[j];
				wb[j] = wc[j];
				wc[j] = wd[j];
			}
			hb = hc;
			hc = hd;
		}
	} else {
		for (int x = xb; x < xe; x += 4) {
			row(x, wa, wb, wc, wd, hb, hd);
			if (x + 1 >= xe) break;
			row(x + 1, wb, wc, wd, wa, hc, ha);
			if (x + 2 >= xe) break;
			row(x + 2, wc, wd, wa, wb, hd, hb);
			if (x + 3 >= xe) break;
			row(x + 3, wd, wa, wb, wc, ha, hc);
		}
	}
}

// ---- 2-D grain growth, split wave strips: W waves per (strip, chunk) --------
// The strip_wave layout (lane l on column c0 + l of a 64-column strip, row
// window in registers, column neighbours by DPP), with the Q fields split
// between the W (2 or 4) waves of a workgroup: wave h holds fields
// h*Q/W .. (h+1)*Q/W - 1, a W-th of the registers, so more waves fit a SIMD.
// The sum of phi_j^2 runs in the reference's field order across the waves:
// wave 0 sums its fields from zero and passes the partial sum through LDS,
// each next wave continues it over its fields, and the last passes the total
// back (W barriers per row). KIND 1: grain growth (graingrowth.cpp:91-107);
// KIND 3: Zener pinning (field 0 held, zener.cpp:117-135).
template <int PREC, int Q, int KIND, int W>
__device__ __forceinline__ void strip_wave_split(const typename Form<PREC>::S* __restrict__ a,
                                                 typename Form<PREC>::S* __restrict__ b, const Geom& g,
                                                 const ModelParams& p, int c0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	using Sum = typename F::Sum;
	constexpr int H = Q / W;
	static_assert(Q % W == 0 && 2 * H <= 64, "the split strips hold Q/W fields per wave");
	__shared__ Sum part[64], total[64];
	const int lane = threadIdx.x & 63;
	const int wv = (int)(threadIdx.x >> 6);  // the wave's share of the fields, 0 .. W-1
	const int f0 = wv * H;
	const int nx = g.n[0], ny = g.n[1];
	const size_t N = (size_t)nx * ny;
	const int c = c0 + lane;
	const bool out = c < ny;
	const unsigned cm = (unsigned)edge_map(c, ny, g.b0[1], g.b1[1]);
	// beside-the-strip loads: lane k < H field f0 + k left of the strip, lane
	// H + k the same field right of it; other lanes load field f0, unused
	const unsigned ch = (unsigned)edge_map(lane < H ? c0 - 1 : c0 + 64, ny, g.b0[1], g.b1[1]);
	const S* __restrict__ ah = a + (size_t)(f0 + (lane < 2 * H ? lane % H : 0)) * N;
	const S* __restrict__ af = a + (size_t)f0 * N;
	S* __restrict__ bf = b + (size_t)f0 * N;
	auto rrow = [&](int x) { return (unsigned)edge_map(x, nx, g.b0[0], g.b1[0]) * (unsigned)ny; };
	S wa[H], wb[H], wc[H], wd[H];
	S ha = 0, hb = 0, hc = 0, hd = 0;
	auto load_row = [&](int x, S* w, S& h) {
		const unsigned r = rrow(min(x, xe));
#pragma unroll
		for (int j = 0; j < H; j++) w[j] = (af + j * N)[r + cm];
		h = ah[r + ch];
	};
#pragma unroll
	for (int j = 0; j < H; j++) wd[j] = 0;
	load_row(xb - 1, wa, ha);
	load_row(xb, wb, hb);
	load_row(xb + 1, wc, hc);
	auto row = [&](int x, const S* w0, const S* w1, const S* w2, S* w3, const S h1, S& h3) {
		load_row(x + 2, w3, h3);
		// sum_j phi_j^2 of row x in field order: wave 0 from zero over its
		// fields, each next wave on from there; all then hold the total
		Sum acc = 0;
		if (wv == 0) {
#pragma unroll
			for (int j = 0; j < H; j++) acc += w1[j] * w1[j];
			part[lane] = acc;
		}
#pragma unroll
		for (int k = 1; k < W; k++) {
			__syncthreads();
			if (wv == k) {
				acc = part[lane];
#pragma unroll
				for (int j = 0; j < H; j++) acc += w1[j] * w1[j];
				if (k + 1 < W)
					part[lane] = acc;
				else
					total[lane] = acc;
			}
		}
		__syncthreads();
		if (wv + 1 < W) acc = total[lane];
		const unsigned o = (unsigned)x * (unsigned)ny + (unsigned)c;
#pragma unroll
		for (int j = 0; j < H; j++) {
			const S y = w1[j];
			const S yl = from_prev(y, lane_value(h1, j)), yh = from_next(y, lane_value(h1, H + j));
			typename F::L lap = F::lap_zero();
			F::lap_add(lap, g.w[0], g.wf[0], w2[j], y, w0[j]);
			F::lap_add(lap, g.w[1], g.wf[1], yh, y, yl);
			if (out) (bf + j * N)[o] = (KIND == 3 && f0 + j == 0) ? y : F::gg(y, acc, lap, p);
		}
	};
	for (int x = xb; x < xe; x += 4) {
		row(x, wa, wb, wc, wd, hb, hd);
		if (x + 1 >= xe) break;
		row(x + 1, wb, wc, wd, wa, hc, ha);
		if (x + 2 >= xe) break;
		row(x + 2, wc, wd, wa, wb, hd, hb);
		if (x + 3 >= xe) break;
		row(x + 3, wd, wa, wb, wc, ha, hc);
	}
}

// ---- 2-D grain growth, two steps per pass: split wave strips (periodic) ------
// The split strips' field split (wave h of W holds fields h*Q/W ..), with the
// Cahn-Hilliard wave strips' lane layout: lane l holds column c0 - 2 + l
// (wrapped), phi^{n+1} is formed on lanes 1..62 and phi^{n+2} written on lanes
// 2..61 (kChWaveOut columns per strip), column neighbours by DPP. Rows: phi^n
// in a four-slot ring (x .. x+2, x+3 in flight) and phi^{n+1} in another
// (x-1 .. x+1), both rotated over a four-row unrolled march; phi^{n+1} never
// goes to memory. An iteration at row x first forms sum_j phi_j^2 of phi^n at
// row x+1 and of phi^{n+1} at row x in one pass of the field-order chain
// across the waves (W barriers, as the one-step strips take per row; none
// when one wave holds all the fields, W = 1), then phi^{n+1}(x+1) and
// phi^{n+2}(x). Every value is the one-step kernels' for
// that node and step (graingrowth.cpp:91-107; KIND 3 holds field 0,
// zener.cpp:117-135), so the result is bitwise two single steps. Periodic
// grids only: a lane past the grid's edge holds the wrapped column, whose
// neighbours are its lanes' neighbours.
template <int PREC, int Q, int KIND, int W>
__device__ __forceinline__ void strip_wave_split2(const typename Form<PREC>::S* __restrict__ a,
                                                  typename Form<PREC>::S* __restrict__ b, const Geom& g,
                                                  const ModelParams& p, int c0, int xb, int xe)
{
	using F = Form<PREC>;
	using S = typename F::S;
	using Sum = typename F::Sum;
	constexpr int H = Q / W;
	static_assert(Q % W == 0, "the two-step split strips hold Q/W fields per wave");
	__shared__ Sum part[2][64], total[2][64];
	const int lane = threadIdx.x & 63;
	const int wv = (int)(threadIdx.x >> 6);
	const int f0 = wv * H;
	const int nx = g.n[0], ny = g.n[1];
	const size_t N = (size_t)nx * ny;
	const int c = c0 - 2 + lane;
	const bool out = lane >= 2 && lane < 62 && c < ny;
	const unsigned cm = (unsigned)wrap(c, ny);
	const S* __restrict__ af = a + (size_t)f0 * N;
	S* __restrict__ bf = b + (size_t)f0 * N;
	// rows past the last one the chunk reads (xe + 1) load that row again
	auto load_row = [&](int x, S* w) {
		const unsigned r = (unsigned)wrap(min(x, xe + 1), nx) * (unsigned)ny;
#pragma unroll
		for (int j = 0; j < H; j++) w[j] = (af + j * N)[r + cm];
	};
	// sum_j u_j^2 and sum_j v_j^2 in field order over all Q fields: wave 0
	// from zero over its fields, each next wave on from there
	auto sums = [&](const S* u, const S* v, Sum& su, Sum& sv) {
		Sum au = 0, av = 0;
		if constexpr (W == 1) {  // one wave holds every field: no exchange
#pragma unroll
			for (int j = 0; j < H; j++) {
				au += u[j] * u[j];
				av += v[j] * v[j];
			}
			su = au;
			sv = av;
			return;
		}
		if (wv == 0) {
#pragma unroll
			for (int j = 0; j < H; j++) {
				au += u[j] * u[j];
				av += v[j] * v[j];
			}
			part[0][lane] = au;
			part[1][lane] = av;
		}
#pragma unroll
		for (int k = 1; k < W; k++) {
			__syncthreads();
			if (wv == k) {
				au = part[0][lane];
				av = part[1][lane];
#pragma unroll
				for (int j = 0; j < H; j++) {
					au += u[j] * u[j];
					av += v[j] * v[j];
				}
				if (k + 1 < W) {
					part[0][lane] = au;
					part[1][lane] = av;
				} else {
					total[0][lane] = au;
					total[1][lane] = av;
				}
			}
		}
		__syncthreads();
		if (wv + 1 < W) {
			au = total[0][lane];
			av = total[1][lane];
		}
		su = au;
		sv = av;
	};
	// the update of row x from rows x-1 (w0), x (w1), x+1 (w2), sum s
	auto update = [&](const S* w0, const S* w1, const S* w2, Sum s, int j) -> S {
		const S y = w1[j];
		const S yl = from_prev(y, y), yh = from_next(y, y);
		typename F::L lap = F::lap_zero();
		F::lap_add(lap, g.w[0], g.wf[0], w2[j], y, w0[j]);
		F::lap_add(lap, g.w[1], g.wf[1], yh, y, yl);
		return (KIND == 3 && f0 + j == 0) ? y : F::gg(y, s, lap, p);
	};
	S n0[H], n1[H], n2[H], n3[H];  // phi^n ring
	S q0[H], q1[H], q2[H], q3[H];  // phi^{n+1} ring
#pragma unroll
	for (int j = 0; j < H; j++) n3[j] = q2[j] = q3[j] = 0;
	{
		// phi^{n+1} of rows xb-1 (q0) and xb (q1) from phi^n rows xb-2 .. xb+1
		S m2[H], m1[H];
		load_row(xb - 2, m2);
		load_row(xb - 1, m1);
		load_row(xb, n0);
		load_row(xb + 1, n1);
		load_row(xb + 2, n2);
		Sum sa, sb;
		sums(m1, n0, sa, sb);
#pragma unroll
		for (int j = 0; j < H; j++) {
			q0[j] = update(m2, m1, n0, sa, j);
			q1[j] = update(m1, n0, n1, sb, j);
		}
	}
	// row x: phi^n x (na), x+1 (nb), x+2 (nc), x+3 into nd; phi^{n+1} x-1
	// (qa), x (qb), x+1 into qc
	auto row = [&](int x, const S* na, const S* nb, const S* nc, S* nd, const S* qa, const S* qb, S* qc) {
		load_row(x + 3, nd);
		Sum s1, s2;
		sums(nb, qb, s1, s2);
#pragma unroll
		for (int j = 0; j < H; j++) qc[j] = update(na, nb, nc, s1, j);
		const unsigned o = (unsigned)x * (unsigned)ny + (unsigned)c;
#pragma unroll
		for (int j = 0; j < H; j++) {
			const S v = update(qa, qb, qc, s2, j);
			if (out) (bf + j * N)[o] = v;
		}
	};
	for (int x = xb; x < xe; x += 4) {
		row(x, n0, n1, n2, n3, q0, q1, q2);
		if (x + 1 >= xe) break;
		row(x + 1, n1, n2, n3, n0, q1, q2, q3);
		if (x + 2 >= xe) break;
		row(x + 2, n2, n3, n0, n1, q2, q3, q0);
		if (x + 3 >= xe) break;
		row(x + 3, n3, n0, n1, n2, q3, q0, q1);
	}
}

}  // namespace mmspx
