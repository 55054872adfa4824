// kernels_fused.hip -- tiled kernels for single grids: fused single-pass
// Cahn-Hilliard (2-D/3-D, periodic) and marching-strip Allen-Cahn and grain
// growth (2-D, any boundary condition). Tile bodies live in tiles.hpp.
//
// Cahn-Hilliard: the reference takes two full sweeps per step
// (cahn-hilliard.cpp:69-76): mu into a temporary grid, then c + dt*M*lap(mu).
// Unfused, that moves 40 bytes per FP64 cell-update through HBM (c in, mu out,
// mu in, c in, c' out); the fused tile moves the compulsory 16 (c in, c' out)
// and recomputes mu on a one-cell halo from the same wrapped values, so the
// result is bitwise that of the two-pass form. Only periodic boundaries take
// it; other boundaries keep the two-pass kernels, where mu at a mapped
// neighbour is mu of the mapped node exactly as in the reference.
//
// Workgroup ids are remapped so that the workgroups one XCD receives (ids are
// dealt round-robin over the 8 XCDs) form a contiguous run of tiles with the
// axis-0 chunk slowest: an XCD sweeps whole planes of one chunk, so halo reads
// of neighbouring tiles hit its own L2.
#include <algorithm>

#include "tiles.hpp"

namespace mmspx {

__device__ __forceinline__ int xcd_tile(int ntiles)
{
	int tile = blockIdx.x;
	if ((gridDim.x & 7) == 0) tile = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
	return tile < ntiles ? tile : -1;
}

// BODY 0: LDS rings, 1: register queues; MU as mu_of (0 polynomial, 1 double
// well, 2 polynomial plus noise)
template <int PREC, int MU, int DIM, int TY, int TZ, int NT, int BODY>
__global__ void __launch_bounds__(NT) ch_fused(const typename Form<PREC>::S* __restrict__ a,
                                               typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                               int chunk_len, int tiles_y, int tiles_z, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int tz = tile % tiles_z;
	const int t2 = tile / tiles_z;
	const int ty = t2 % tiles_y;
	const int chunk = t2 / tiles_y;
	const int xb = chunk * chunk_len;
	const int xe = min(xb + chunk_len, g.n[0]);
	if constexpr (BODY == 1)
		ch_tile_rq<PREC, MU, DIM, TY, TZ, NT>(a, b, g, p, ty * TY, tz * TZ, xb, xe);
	else
		ch_tile<PREC, MU, DIM, TY, TZ, NT>(a, b, g, p, ty * TY, tz * TZ, xb, xe);
}

// Two steps per pass (3-D, periodic, deterministic chemical potentials).
// FP32 on 32 x 32 tiles asks for six waves per SIMD (80 VGPRs, no spill), so
// two 768-thread workgroups share a CU.
template <int PREC, int MU, int TY, int TZ, int NT, bool PHI2, bool UNIT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu((PREC != MMSPX_F64 && TZ == 32) ? 6 : 1)))
ch_fused2(const typename Form<PREC>::S* __restrict__ a,
                                                typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                int chunk_len, int tiles_y, int tiles_z, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int tz = tile % tiles_z;
	const int t2 = tile / tiles_z;
	const int ty = t2 % tiles_y;
	const int chunk = t2 / tiles_y;
	const int xb = chunk * chunk_len;
	const int xe = min(xb + chunk_len, g.n[0]);
	ch_tile_rq2<PREC, MU, TY, TZ, NT, PHI2, UNIT>(a, b, g, p, ty * TY, tz * TZ, xb, xe);
}

// 3-D plane march (Allen-Cahn, model A, grain growth, Zener pinning): one
// workgroup per (tile, chunk), XCD-contiguous runs as the Cahn-Hilliard tiles.
template <int PREC, int Q, int KIND, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT) march3d_k(const typename Form<PREC>::S* __restrict__ a,
                                                typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                int chunk_len, int tiles_y, int tiles_z, int ntiles,
                                                const typename Form<PREC>::S* __restrict__ cin,
                                                typename Form<PREC>::S* __restrict__ wout)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int tz = tile % tiles_z;
	const int t2 = tile / tiles_z;
	const int ty = t2 % tiles_y;
	const int chunk = t2 / tiles_y;
	const int xb = chunk * chunk_len;
	march3d<PREC, Q, KIND, TY, TZ, NT>(a, b, g, p, ty * TY, tz * TZ, xb, min(xb + chunk_len, g.n[0]), cin, wout);
}

// The plane march two steps per pass (periodic Allen-Cahn, grain growth,
// Zener pinning): one workgroup per (tile, chunk), as march3d_k.
template <int PREC, int Q, int KIND, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT) march3d2_k(const typename Form<PREC>::S* __restrict__ a,
                                                 typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                 int chunk_len, int tiles_y, int tiles_z, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int tz = tile % tiles_z;
	const int t2 = tile / tiles_z;
	const int ty = t2 % tiles_y;
	const int chunk = t2 / tiles_y;
	const int xb = chunk * chunk_len;
	march3d2<PREC, Q, KIND, TY, TZ, NT>(a, b, g, p, ty * TY, tz * TZ, xb, min(xb + chunk_len, g.n[0]));
}

// 3-D Ostwald ripening in one pass (periodic): one workgroup per (tile, chunk).
template <int PREC, int Q, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT) march3d_ow_k(const typename Form<PREC>::S* __restrict__ a,
                                                   typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                   int chunk_len, int tiles_y, int tiles_z, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int tz = tile % tiles_z;
	const int t2 = tile / tiles_z;
	const int ty = t2 % tiles_y;
	const int chunk = t2 / tiles_y;
	const int xb = chunk * chunk_len;
	march3d_ow<PREC, Q, TY, TZ, NT>(a, b, g, p, ty * TY, tz * TZ, xb, min(xb + chunk_len, g.n[0]));
}

template <int PREC, int TZ>
__global__ void __launch_bounds__(TZ) ac_tiled(const typename Form<PREC>::S* __restrict__ a,
                                               typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                               int chunk_len, int strips, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int s = tile % strips, chunk = tile / strips;
	const int xb = chunk * chunk_len;
	ac_strip<PREC, TZ>(a, b, g, p, s * TZ, xb, min(xb + chunk_len, g.n[0]));
}

template <int PREC, int Q, int TZ>
__global__ void __launch_bounds__(TZ) gg_tiled(const typename Form<PREC>::S* __restrict__ a,
                                               typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                               int chunk_len, int strips, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int s = tile % strips, chunk = tile / strips;
	const int xb = chunk * chunk_len;
	gg_strip<PREC, Q, TZ>(a, b, g, p, s * TZ, xb, min(xb + chunk_len, g.n[0]));
}

template <int PREC, int Q, int TX, int TC, int NT>
__global__ void __launch_bounds__(NT) gg_tile2d(const typename Form<PREC>::S* __restrict__ a,
                                                typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p, int tiles_c,
                                                int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int tc = tile % tiles_c, tr = tile / tiles_c;
	gg_tile<PREC, Q, TX, TC, NT>(a, b, g, p, tr * TX, tc * TC);
}

// Wave strips (2-D Allen-Cahn and grain growth): one wave per (strip, chunk)
// item, four waves per workgroup, strips kWaveCols columns wide. Items run
// strip-fastest, so a workgroup's waves and an XCD's run of workgroups cover
// neighbouring strips of the same rows (shared halo columns in one L2).
template <int PREC, int Q, int KIND>
__global__ void __launch_bounds__(256) wave_strips(const typename Form<PREC>::S* __restrict__ a,
                                                   typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                   int chunk_len, int strips, int nitems, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int item = tile * 4 + (int)(threadIdx.x >> 6);
	if (item >= nitems) return;  // a whole wave; the waves share nothing
	const int s = item % strips, chunk = item / strips;
	const int xb = chunk * chunk_len;
	strip_wave<PREC, Q, KIND>(a, b, g, p, s * kWaveCols, xb, min(xb + chunk_len, g.n[0]));
}

// Split wave strips (2-D grain growth, Zener pinning): one workgroup of W
// waves per (strip, chunk) item, each wave holding Q/W of the fields. Items
// run strip-fastest, as the wave strips.
template <int PREC, int Q, int KIND, int W>
__global__ void __launch_bounds__(64 * W) wave_strips_split(const typename Form<PREC>::S* __restrict__ a,
                                                            typename Form<PREC>::S* __restrict__ b, Geom g,
                                                            ModelParams p, int chunk_len, int strips, int nitems)
{
	const int item = xcd_tile(nitems);
	if (item < 0) return;  // the whole workgroup
	const int s = item % strips, chunk = item / strips;
	const int xb = chunk * chunk_len;
	strip_wave_split<PREC, Q, KIND, W>(a, b, g, p, s * kWaveCols, xb, min(xb + chunk_len, g.n[0]));
}

// Two steps per pass on the split strips (periodic): one workgroup of W
// waves per (strip of kChWaveOut columns, chunk), strip-fastest.
template <int PREC, int Q, int KIND, int W>
__global__ void __launch_bounds__(64 * W) wave_strips_split2(const typename Form<PREC>::S* __restrict__ a,
                                                             typename Form<PREC>::S* __restrict__ b, Geom g,
                                                             ModelParams p, int chunk_len, int strips, int nitems)
{
	const int item = xcd_tile(nitems);
	if (item < 0) return;  // the whole workgroup
	const int s = item % strips, chunk = item / strips;
	const int xb = chunk * chunk_len;
	strip_wave_split2<PREC, Q, KIND, W>(a, b, g, p, s * kChWaveOut, xb, min(xb + chunk_len, g.n[0]));
}

// Temporal blocks (2-D periodic Allen-Cahn): one workgroup per TX x kTbCols
// tile, `steps` (<= kTbMaxSteps) steps per launch.
template <int PREC, int TX, int NT>
__global__ void __launch_bounds__(NT) ac_tblocks(const typename Form<PREC>::S* __restrict__ a,
                                                 typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                 int tiles_y, int ntiles, int steps)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int ty = tile % tiles_y, tx = tile / tiles_y;
	ac_tblock<PREC, TX, kTbMaxSteps, NT>(a, b, g, p, tx * TX, ty * kTbCols, steps);
}

// Cahn-Hilliard wave strips (2-D periodic): one wave per (strip, chunk) item,
// four waves per workgroup, kChWaveOut columns per strip; odd chunks march
// downward.
template <int PREC, int MU>
__global__ void __launch_bounds__(256) ch_waves(const typename Form<PREC>::S* __restrict__ a,
                                                typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                int chunk_len, int strips, int nitems, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int item = tile * 4 + (int)(threadIdx.x >> 6);
	if (item >= nitems) return;  // a whole wave; the waves share nothing
	const int s = item % strips, chunk = item / strips;
	const int xb = chunk * chunk_len;
	ch_wave<PREC, MU>(a, b, g, p, s * kChWaveOut, xb, min(xb + chunk_len, g.n[0]), chunk & 1);
}

// Ostwald ripening wave strips (2-D periodic), laid out as the Cahn-Hilliard
// ones (all rows march upward)
template <int PREC, int Q>
__global__ void __launch_bounds__(256) ow_waves(const typename Form<PREC>::S* __restrict__ a,
                                                typename Form<PREC>::S* __restrict__ b, Geom g, ModelParams p,
                                                int chunk_len, int strips, int nitems, int ntiles)
{
	const int tile = xcd_tile(ntiles);
	if (tile < 0) return;
	const int item = tile * 4 + (int)(threadIdx.x >> 6);
	if (item >= nitems) return;  // a whole wave; the waves share nothing
	const int s = item % strips, chunk = item / strips;
	const int xb = chunk * chunk_len;
	ow_wave<PREC, Q>(a, b, g, p, s * kChWaveOut, xb, min(xb + chunk_len, g.n[0]));
}

static bool all_periodic(const Geom& g)
{
	for (int k = 0; k < g.dim; k++)
		if (g.b0[k] != MMSPX_PERIODIC || g.b1[k] != MMSPX_PERIODIC) return false;
	return true;
}

static bool gg_q_compiled(int q) { return q == 2 || q == 4 || q == 8 || q == 16 || q == 32; }

// every axis periodic on both sides or on neither (the fused Cahn-Hilliard
// bodies map the non-periodic sides; an axis periodic on one side only takes
// the two-pass kernels)
static bool axes_uniform(const Geom& g)
{
	for (int k = 0; k < g.dim; k++)
		if ((g.b0[k] == MMSPX_PERIODIC) != (g.b1[k] == MMSPX_PERIODIC)) return false;
	return true;
}

// the Cahn-Hilliard bodies that map non-periodic sides: the register queues
// (3-D and the 2-D row strips) and the 2-D wave strips
static bool ch_body_maps_edges(const Geom& g, const TuneChoice& t)
{
	if (g.dim == 3) return t.opt == 1 || t.opt == 2 || t.opt == 3;
	return t.tz == kChWaveOut || t.opt == 1;
}

// field counts of the 3-D march for grain growth and Zener pinning
static bool march_q_compiled(int q) { return q == 2 || q == 4 || q == 8; }

bool fused_supported(const Geom& g, int prec, int model)
{
	(void)prec;
	if (g.dim == 3) {
		if (model == MMSPX_AC || model == MMSPX_AC_NOISE) return true;
		if (model == MMSPX_GG) return march_q_compiled(g.fields);
		if (model == MMSPX_ZENER || model == MMSPX_GG_ANISO) return march_q_compiled(g.fields) && prec != MMSPX_F32REF;
		if (model == MMSPX_OSTWALD) return g.fields >= 2 && g.fields <= 4 && prec != MMSPX_F32REF;
	}
	switch (model) {
	case MMSPX_CH:
	case MMSPX_CH_DW:
	case MMSPX_CH_NOISE: return g.dim >= 2 && axes_uniform(g);
	case MMSPX_AC:
	case MMSPX_AC_NOISE: return g.dim == 2;
	case MMSPX_GG: return g.dim == 2 && gg_q_compiled(g.fields);
	// Ostwald ripening: the wave strips for 2 to 8 fields on periodic grids
	case MMSPX_OSTWALD:
		return g.dim == 2 && all_periodic(g) && g.fields >= 2 && g.fields <= 8 && prec != MMSPX_F32REF;
	// multi-phase grain growth: the wave strips for 2, 4 and 8 fields (the
	// per-node arrays of more fields would not fit the registers)
	case MMSPX_GG_ANISO:
		return g.dim == 2 && (g.fields == 2 || g.fields == 4 || g.fields == 8) && prec != MMSPX_F32REF;
	// Zener pinning: the wave strips, where their registers hold the fields
	case MMSPX_ZENER:
		return g.dim == 2 && gg_q_compiled(g.fields) && prec != MMSPX_F32REF && (g.fields < 32 || prec != MMSPX_F64);
	}
	return false;
}

// Periodic 3-D Cahn-Hilliard with a deterministic chemical potential: two
// steps per pass on 32 x 32 tiles (768 threads) or, in FP32, 32 x 64 tiles
// (896 threads), whichever the cost model below favours. A workgroup's time
// grows with its planes (chunk + 6: three halo planes of the two steps on
// each side are recomputed) times the R3 points each thread owns; a launch
// takes as many such rounds as its workgroups fill the 256 CUs. One
// workgroup fits a CU, except FP32 on 32 x 32 tiles (80 VGPRs, 44 KB of LDS:
// two per CU), where a launch of more than 256 workgroups takes rounds of
// 512, each 1.7 times as long (the two share the CU; fitted to the FP32
// sweep, which the model then matches at every side). The chunk with the
// least rounds x time wins (ties: the longer chunk).
// Against a sweep of every chunk on cubes of side 48 to 256 (us per step):
// the model picks the measured best in FP64 at every side (128^3: 8-plane
// chunks, 10.8 against 22.9 for the former 8 x 32 LDS-ring default; 192^3:
// 32, 25.8 against 41.5; 224^3: 48, 37.8 against 59.7; 256^3: 64, 50.5) and
// in FP32 (128^3 32 x 32 tiles with 8-plane chunks 8.8 against 18.8; 192^3
// 32 x 64 with 16, 16.9 against 34.0; 224^3 32 x 32 with 24, 25.9 against
// 47.1; 256^3 32 x 64 with 32, 29.6).
static TuneChoice two_step_choice(const Geom& g, int prec)
{
	constexpr int kCUs = 256;
	static const int chunks[] = {6, 8, 12, 16, 24, 32, 48, 64};
	TuneChoice best{2, 32, 32, 64, 7};
	double best_cost = 1e300;
	for (int body = 7; body >= 6; body--) {
		if (body == 6 && prec != MMSPX_F32) continue;  // FP64 32 x 64 tiles spill
		const int ty = 32, tz = body == 7 ? 32 : 64, nt = body == 7 ? 768 : 896;
		const double own = (double)(ty + 6) * (tz + 6) / nt;  // R3 points per thread
		const bool two_per_cu = body == 7 && prec == MMSPX_F32;
		const int64_t tiles = (int64_t)((g.n[1] + ty - 1) / ty) * ((g.n[2] + tz - 1) / tz);
		for (int c : chunks) {
			const int64_t wgs = tiles * ((g.n[0] + c - 1) / c);
			const double rounds = (two_per_cu && wgs > kCUs) ? 1.7 * (double)((wgs + 2 * kCUs - 1) / (2 * kCUs))
			                                                 : (double)((wgs + kCUs - 1) / kCUs);
			const double cost = rounds * (c + 6) * own;
			if (cost <= best_cost) {
				best_cost = cost;
				best = TuneChoice{2, ty, tz, c, body};
			}
		}
	}
	return best;
}

// Periodic 3-D Allen-Cahn, grain growth and Zener pinning: the plane march
// two steps per pass (body 12) where its 64-node tile rows fit the last axis
// (padding at most an eighth of it: 160^3 Allen-Cahn FP64 pads to 192 and
// takes 15.5 us per step against 14.1 for the one-step march).
// Allen-Cahn: 16 x 64 or 32 x 64 tiles and the chunk from the cost model of
// two_step_choice: rounds x (chunk + 3 planes: the halo plane on each side
// and the queue's first load) x owned points per thread (16 x 64: 1188 on
// 640 threads, two workgroups per CU in rounds of 512 each 1.7 times as
// long; 32 x 64: 2244 on 768, one per CU). It picks the measured best, or
// one within 5 %, at every side swept (us per step, against the one-step
// march): FP64 64^3 16 x 64 with 4-plane chunks 4.3 (5.7), 128^3 16 x 64
// with 8 7.2 (10.3), 192^3 32 x 64 with 16 16.7 (21.9), 256^3 32 x 64 with
// 32 30.0 (43.9; 16 x 64 29.8); FP32 64^3 3.6 (6.9), 128^3 5.6 (8.0; 32 x 64
// with 4: 5.3), 192^3 11.2 (15.7), 256^3 19.0 (24.3).
// Grain growth and Zener pinning (2 or 4 fields; 8 in FP32 from 2^22 nodes):
// 8 x 64 tiles (one node and its halo per thread, 704 threads), 64-plane
// chunks halved until a launch has 256 workgroups per CU that fit (two for
// 2 fields, one otherwise by registers): 256^3 x 2 FP64 81.0 (117.0), x 4
// FP64 178.6 (239.6), x 4 FP32 111.8 (125.2), x 8 FP32 222.0 (253.3); Zener
// x 2 FP32 34.7 (54.8), x 4 107.0 (122.4); 192^3 x 8 FP32 109.1 (125.6);
// 128^3 x 2 FP32 8.6 (12.2), x 4 FP64 24.6 (27.4); 64^3 x 4 FP32 5.6 (7.1).
// 128^3 x 8 FP32 gains nothing (29.5 against 28.6), and 8 FP64 fields fit
// the LDS on 8 x 32 tiles only (256^3: 488.5 against 481.0): those stay on
// the one-step march.
static bool march2_choice(const Geom& g, int prec, int model, TuneChoice* out)
{
	constexpr int kCUs = 256;
	if (!all_periodic(g) || (prec != MMSPX_F64 && prec != MMSPX_F32)) return false;
	if (model != MMSPX_AC && model != MMSPX_GG && model != MMSPX_ZENER) return false;
	const int nz = g.n[2], pad = (nz + 63) / 64 * 64 - nz;
	if (pad * 8 > nz) return false;
	if (model == MMSPX_AC) {
		double best_cost = 1e300;
		for (int ty : {16, 32}) {
			const double own = ty == 16 ? 1188.0 / 640 : 2244.0 / 768;
			const int64_t tiles = (int64_t)((g.n[1] + ty - 1) / ty) * ((nz + 63) / 64);
			for (int c : {4, 8, 16, 32, 64}) {
				const int64_t wgs = tiles * ((g.n[0] + c - 1) / c);
				const double rounds = (ty == 16 && wgs > kCUs) ? 1.7 * (double)((wgs + 2 * kCUs - 1) / (2 * kCUs))
				                                               : (double)((wgs + kCUs - 1) / kCUs);
				const double cost = rounds * (c + 3) * own;
				if (cost < best_cost) {
					best_cost = cost;
					*out = TuneChoice{2, ty, 64, c, 12};
				}
			}
		}
		return true;
	}
	const int q = g.fields;
	if (!(q == 2 || q == 4 || (q == 8 && prec == MMSPX_F32 && g.nodes >= ((int64_t)1 << 22)))) return false;
	const int per_cu = q == 2 ? 2 : 1;
	const int64_t tiles = (int64_t)((g.n[1] + 7) / 8) * ((nz + 63) / 64);
	int chunk = 64;
	while (chunk > 4 && tiles * ((g.n[0] + chunk - 1) / chunk) < (int64_t)kCUs * per_cu) chunk /= 2;
	*out = TuneChoice{2, 8, 64, chunk, 12};
	return true;
}

// Defaults come from sweeps of the candidates below on MI355X (the figures
// quoted are microseconds per step from those sweeps).
TuneChoice default_choice(const Geom& g, int prec, int model)
{
	if (!fused_supported(g, prec, model)) return TuneChoice{1, 0, 0, 0, 0};
	const int64_t nodes = g.nodes;
	// 3-D Allen-Cahn, model A, grain growth, Zener pinning and multi-phase
	// grain growth: the plane march.
	// 256^3 sweep (us per step): Allen-Cahn FP64 16 x 32 tiles with 32-plane
	// chunks 44.5 (6.0 TB/s; 8 x 64 45.0, 32 x 64 46.5, one thread per node
	// 93.5), FP32 32 x 64 with 16-plane chunks 24.6 (16 x 64 24.8; 32-plane
	// chunks 31.7); grain growth 8 fields 8 x 64 with 64-plane chunks, FP32
	// 252 (16 x 32 275, one thread per node 430), FP64 484 (935); Zener 4
	// fields FP32 122 (211). Chunks are halved until the launch has 512
	// workgroups (two per CU) or more.
	// Multi-phase grain growth is bound by its per-node arithmetic, not by
	// bytes: one-hot grains with 8 fields, 128^3 FP64 200 us per step on 8 x
	// 64 tiles with 16-plane chunks (256 workgroups; 206 with 8-plane chunks,
	// 374 with 32) against 253 for one thread per node, FP32 87 against 155;
	// at 64^3 one thread per node wins (41.5 against 54.1), so the march
	// starts at 2^21 nodes, with 16-plane chunks halved to no less than 8
	// until the launch has 256 workgroups.
	if (g.dim == 3 && model == MMSPX_GG_ANISO) {
		if (nodes < ((int64_t)1 << 21)) return TuneChoice{1, 0, 0, 0, 0};
		const int64_t tiles = (int64_t)((g.n[1] + 7) / 8) * ((g.n[2] + 63) / 64);
		int chunk = 16;
		while (chunk > 8 && tiles * ((g.n[0] + chunk - 1) / chunk) < 256) chunk /= 2;
		return TuneChoice{2, 8, 64, chunk, 8};
	}
	// Ostwald ripening: the two passes as marches, 8 x 64 tiles with 32-plane
	// chunks halved (to no less than 4) until a launch has 1024 workgroups:
	// 256^3 x 2 fields FP64 215 us per step against 340 for the generic
	// passes (FP32 121.5 against 172.6), 128^3 x 4 FP64 with 4-plane chunks
	// 40.4 against 57.4; below 2^20 nodes the generic passes win (the
	// example's 128 x 64 x 16 grid: 8.4 against 12.9)
	// Periodic with 2 fields and 64-node tile rows that pad the last axis by at
	// most an eighth: the one-pass march (body 12, w kept in LDS), 8 x 64
	// tiles with 64-plane chunks halved until a launch has 512 workgroups
	// (two per CU): 256^3 FP64 156.8 us per step against 213.7 for the two
	// marches, FP32 75.1 against 121.6; 128^3 FP64 24.0 against 28.7. With 3
	// or 4 fields its registers leave one workgroup per CU in FP64 (192^3 x 3
	// FP64 134.0 against 125.6; 128^3 x 4 FP64 40.0 against 40.3), so those
	// stay on the two marches.
	if (g.dim == 3 && model == MMSPX_OSTWALD) {
		if (nodes < ((int64_t)1 << 20)) return TuneChoice{1, 0, 0, 0, 0};
		const int64_t tiles = (int64_t)((g.n[1] + 7) / 8) * ((g.n[2] + 63) / 64);
		const int pad = (g.n[2] + 63) / 64 * 64 - g.n[2];
		if (g.fields == 2 && all_periodic(g) && (prec == MMSPX_F64 || prec == MMSPX_F32) && pad * 8 <= g.n[2]) {
			int c = 64;
			while (c > 4 && tiles * ((g.n[0] + c - 1) / c) < 512) c /= 2;
			return TuneChoice{2, 8, 64, c, 12};
		}
		int chunk = 32;
		while (chunk > 4 && tiles * ((g.n[0] + chunk - 1) / chunk) < 1024) chunk /= 2;
		return TuneChoice{2, 8, 64, chunk, 8};
	}
	TuneChoice two;
	if (g.dim == 3 && march2_choice(g, prec, model, &two)) return two;
	if (g.dim == 3 && (model == MMSPX_AC || model == MMSPX_AC_NOISE || model == MMSPX_GG || model == MMSPX_ZENER)) {
		const bool scalar = model == MMSPX_AC || model == MMSPX_AC_NOISE;
		const int ty = scalar ? (prec == MMSPX_F64 ? 16 : 32) : 8;
		const int tz = scalar ? (prec == MMSPX_F64 ? 32 : 64) : 64;
		int chunk = scalar ? (prec == MMSPX_F64 ? 32 : 16) : 64;
		const int64_t tiles = (int64_t)((g.n[1] + ty - 1) / ty) * ((g.n[2] + tz - 1) / tz);
		while (chunk > 4 && tiles * ((g.n[0] + chunk - 1) / chunk) < 512) chunk /= 2;
		return TuneChoice{2, ty, tz, chunk, 8};
	}
	// Allen-Cahn: periodic grids take temporal blocks of 48 columns and 8
	// steps per pass (512^2: 2.05 us per step in FP64 against 4.58 us for one
	// thread per node). Block heights from a sweep of 8, 16, 32 rows (on 512,
	// 1024, 1024 threads) x 4, 8 steps, sides 256 to 4096: 16 rows up to 2^16
	// nodes (256^2 FP64 1.10 us, FP32 0.97), 32 rows above (512^2 FP64 1.48
	// against 1.57 with 16 rows and 1.69 with 8, FP32 1.26; 1024^2 3.09 and
	// 2.22; 2048^2 10.6 against 13.6 with 16 rows, FP32 6.95 against 9.41;
	// 4096^2 36.7 and 24.6 against 47.4 and 33.5).
	// Other boundaries and model A: the wave strips, 4-row chunks below 2^21
	// nodes, 8 below 2^23, 16 above (Neumann FP64: 1024^2 5.5 us per step
	// against 8.6 for one thread per node, 2048^2 12.6 against 16.5, 4096^2
	// 43.3 against 61.1; model A FP64 1024^2 7.7 us, 4096^2 60.8)
	const int wave_chunk = nodes < ((int64_t)1 << 21) ? 4 : nodes < ((int64_t)1 << 23) ? 8 : 16;
	const int tb_rows = nodes <= ((int64_t)1 << 16) ? 16 : 32;
	if (model == MMSPX_AC)
		return all_periodic(g) ? TuneChoice{2, tb_rows, kTbCols, 8, 3} : TuneChoice{2, 1, kWaveCols, wave_chunk, 2};
	if (model == MMSPX_AC_NOISE) return TuneChoice{2, 1, kWaveCols, wave_chunk, 2};
	// multi-phase grain growth (one-hot grains, 8 fields): 2-row chunks below
	// 2^21 nodes (1024^2 FP64 36.7 us per step against 39.1 with 4 rows, 40.0
	// with 8 and 77.3 for one thread per node), 8 rows above (2048^2 FP64 139
	// us, FP32 70; 4096^2 FP64 500, FP32 265: within 2 % of the best of 2..16)
	if (model == MMSPX_GG_ANISO) return TuneChoice{2, 1, kWaveCols, nodes < ((int64_t)1 << 21) ? 2 : 8, 2};
	// Ostwald ripening (4 fields): with the row rings, 4-row chunks below 2^21
	// nodes (1024^2 FP64 17.0 us per step against 18.0 with 8 rows; the
	// generic two passes: 28 us), 8 above (2048^2 FP64 51.1 against 52.2, FP32
	// 27.2 against 28.5)
	if (model == MMSPX_OSTWALD) return TuneChoice{2, 1, kChWaveOut, nodes < ((int64_t)1 << 21) ? 4 : 8, 0};
	// Grain growth and Zener pinning with 8, 16 or 32 fields: the split strips.
	// Below 2^21 nodes four waves per strip (a quarter of the fields each)
	// with 8-row chunks in FP32 and 16 in FP64; above, two waves with 64-row
	// chunks. Measured per step: 2048^2 x 32 FP32 233.7 us against 248.2 for
	// the wave strips (32-row chunks) and 388 for one thread per node; 2048^2
	// x 32 FP64 459.5 against 835.9 for the 8 x 64 tile; 2048^2 x 16 Zener
	// FP32 115.4 against 118.3; 512^2 x 32 FP32 15.7 against 20.3 on two
	// waves; 512^2 x 8 FP32 7.9 against 9.4 for one thread per node; 1024^2 x
	// 16 FP64 47.1 against 82.0.
	// Periodic grids take the split strips two steps per pass on four waves
	// (body 10): 2048^2 x 32 FP32 149.6 us per step with 16-row chunks
	// against 241.2 for the one-step split strips, FP64 275.0 with 32-row
	// chunks against 465.6; 2048^2 x 16 Zener FP32 73.0 against 117.3;
	// 1024^2 x 32 FP32 30.7 with 8-row chunks against 49.6; 512^2 x 8 FP32 5.0
	// against 8.1; 1024^2 x 8 FP64 15.8 against 24.7 (two waves: 1-8 % slower).
	if ((model == MMSPX_GG || model == MMSPX_ZENER) && (g.fields == 8 || g.fields == 16 || g.fields == 32)) {
		if (all_periodic(g))
			return TuneChoice{2, 1, kChWaveOut, nodes < ((int64_t)1 << 21) ? 8 : prec == MMSPX_F64 ? 32 : 16, 10};
		if (nodes < ((int64_t)1 << 21)) return TuneChoice{2, 1, kWaveCols, prec == MMSPX_F64 ? 16 : 8, 5};
		return TuneChoice{2, 1, kWaveCols, 64, 4};
	}
	// Grain growth and Zener pinning with 2 or 4 fields: the wave strips,
	// 4-row chunks below 2^21 nodes, 16 above (2 and 4 fields, 512^2 .. 4096^2,
	// FP32 and FP64: 1.4 to 1.9 times faster than one thread per node; 1024^2
	// x 4 FP32 8.0 us against 12.0, 4096^2 x 4 FP64 219.7 against 354.1)
	// Periodic: two steps per pass on one wave per strip (body 11), 8-row
	// chunks below 2^21 nodes and 16 above (2048^2 x 2 FP64 13.8 us per step
	// against 22.2 for the wave strips, 2048^2 x 4 FP32 14.4 against 23.0,
	// 1024^2 x 4 FP64 9.7 against 14.1, 512^2 x 2 FP64 3.2 against 4.6; Zener
	// 2048^2 x 4 FP32 13.5 against 22.5)
	if ((model == MMSPX_GG || model == MMSPX_ZENER) && all_periodic(g))
		return TuneChoice{2, 1, kChWaveOut, nodes < ((int64_t)1 << 21) ? 8 : 16, 11};
	if (model == MMSPX_GG || model == MMSPX_ZENER)
		return TuneChoice{2, 1, kWaveCols, nodes < ((int64_t)1 << 21) ? 4 : 16, 2};
	// 2-D Cahn-Hilliard: wave strips (1024^2 FP64 6.8 us per step against 13.5
	// us for the LDS row strips; 4096^2 FP64 42.4 us against 69 us). Chunks
	// from a sweep of 3..32 rows with the row ring: 3 below 2^17 nodes (256^2
	// FP64 3.0 us against 3.2 with 4 rows), 4 below 2^21 (1024^2 6.8, FP32
	// 5.2), then 16 in FP64 (2048^2 15.6 against 16.5 with 8; 4096^2 42.4) and
	// 8 in FP32 below 2^23 (2048^2 11.3 against 11.5), 16 above (4096^2 30.7)
	if (g.dim == 2) {
		const int c = nodes < ((int64_t)1 << 17)                          ? 3
		              : nodes < ((int64_t)1 << 21)                        ? 4
		              : nodes < ((int64_t)1 << 23) && prec != MMSPX_F64 ? 8
		                                                                  : 16;
		return TuneChoice{2, 1, kChWaveOut, c, 0};
	}
	// other boundaries than periodic: the register queues (8 x 32 tiles on
	// 256 threads below 2^24 nodes, 32 x 64 on 768 above)
	if (!all_periodic(g))
		return nodes < (int64_t)16 * 64 * 32 * 512 ? TuneChoice{2, 8, 32, 16, 1}
		                                            : TuneChoice{2, 32, 64, prec == MMSPX_F64 ? 32 : 16, 2};
	if (model != MMSPX_CH_NOISE && (prec == MMSPX_F64 || prec == MMSPX_F32)) return two_step_choice(g, prec);
	if (nodes < (int64_t)16 * 64 * 32 * 512) return TuneChoice{2, 8, 32, 16, 0};
	// 256^3: register queues on 32 x 64 tiles with 768 threads. FP64 with
	// 32-plane chunks, one workgroup per CU (52.3 us per step, against 53.0
	// with 512 or 1024 threads, 55.9 with 16-plane chunks, 54.4 for 16 x 64
	// tiles with 24-plane chunks, 64-75 us for chunks of 12, 20 and 24 that
	// leave CUs with uneven work); FP32 with 16-plane chunks, two per CU
	// (31.3 us, against 31.7 with 512 threads, 35.0 with 32-plane chunks)
	// model B (noise, FP64): 32-plane chunks on 1024 threads (95.9 us against
	// 97.6 with 768 in one sweep, 92.6 against 91.6 in another: the two are
	// within the noise; 16-plane chunks take 99-107 us)
	if (model == MMSPX_CH_NOISE && prec == MMSPX_F64) return TuneChoice{2, 32, 64, 32, 3};
	return TuneChoice{2, 32, 64, prec == MMSPX_F64 ? 32 : 16, 2};
}

bool tile_known(const Geom& g, const TuneChoice& t)
{
	if (t.chunk < 2 || t.chunk > 4096) return false;
	if (g.dim == 3)
		return (t.ty == 8 && t.tz == 64) || (t.ty == 16 && t.tz == 64) || (t.ty == 32 && t.tz == 64) ||
		       (t.ty == 8 && t.tz == 32) || (t.ty == 4 && t.tz == 128) || (t.ty == 32 && t.tz == 32) ||
		       (t.ty == 16 && t.tz == 32);
	if (g.dim == 2)
		return (t.ty == 1 && (t.tz == 64 || t.tz == 128 || t.tz == 256 || t.tz == kChWaveOut)) ||
		       ((t.ty == 8 || t.ty == 16 || t.ty == 32) && t.tz == kTbCols && t.chunk <= kTbMaxSteps) ||
		       (t.opt == 1 && (t.ty == 8 || t.ty == 16) && t.tz == 64);
	return false;
}

int tile_candidates(const Geom& g, int prec, int model, TuneChoice* out, int max)
{
	int n = 0;
	auto add = [&](int ty, int tz, int chunk, int opt) {
		if (n < max) out[n++] = TuneChoice{2, ty, tz, chunk, opt};
	};
	const bool ch = ch_family(model);
	if (g.dim == 2 && model == MMSPX_OSTWALD) {
		for (int c : {4, 8, 16, 32, 64}) add(1, kChWaveOut, c, 0);
		return n;
	}
	if (g.dim == 2 && model == MMSPX_GG_ANISO) {
		for (int c : {2, 4, 8, 16, 32}) add(1, kWaveCols, c, 2);
		return n;
	}
	const bool split = (model == MMSPX_GG || model == MMSPX_ZENER) && (g.fields == 8 || g.fields == 16 || g.fields == 32);
	if (g.dim == 2 && (model == MMSPX_AC_NOISE || model == MMSPX_ZENER)) {
		// model A and Zener pinning: wave strips, and the split strips
		for (int c : {8, 12, 16, 24, 32}) add(1, kWaveCols, c, 2);
		if (split)
			for (int c : {8, 16, 32, 64})
				for (int body = 4; body <= 5; body++) add(1, kWaveCols, c, body);
		if (split && model == MMSPX_ZENER && all_periodic(g))
			for (int c : {8, 16, 32, 64})
				for (int body = 9; body <= 10; body++) add(1, kChWaveOut, c, body);
		if (model == MMSPX_ZENER && (g.fields == 2 || g.fields == 4 || g.fields == 8) && all_periodic(g))
			for (int c : {8, 16, 32, 64}) add(1, kChWaveOut, c, 11);
		return n;
	}
	if (g.dim == 3 && !ch) {
		// the plane march: Allen-Cahn on four tiles, the vector models on two
		const bool scalar = model == MMSPX_AC || model == MMSPX_AC_NOISE;
		for (int c : {16, 32, 64}) {
			add(8, 64, c, 8);
			add(16, 32, c, 8);
			if (scalar) {
				add(16, 64, c, 8);
				add(32, 64, c, 8);
			}
		}
		// two steps per pass on periodic grids (Allen-Cahn, grain growth,
		// Zener pinning; the 8-field FP64 planes fit the LDS on 8 x 32 only)
		// (Ostwald ripening: the one-pass march)
		const bool two = all_periodic(g) && model != MMSPX_AC_NOISE && model != MMSPX_GG_ANISO &&
		                 (prec == MMSPX_F64 || prec == MMSPX_F32);
		if (two)
			for (int c : {8, 16, 32, 64}) {
				if (scalar) {
					add(16, 64, c, 12);
					add(32, 64, c, 12);
				} else if (g.fields == 8 && prec == MMSPX_F64) {
					add(8, 32, c, 12);
				} else {
					add(8, 64, c, 12);
					add(16, 32, c, 12);
				}
			}
		return n;
	}
	const bool edges = !all_periodic(g);  // Cahn-Hilliard: the bodies that map non-periodic sides
	if (g.dim == 3) {
		const int tiles[4][2] = {{16, 64}, {8, 64}, {8, 32}, {4, 128}};
		for (auto& t : tiles)
			for (int c : {16, 32, 64})
				for (int opt = edges ? 1 : 0; opt < 2; opt++) add(t[0], t[1], c, opt);
		// 32 x 64: register queues only (the LDS rings would leave one
		// workgroup per CU), on 512 and 768 threads
		for (int c : {12, 16, 22, 32})
			for (int body = 1; body <= 3; body++) add(32, 64, c, body);
		for (int c : {20, 22, 24}) add(16, 64, c, 1);
		// two steps per pass (deterministic chemical potentials, periodic grids)
		if ((model == MMSPX_CH || model == MMSPX_CH_DW) && !edges)
			for (int c : {6, 8, 12, 16, 24, 32, 48, 64}) {
				add(32, 64, c, 6);
				add(32, 32, c, 7);
			}
	} else if (g.dim == 2) {
		// LDS-row strips (Cahn-Hilliard: both tile bodies; non-periodic: the register queues)
		for (int tz : {64, 128, 256})
			for (int c : {16, 32, 64, 128})
				for (int opt = (ch && edges) ? 1 : 0; opt < (ch ? 2 : 1); opt++) add(1, tz, c, opt);
		// Cahn-Hilliard wave strips
		if (ch)
			for (int c : {8, 16, 32, 64}) add(1, kChWaveOut, c, 0);
		// grain growth: 2-D tiles
		if (model == MMSPX_GG && g.fields > 1) {
			add(8, 64, 8, 1);
			add(16, 64, 16, 1);
		}
		// Allen-Cahn and grain growth: wave strips, where compiled
		const bool wave = model == MMSPX_AC || (model == MMSPX_GG && (g.fields < 32 || prec != MMSPX_F64));
		if (wave)
			for (int c : {8, 12, 16, 24, 32}) add(1, kWaveCols, c, 2);
		if (split)
			for (int c : {8, 16, 32, 64})
				for (int body = 4; body <= 5; body++) add(1, kWaveCols, c, body);
		// periodic: the split strips two steps per pass (one wave for 2, 4, 8 fields)
		if (split && all_periodic(g))
			for (int c : {8, 16, 32, 64})
				for (int body = 9; body <= 10; body++) add(1, kChWaveOut, c, body);
		if ((model == MMSPX_GG || model == MMSPX_ZENER) && (g.fields == 2 || g.fields == 4 || g.fields == 8) &&
		    all_periodic(g))
			for (int c : {8, 16, 32, 64}) add(1, kChWaveOut, c, 11);
		// periodic Allen-Cahn: temporal blocks (chunk = steps per pass)
		if (model == MMSPX_AC && all_periodic(g))
			for (int tx : {8, 16, 32})
				for (int k : {4, 8}) add(tx, kTbCols, k, 3);
	}
	return n;
}

template <int PREC, int MU, int DIM, int TY, int TZ, int NT, int BODY>
static int launch_ch_body(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int nx = g.n[0];
	const int ny = DIM == 3 ? g.n[1] : 1;
	const int nz = DIM == 3 ? g.n[2] : g.n[1];
	const int tiles_y = (ny + TY - 1) / TY;
	const int tiles_z = (nz + TZ - 1) / TZ;
	const int chunks = (nx + chunk - 1) / chunk;
	const int ntiles = tiles_y * tiles_z * chunks;
	// round the grid up to a multiple of 8 so the XCD remap applies; extra ids exit
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((ch_fused<PREC, MU, DIM, TY, TZ, NT, BODY>), dim3(blocks), dim3(NT), 0, st, (const S*)a, (S*)b, g,
	                   p, chunk, tiles_y, tiles_z, ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

// The grid weights and the gradient coefficient are all exactly 1 (dx = 1,
// kappa or K = 1): the two-step bodies then leave out the multiplications by them.
static bool unit_weights(const Geom& g, const ModelParams& p, int mu)
{
	for (int k = 0; k < 3; k++)
		if (g.w[k] != 1.0 || g.wf[k] != 1.0f) return false;
	return mu == 1 ? (p.kappa == 1.0 && p.kappaf == 1.0f) : (p.K == 1.0 && p.Kf == 1.0f);
}

template <int PREC, int MU, int TY, int TZ, int NT, bool PHI2>
static int launch_ch2_body(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int tiles_y = (g.n[1] + TY - 1) / TY;
	const int tiles_z = (g.n[2] + TZ - 1) / TZ;
	const int chunks = (g.n[0] + chunk - 1) / chunk;
	const int ntiles = tiles_y * tiles_z * chunks;
	const int blocks = (ntiles + 7) / 8 * 8;
	if (unit_weights(g, p, MU))
		hipLaunchKernelGGL((ch_fused2<PREC, MU, TY, TZ, NT, PHI2, true>), dim3(blocks), dim3(NT), 0, st, (const S*)a,
		                   (S*)b, g, p, chunk, tiles_y, tiles_z, ntiles);
	else
		hipLaunchKernelGGL((ch_fused2<PREC, MU, TY, TZ, NT, PHI2, false>), dim3(blocks), dim3(NT), 0, st, (const S*)a,
		                   (S*)b, g, p, chunk, tiles_y, tiles_z, ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

// opt 1 selects the register-queue body (opt 2 and 3, register queues with
// more threads per tile, are dispatched in launch_ch_prec)
template <int PREC, int MU, int DIM, int TY, int TZ, int NT>
static int launch_ch(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk, int opt)
{
	if (opt == 1) return launch_ch_body<PREC, MU, DIM, TY, TZ, NT, 1>(g, p, a, b, st, chunk);
	return launch_ch_body<PREC, MU, DIM, TY, TZ, NT, 0>(g, p, a, b, st, chunk);
}

template <int PREC, int Q, int KIND, int TY, int TZ, int NT>
static int launch_march_body(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk,
                             const void* cin = nullptr, void* wout = nullptr)
{
	using S = typename Form<PREC>::S;
	const int tiles_y = (g.n[1] + TY - 1) / TY;
	const int tiles_z = (g.n[2] + TZ - 1) / TZ;
	const int chunks = (g.n[0] + chunk - 1) / chunk;
	const int ntiles = tiles_y * tiles_z * chunks;
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((march3d_k<PREC, Q, KIND, TY, TZ, NT>), dim3(blocks), dim3(NT), 0, st, (const S*)a, (S*)b, g,
	                   p, chunk, tiles_y, tiles_z, ntiles, (const S*)cin, (S*)wout);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

// 3-D Ostwald ripening: the two passes as plane marches (8 x 64 or 16 x 32
// tiles, 2, 3 or 4 fields); w goes through `aux` (one scalar plane per node)
template <int PREC>
static int launch_ow3d(const Geom& g, const ModelParams& p, const void* a, void* b, void* aux, hipStream_t st,
                       const TuneChoice& t)
{
	const int chunk = t.chunk > 0 ? t.chunk : 32;
	if (t.opt != 8) return fail(MMSPX_EINVAL, "3-D Ostwald ripening runs tile body 8");
	if (!aux) return fail(MMSPX_EINVAL, "3-D Ostwald ripening needs a scalar plane for w");
	const bool t864 = t.ty == 8 && t.tz == 64, t1632 = t.ty == 16 && t.tz == 32;
	if (!t864 && !t1632) return fail(MMSPX_EINVAL, "no 3-D march compiled for this tile");
	int rc;
	switch (g.fields) {
	case 2:
		rc = t864 ? launch_march_body<PREC, 2, 5, 8, 64, 512>(g, p, a, b, st, chunk, nullptr, aux)
		          : launch_march_body<PREC, 2, 5, 16, 32, 512>(g, p, a, b, st, chunk, nullptr, aux);
		break;
	case 3:
		rc = t864 ? launch_march_body<PREC, 3, 5, 8, 64, 512>(g, p, a, b, st, chunk, nullptr, aux)
		          : launch_march_body<PREC, 3, 5, 16, 32, 512>(g, p, a, b, st, chunk, nullptr, aux);
		break;
	case 4:
		rc = t864 ? launch_march_body<PREC, 4, 5, 8, 64, 512>(g, p, a, b, st, chunk, nullptr, aux)
		          : launch_march_body<PREC, 4, 5, 16, 32, 512>(g, p, a, b, st, chunk, nullptr, aux);
		break;
	default: return fail(MMSPX_EUNSUPPORTED, "the 3-D Ostwald march takes 2, 3 or 4 fields");
	}
	if (rc) return rc;
	return t864 ? launch_march_body<PREC, 1, 6, 8, 64, 256>(g, p, aux, b, st, chunk, a, nullptr)
	            : launch_march_body<PREC, 1, 6, 16, 32, 256>(g, p, aux, b, st, chunk, a, nullptr);
}

// 3-D Ostwald ripening in one pass (tile body 12, periodic): 2, 3 or 4
// fields on 8 x 64 tiles (704 threads) or 16 x 32 (640), one point of the
// tile and its halo per thread
template <int PREC, int Q, int TY, int TZ, int NT>
static int launch_ow1_body(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int tiles_y = (g.n[1] + TY - 1) / TY;
	const int tiles_z = (g.n[2] + TZ - 1) / TZ;
	const int chunks = (g.n[0] + chunk - 1) / chunk;
	const int ntiles = tiles_y * tiles_z * chunks;
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((march3d_ow_k<PREC, Q, TY, TZ, NT>), dim3(blocks), dim3(NT), 0, st, (const S*)a, (S*)b, g, p,
	                   chunk, tiles_y, tiles_z, ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int Q>
static int launch_ow1_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st,
                        const TuneChoice& t, int chunk)
{
	if (t.ty == 8 && t.tz == 64) return launch_ow1_body<PREC, Q, 8, 64, 704>(g, p, a, b, st, chunk);
	if (t.ty == 16 && t.tz == 32) return launch_ow1_body<PREC, Q, 16, 32, 640>(g, p, a, b, st, chunk);
	return fail(MMSPX_EINVAL, "no one-pass 3-D Ostwald march compiled for this tile");
}

template <int PREC>
static int launch_ow1(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, const TuneChoice& t)
{
	const int chunk = t.chunk > 0 ? t.chunk : 32;
	if (!all_periodic(g)) return fail(MMSPX_EUNSUPPORTED, "the one-pass 3-D Ostwald march needs periodic boundaries");
	switch (g.fields) {
	case 2: return launch_ow1_q<PREC, 2>(g, p, a, b, st, t, chunk);
	case 3: return launch_ow1_q<PREC, 3>(g, p, a, b, st, t, chunk);
	case 4: return launch_ow1_q<PREC, 4>(g, p, a, b, st, t, chunk);
	}
	return fail(MMSPX_EUNSUPPORTED, "the 3-D Ostwald march takes 2, 3 or 4 fields");
}

// tile body 8 on 3-D grids: Allen-Cahn on 8, 16 or 32 x 64 and 16 x 32 tiles;
// the vector models (2, 4 or 8 fields) on 8 x 64 and 16 x 32 tiles, one node
// per thread
template <int PREC, int KIND>
static int launch_march(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st,
                        const TuneChoice& t)
{
	const int chunk = t.chunk > 0 ? t.chunk : 32;
	if (t.opt != 8) return fail(MMSPX_EINVAL, "3-D Allen-Cahn, grain growth, Zener pinning and multi-phase grain growth run tile body 8");
	if constexpr (KIND == 0 || KIND == 2) {
		if (t.ty == 8 && t.tz == 64) return launch_march_body<PREC, 1, KIND, 8, 64, 256>(g, p, a, b, st, chunk);
		if (t.ty == 16 && t.tz == 64) return launch_march_body<PREC, 1, KIND, 16, 64, 256>(g, p, a, b, st, chunk);
		if (t.ty == 32 && t.tz == 64) return launch_march_body<PREC, 1, KIND, 32, 64, 512>(g, p, a, b, st, chunk);
		if (t.ty == 16 && t.tz == 32) return launch_march_body<PREC, 1, KIND, 16, 32, 256>(g, p, a, b, st, chunk);
	} else {
		const bool t864 = t.ty == 8 && t.tz == 64, t1632 = t.ty == 16 && t.tz == 32;
		switch (g.fields) {
		case 2:
			if (t864) return launch_march_body<PREC, 2, KIND, 8, 64, 512>(g, p, a, b, st, chunk);
			if (t1632) return launch_march_body<PREC, 2, KIND, 16, 32, 512>(g, p, a, b, st, chunk);
			break;
		case 4:
			if (t864) return launch_march_body<PREC, 4, KIND, 8, 64, 512>(g, p, a, b, st, chunk);
			if (t1632) return launch_march_body<PREC, 4, KIND, 16, 32, 512>(g, p, a, b, st, chunk);
			break;
		case 8:
			if (t864) return launch_march_body<PREC, 8, KIND, 8, 64, 512>(g, p, a, b, st, chunk);
			if (t1632) return launch_march_body<PREC, 8, KIND, 16, 32, 512>(g, p, a, b, st, chunk);
			break;
		default: return fail(MMSPX_EUNSUPPORTED, "the 3-D march takes 2, 4 or 8 fields");
		}
	}
	return fail(MMSPX_EINVAL, "no 3-D march compiled for this tile");
}

// Tile body 12 on periodic 3-D grids: the plane march two steps per pass.
// The threads own the tile and its one-cell halo, one point each for the
// vector models (8 x 64 tiles on 704 threads, 16 x 32 on 640, 8 x 32 on
// 384), two or three for Allen-Cahn (16 x 32 on 320, 16 x 64 on 640, 32 x 64
// on 768). LDS: the phi^n plane with a two-cell halo and the phi^{n+1} plane
// on the owned points, both double-buffered, Q fields each.
template <int PREC, int Q, int KIND, int TY, int TZ, int NT>
static int launch_march2_body(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	constexpr size_t lds = (size_t)2 * Q * ((TY + 4) * (TZ + 4) + (TY + 2) * (TZ + 2)) * sizeof(S);
	if constexpr (lds > (size_t)160 * 1024) {
		return fail(MMSPX_EINVAL, "the planes of this tile and field count do not fit the LDS");
	} else {
		const int tiles_y = (g.n[1] + TY - 1) / TY;
		const int tiles_z = (g.n[2] + TZ - 1) / TZ;
		const int chunks = (g.n[0] + chunk - 1) / chunk;
		const int ntiles = tiles_y * tiles_z * chunks;
		const int blocks = (ntiles + 7) / 8 * 8;
		hipLaunchKernelGGL((march3d2_k<PREC, Q, KIND, TY, TZ, NT>), dim3(blocks), dim3(NT), 0, st, (const S*)a,
		                   (S*)b, g, p, chunk, tiles_y, tiles_z, ntiles);
		MMSPX_HIP(hipGetLastError());
		return 0;
	}
}

template <int PREC, int Q, int KIND>
static int launch_march2_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st,
                           const TuneChoice& t, int chunk)
{
	if (t.ty == 8 && t.tz == 64) return launch_march2_body<PREC, Q, KIND, 8, 64, 704>(g, p, a, b, st, chunk);
	if (t.ty == 16 && t.tz == 32) return launch_march2_body<PREC, Q, KIND, 16, 32, 640>(g, p, a, b, st, chunk);
	if (t.ty == 8 && t.tz == 32) return launch_march2_body<PREC, Q, KIND, 8, 32, 384>(g, p, a, b, st, chunk);
	return fail(MMSPX_EINVAL, "no two-step 3-D march compiled for this tile");
}

template <int PREC, int KIND>
static int launch_march2(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st,
                         const TuneChoice& t, int nsteps)
{
	const int chunk = t.chunk > 0 ? t.chunk : 32;
	if (!all_periodic(g)) return fail(MMSPX_EUNSUPPORTED, "tile body 12 (two steps per pass) needs periodic boundaries");
	if (PREC == MMSPX_F32REF) return fail(MMSPX_EUNSUPPORTED, "tile body 12 runs the MMSPX_F64 and MMSPX_F32 forms");
	if (nsteps == 1) {
		// an odd last step: the one-step march
		const TuneChoice t1{2, KIND == 0 ? 16 : 8, KIND == 0 ? 32 : 64, chunk, 8};
		return launch_march<PREC, KIND>(g, p, a, b, st, t1);
	}
	if (nsteps != 2) return fail(MMSPX_EINVAL, "tile body 12 takes one or two steps per pass");
	if constexpr (PREC == MMSPX_F32REF) {
		return fail(MMSPX_EUNSUPPORTED, "tile body 12 runs the MMSPX_F64 and MMSPX_F32 forms");
	} else if constexpr (KIND == 0) {
		if (t.ty == 16 && t.tz == 32) return launch_march2_body<PREC, 1, 0, 16, 32, 320>(g, p, a, b, st, chunk);
		if (t.ty == 16 && t.tz == 64) return launch_march2_body<PREC, 1, 0, 16, 64, 640>(g, p, a, b, st, chunk);
		if (t.ty == 32 && t.tz == 64) return launch_march2_body<PREC, 1, 0, 32, 64, 768>(g, p, a, b, st, chunk);
		return fail(MMSPX_EINVAL, "no two-step 3-D march compiled for this tile");
	} else {
		switch (g.fields) {
		case 2: return launch_march2_q<PREC, 2, KIND>(g, p, a, b, st, t, chunk);
		case 4: return launch_march2_q<PREC, 4, KIND>(g, p, a, b, st, t, chunk);
		case 8: return launch_march2_q<PREC, 8, KIND>(g, p, a, b, st, t, chunk);
		}
		return fail(MMSPX_EUNSUPPORTED, "the two-step 3-D march takes 2, 4 or 8 fields");
	}
}

template <int PREC, int MU>
static int launch_ch_waves(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int strips = (g.n[1] + kChWaveOut - 1) / kChWaveOut;
	const int nitems = strips * ((g.n[0] + chunk - 1) / chunk);
	const int ntiles = (nitems + 3) / 4;
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((ch_waves<PREC, MU>), dim3(blocks), dim3(256), 0, st, (const S*)a, (S*)b, g, p, chunk, strips,
	                   nitems, ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int MU>
static int launch_ch_prec(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st,
                          const TuneChoice& t, int nsteps)
{
	const int chunk = t.chunk > 0 ? t.chunk : 32;
	if (!all_periodic(g) && !ch_body_maps_edges(g, t))
		return fail(MMSPX_EUNSUPPORTED, "non-periodic boundaries run the register-queue bodies (1, 2, 3) or the "
		                                "2-D wave strips");
	if (g.dim == 3 && (t.opt == 6 || t.opt == 7)) {
		// two steps per pass: body 6 on 32 x 64 tiles (896 threads; the c^n plane
		// double-buffered in FP32, single in FP64 where the LDS would not hold
		// two), body 7 on 32 x 32 tiles (768 threads, every plane
		// double-buffered); an odd last step runs the one-step register queues
		if constexpr (MU == 2) {
			return fail(MMSPX_EUNSUPPORTED, "the two-step bodies take the deterministic chemical potentials");
		} else {
			if (nsteps == 1)
				return launch_ch_body<PREC, MU, 3, 32, 64, 768, 1>(g, p, a, b, st, PREC == MMSPX_F64 ? 32 : 16);
			if (nsteps != 2) return fail(MMSPX_EINVAL, "the two-step bodies take one or two steps per pass");
			if (t.opt == 6) {
				if (t.ty != 32 || t.tz != 64) return fail(MMSPX_EINVAL, "tile body 6 runs 32 x 64 tiles");
				return launch_ch2_body<PREC, MU, 32, 64, 896, PREC != MMSPX_F64>(g, p, a, b, st, chunk);
			}
			if (t.ty != 32 || t.tz != 32) return fail(MMSPX_EINVAL, "tile body 7 runs 32 x 32 tiles");
			return launch_ch2_body<PREC, MU, 32, 32, 768, true>(g, p, a, b, st, chunk);
		}
	}
	if (g.dim == 3 && t.ty == 32 && t.tz == 32) return fail(MMSPX_EINVAL, "32 x 32 tiles run tile body 7");
	if (g.dim == 3 && (t.opt == 2 || t.opt == 3)) {
		// register queues on 32 x 64 tiles with 768 or 1024 threads (fewer own points each)
		if (t.ty != 32 || t.tz != 64) return fail(MMSPX_EINVAL, "tile bodies 2 and 3 run 32 x 64 tiles");
		if (t.opt == 3) return launch_ch_body<PREC, MU, 3, 32, 64, 1024, 1>(g, p, a, b, st, chunk);
		return launch_ch_body<PREC, MU, 3, 32, 64, 768, 1>(g, p, a, b, st, chunk);
	}
	if (g.dim == 3) {
		if (t.ty == 16 && t.tz == 64) return launch_ch<PREC, MU, 3, 16, 64, 256>(g, p, a, b, st, chunk, t.opt);
		if (t.ty == 32 && t.tz == 64) return launch_ch<PREC, MU, 3, 32, 64, 512>(g, p, a, b, st, chunk, t.opt);
		if (t.ty == 8 && t.tz == 32) return launch_ch<PREC, MU, 3, 8, 32, 256>(g, p, a, b, st, chunk, t.opt);
		if (t.ty == 4 && t.tz == 128) return launch_ch<PREC, MU, 3, 4, 128, 256>(g, p, a, b, st, chunk, t.opt);
		return launch_ch<PREC, MU, 3, 8, 64, 256>(g, p, a, b, st, chunk, t.opt);
	}
	if (t.tz == kChWaveOut) return launch_ch_waves<PREC, MU>(g, p, a, b, st, chunk);
	if (t.tz == 64) return launch_ch<PREC, MU, 2, 1, 64, 64>(g, p, a, b, st, chunk, t.opt);
	if (t.tz == 128) return launch_ch<PREC, MU, 2, 1, 128, 128>(g, p, a, b, st, chunk, t.opt);
	return launch_ch<PREC, MU, 2, 1, 256, 256>(g, p, a, b, st, chunk, t.opt);
}

template <int PREC, int TZ>
static int launch_ac(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int strips = (g.n[1] + TZ - 1) / TZ;
	const int ntiles = strips * ((g.n[0] + chunk - 1) / chunk);
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((ac_tiled<PREC, TZ>), dim3(blocks), dim3(TZ), 0, st, (const S*)a, (S*)b, g, p, chunk, strips,
	                   ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int Q, int TZ>
static int launch_gg_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int strips = (g.n[1] + TZ - 1) / TZ;
	const int ntiles = strips * ((g.n[0] + chunk - 1) / chunk);
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((gg_tiled<PREC, Q, TZ>), dim3(blocks), dim3(TZ), 0, st, (const S*)a, (S*)b, g, p, chunk, strips,
	                   ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int Q, int TX>
static int launch_gg_tile_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st)
{
	using S = typename Form<PREC>::S;
	constexpr int TC = 64, NT = 256;
	const int tiles_c = (g.n[1] + TC - 1) / TC;
	const int ntiles = tiles_c * ((g.n[0] + TX - 1) / TX);
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((gg_tile2d<PREC, Q, TX, TC, NT>), dim3(blocks), dim3(NT), 0, st, (const S*)a, (S*)b, g, p,
	                   tiles_c, ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int TX>
static int launch_gg_tile(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st)
{
	switch (g.fields) {
	case 2: return launch_gg_tile_q<PREC, 2, TX>(g, p, a, b, st);
	case 4: return launch_gg_tile_q<PREC, 4, TX>(g, p, a, b, st);
	case 8: return launch_gg_tile_q<PREC, 8, TX>(g, p, a, b, st);
	case 16: return launch_gg_tile_q<PREC, 16, TX>(g, p, a, b, st);
	case 32: return launch_gg_tile_q<PREC, 32, TX>(g, p, a, b, st);
	}
	return fail(MMSPX_EUNSUPPORTED, "no tiled grain-growth kernel for this field count");
}

template <int PREC, int TZ>
static int launch_gg_tz(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	switch (g.fields) {
	case 2: return launch_gg_q<PREC, 2, TZ>(g, p, a, b, st, chunk);
	case 4: return launch_gg_q<PREC, 4, TZ>(g, p, a, b, st, chunk);
	case 8: return launch_gg_q<PREC, 8, TZ>(g, p, a, b, st, chunk);
	case 16: return launch_gg_q<PREC, 16, TZ>(g, p, a, b, st, chunk);
	case 32: return launch_gg_q<PREC, 32, TZ>(g, p, a, b, st, chunk);
	}
	return fail(MMSPX_EUNSUPPORTED, "no tiled grain-growth kernel for this field count");
}

template <int PREC, int TX>
static int launch_tb(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int steps)
{
	using S = typename Form<PREC>::S;
	const int tiles_y = (g.n[1] + kTbCols - 1) / kTbCols;
	const int ntiles = tiles_y * ((g.n[0] + TX - 1) / TX);
	const int blocks = (ntiles + 7) / 8 * 8;
	// 8-row blocks on 8 waves, 16- and 32-row blocks on 16 (a wave takes 3
	// of the 24, 32 or 48 region rows per step): 512^2 FP64 with 32 rows 1.48
	// us per step on 1024 threads against 2.50 on 256
	constexpr int NT = TX == 8 ? 512 : 1024;
	hipLaunchKernelGGL((ac_tblocks<PREC, TX, NT>), dim3(blocks), dim3(NT), 0, st, (const S*)a, (S*)b, g, p, tiles_y,
	                   ntiles, steps);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int Q, int KIND>
static int launch_wave_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int strips = (g.n[1] + kWaveCols - 1) / kWaveCols;
	const int nitems = strips * ((g.n[0] + chunk - 1) / chunk);
	const int ntiles = (nitems + 3) / 4;
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((wave_strips<PREC, Q, KIND>), dim3(blocks), dim3(256), 0, st, (const S*)a, (S*)b, g, p, chunk,
	                   strips, nitems, ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

// The wave strips keep four rows of every field in registers: compiled for up
// to 128 bytes of field values per cell (FP32 up to 32 fields, FP64 up to 16).
template <int PREC>
static int launch_wave(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	constexpr bool wide = sizeof(typename Form<PREC>::S) == 4;
	if (p.model == MMSPX_AC) return launch_wave_q<PREC, 1, 0>(g, p, a, b, st, chunk);
	if (p.model == MMSPX_AC_NOISE) return launch_wave_q<PREC, 1, 2>(g, p, a, b, st, chunk);
	if (p.model == MMSPX_GG_ANISO) {
		if constexpr (PREC != MMSPX_F32REF) {
			switch (g.fields) {
			case 2: return launch_wave_q<PREC, 2, 4>(g, p, a, b, st, chunk);
			case 4: return launch_wave_q<PREC, 4, 4>(g, p, a, b, st, chunk);
			case 8: return launch_wave_q<PREC, 8, 4>(g, p, a, b, st, chunk);
			}
		}
		return fail(MMSPX_EUNSUPPORTED, "multi-phase wave strips take 2, 4 or 8 fields in the MMSPX_F64 and MMSPX_F32 forms");
	}
	if (p.model == MMSPX_ZENER) {
		if constexpr (PREC != MMSPX_F32REF) {
			switch (g.fields) {
			case 2: return launch_wave_q<PREC, 2, 3>(g, p, a, b, st, chunk);
			case 4: return launch_wave_q<PREC, 4, 3>(g, p, a, b, st, chunk);
			case 8: return launch_wave_q<PREC, 8, 3>(g, p, a, b, st, chunk);
			case 16: return launch_wave_q<PREC, 16, 3>(g, p, a, b, st, chunk);
			case 32:
				if constexpr (wide) return launch_wave_q<PREC, 32, 3>(g, p, a, b, st, chunk);
				break;
			}
		}
		return fail(MMSPX_EUNSUPPORTED, "no wave-strip kernel for this field count and precision");
	}
	switch (g.fields) {
	case 2: return launch_wave_q<PREC, 2, 1>(g, p, a, b, st, chunk);
	case 4: return launch_wave_q<PREC, 4, 1>(g, p, a, b, st, chunk);
	case 8: return launch_wave_q<PREC, 8, 1>(g, p, a, b, st, chunk);
	case 16: return launch_wave_q<PREC, 16, 1>(g, p, a, b, st, chunk);
	case 32:
		if constexpr (wide) return launch_wave_q<PREC, 32, 1>(g, p, a, b, st, chunk);
		break;
	}
	return fail(MMSPX_EUNSUPPORTED, "no wave-strip kernel for this field count and precision");
}

template <int PREC, int Q, int KIND, int W>
static int launch_split_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int strips = (g.n[1] + kWaveCols - 1) / kWaveCols;
	const int nitems = strips * ((g.n[0] + chunk - 1) / chunk);
	const int blocks = (nitems + 7) / 8 * 8;
	hipLaunchKernelGGL((wave_strips_split<PREC, Q, KIND, W>), dim3(blocks), dim3(64 * W), 0, st, (const S*)a, (S*)b, g,
	                   p, chunk, strips, nitems);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int KIND, int W>
static int launch_split_w(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	switch (g.fields) {
	case 8: return launch_split_q<PREC, 8, KIND, W>(g, p, a, b, st, chunk);
	case 16: return launch_split_q<PREC, 16, KIND, W>(g, p, a, b, st, chunk);
	case 32: return launch_split_q<PREC, 32, KIND, W>(g, p, a, b, st, chunk);
	}
	return fail(MMSPX_EUNSUPPORTED, "split wave strips take 8, 16 or 32 fields");
}

// The split strips: grain growth and Zener pinning with 8, 16 or 32 fields
// on W = 2 (body 4) or 4 (body 5) waves per strip
template <int PREC>
static int launch_split(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk,
                        int waves)
{
	if (p.model == MMSPX_ZENER) {
		if constexpr (PREC != MMSPX_F32REF)
			return waves == 4 ? launch_split_w<PREC, 3, 4>(g, p, a, b, st, chunk)
			                  : launch_split_w<PREC, 3, 2>(g, p, a, b, st, chunk);
		return fail(MMSPX_EUNSUPPORTED, "Zener pinning runs the MMSPX_F64 and MMSPX_F32 forms");
	}
	return waves == 4 ? launch_split_w<PREC, 1, 4>(g, p, a, b, st, chunk)
	                  : launch_split_w<PREC, 1, 2>(g, p, a, b, st, chunk);
}

template <int PREC, int Q, int KIND, int W>
static int launch_split2_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int strips = (g.n[1] + kChWaveOut - 1) / kChWaveOut;
	const int nitems = strips * ((g.n[0] + chunk - 1) / chunk);
	const int blocks = (nitems + 7) / 8 * 8;
	hipLaunchKernelGGL((wave_strips_split2<PREC, Q, KIND, W>), dim3(blocks), dim3(64 * W), 0, st, (const S*)a, (S*)b,
	                   g, p, chunk, strips, nitems);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC, int KIND, int W>
static int launch_split2_w(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	if constexpr (W == 1) {
		// one wave holds every field: 2, 4 or 8 of them
		switch (g.fields) {
		case 2: return launch_split2_q<PREC, 2, KIND, 1>(g, p, a, b, st, chunk);
		case 4: return launch_split2_q<PREC, 4, KIND, 1>(g, p, a, b, st, chunk);
		case 8: return launch_split2_q<PREC, 8, KIND, 1>(g, p, a, b, st, chunk);
		}
		return fail(MMSPX_EUNSUPPORTED, "the one-wave two-step strips take 2, 4 or 8 fields");
	} else {
		switch (g.fields) {
		case 8: return launch_split2_q<PREC, 8, KIND, W>(g, p, a, b, st, chunk);
		case 16: return launch_split2_q<PREC, 16, KIND, W>(g, p, a, b, st, chunk);
		case 32: return launch_split2_q<PREC, 32, KIND, W>(g, p, a, b, st, chunk);
		}
		return fail(MMSPX_EUNSUPPORTED, "split wave strips take 8, 16 or 32 fields");
	}
}

// Grain growth and Zener pinning, periodic 2-D: two steps per pass on the
// split strips, W = 2 (body 9) or 4 (body 10) waves per strip for 8, 16 or
// 32 fields, one wave (body 11) for 2, 4 or 8; a single step (an odd last
// one) runs the one-step split strips (W = 2, 4) or wave strips (W = 1)
template <int PREC>
static int launch_split2(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk,
                         int waves, int nsteps)
{
	if (!all_periodic(g)) return fail(MMSPX_EUNSUPPORTED, "the two-step split strips need periodic boundaries");
	if (nsteps == 1)
		return waves == 1 ? launch_wave<PREC>(g, p, a, b, st, chunk) : launch_split<PREC>(g, p, a, b, st, chunk, waves);
	if (nsteps != 2) return fail(MMSPX_EINVAL, "the two-step split strips take one or two steps per pass");
	const bool zener = p.model == MMSPX_ZENER;
	if constexpr (PREC == MMSPX_F32REF) {
		if (zener) return fail(MMSPX_EUNSUPPORTED, "Zener pinning runs the MMSPX_F64 and MMSPX_F32 forms");
		return waves == 1 ? launch_split2_w<PREC, 1, 1>(g, p, a, b, st, chunk)
		     : waves == 4 ? launch_split2_w<PREC, 1, 4>(g, p, a, b, st, chunk)
		                  : launch_split2_w<PREC, 1, 2>(g, p, a, b, st, chunk);
	} else {
		if (waves == 1)
			return zener ? launch_split2_w<PREC, 3, 1>(g, p, a, b, st, chunk) : launch_split2_w<PREC, 1, 1>(g, p, a, b, st, chunk);
		if (waves == 4)
			return zener ? launch_split2_w<PREC, 3, 4>(g, p, a, b, st, chunk) : launch_split2_w<PREC, 1, 4>(g, p, a, b, st, chunk);
		return zener ? launch_split2_w<PREC, 3, 2>(g, p, a, b, st, chunk) : launch_split2_w<PREC, 1, 2>(g, p, a, b, st, chunk);
	}
}

template <int PREC, int Q>
static int launch_ow_q(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	using S = typename Form<PREC>::S;
	const int strips = (g.n[1] + kChWaveOut - 1) / kChWaveOut;
	const int nitems = strips * ((g.n[0] + chunk - 1) / chunk);
	const int ntiles = (nitems + 3) / 4;
	const int blocks = (ntiles + 7) / 8 * 8;
	hipLaunchKernelGGL((ow_waves<PREC, Q>), dim3(blocks), dim3(256), 0, st, (const S*)a, (S*)b, g, p, chunk, strips,
	                   nitems, ntiles);
	MMSPX_HIP(hipGetLastError());
	return 0;
}

template <int PREC>
static int launch_ow(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, int chunk)
{
	if constexpr (PREC != MMSPX_F32REF) {
		switch (g.fields) {
		case 2: return launch_ow_q<PREC, 2>(g, p, a, b, st, chunk);
		case 3: return launch_ow_q<PREC, 3>(g, p, a, b, st, chunk);
		case 4: return launch_ow_q<PREC, 4>(g, p, a, b, st, chunk);
		case 5: return launch_ow_q<PREC, 5>(g, p, a, b, st, chunk);
		case 6: return launch_ow_q<PREC, 6>(g, p, a, b, st, chunk);
		case 7: return launch_ow_q<PREC, 7>(g, p, a, b, st, chunk);
		case 8: return launch_ow_q<PREC, 8>(g, p, a, b, st, chunk);
		}
	}
	return fail(MMSPX_EUNSUPPORTED, "Ostwald wave strips take 2 to 8 fields in the MMSPX_F64 and MMSPX_F32 forms");
}

template <int PREC>
static int launch_prec(const Geom& g, const ModelParams& p, const void* a, void* b, hipStream_t st, const TuneChoice& t,
                       int nsteps, void* aux)
{
	const int chunk = t.chunk > 0 ? t.chunk : 32;
	switch (p.model) {
	case MMSPX_CH: return launch_ch_prec<PREC, 0>(g, p, a, b, st, t, nsteps);
	case MMSPX_CH_DW: return launch_ch_prec<PREC, 1>(g, p, a, b, st, t, nsteps);
	case MMSPX_CH_NOISE: return launch_ch_prec<PREC, 2>(g, p, a, b, st, t, nsteps);
	case MMSPX_OSTWALD:
		if (g.dim == 3) {
			if constexpr (PREC == MMSPX_F32REF) return fail(MMSPX_EUNSUPPORTED, "no 3-D march for this form");
			else return t.opt == 12 ? launch_ow1<PREC>(g, p, a, b, st, t) : launch_ow3d<PREC>(g, p, a, b, aux, st, t);
		}
		if (g.dim == 2 && t.ty == 1 && t.tz == kChWaveOut) {
			if (!all_periodic(g)) return fail(MMSPX_EUNSUPPORTED, "Ostwald wave strips need periodic boundaries");
			return launch_ow<PREC>(g, p, a, b, st, chunk);
		}
		return fail(MMSPX_EUNSUPPORTED, "Ostwald ripening runs the wave strips (1 x 60) or the generic kernels");
	case MMSPX_ZENER:
		if (g.dim == 3) return t.opt == 12 ? launch_march2<PREC, 3>(g, p, a, b, st, t, nsteps)
		                                   : launch_march<PREC, 3>(g, p, a, b, st, t);
		if (t.opt >= 9 && t.opt <= 11 && t.ty == 1 && t.tz == kChWaveOut)
			return launch_split2<PREC>(g, p, a, b, st, chunk, t.opt == 11 ? 1 : t.opt == 10 ? 4 : 2, nsteps);
		if ((t.opt == 4 || t.opt == 5) && t.ty == 1 && t.tz == kWaveCols)
			return launch_split<PREC>(g, p, a, b, st, chunk, t.opt == 5 ? 4 : 2);
		if (t.opt == 2 && t.ty == 1 && t.tz == kWaveCols) return launch_wave<PREC>(g, p, a, b, st, chunk);
		return fail(MMSPX_EUNSUPPORTED, "Zener pinning runs the wave strips (1 x 64, body 2) or the generic kernel");
	case MMSPX_GG_ANISO:
		if (g.dim == 3) {
			if constexpr (PREC == MMSPX_F32REF) return fail(MMSPX_EUNSUPPORTED, "no 3-D march for this form");
			else return launch_march<PREC, 4>(g, p, a, b, st, t);
		}
		if (t.opt == 2 && t.ty == 1 && t.tz == kWaveCols) return launch_wave<PREC>(g, p, a, b, st, chunk);
		return fail(MMSPX_EUNSUPPORTED, "multi-phase grain growth runs the wave strips (1 x 64, body 2) or the generic kernel");
	case MMSPX_AC_NOISE:
		if (g.dim == 3) return launch_march<PREC, 2>(g, p, a, b, st, t);
		// model A: wave strips only (the noise is drawn per node and step)
		if (t.opt == 2 && t.ty == 1 && t.tz == kWaveCols) return launch_wave<PREC>(g, p, a, b, st, chunk);
		return fail(MMSPX_EUNSUPPORTED, "Allen-Cahn with noise runs the wave strips (1 x 64, body 2) or the generic kernel");
	case MMSPX_AC:
		if (g.dim == 3) return t.opt == 12 ? launch_march2<PREC, 0>(g, p, a, b, st, t, nsteps)
		                                   : launch_march<PREC, 0>(g, p, a, b, st, t);
		if (t.tz == kTbCols) {
			if (!all_periodic(g)) return fail(MMSPX_EUNSUPPORTED, "temporal blocks need periodic boundaries");
			if (nsteps < 1 || nsteps > kTbMaxSteps) return fail(MMSPX_EINVAL, "temporal blocks take 1..8 steps per pass");
			if (t.ty == 8) return launch_tb<PREC, 8>(g, p, a, b, st, nsteps);
			return t.ty == 32 ? launch_tb<PREC, 32>(g, p, a, b, st, nsteps) : launch_tb<PREC, 16>(g, p, a, b, st, nsteps);
		}
		if (t.opt == 2 && t.ty == 1 && t.tz == kWaveCols) return launch_wave<PREC>(g, p, a, b, st, chunk);
		if (t.tz == 64) return launch_ac<PREC, 64>(g, p, a, b, st, chunk);
		if (t.tz == 128) return launch_ac<PREC, 128>(g, p, a, b, st, chunk);
		return launch_ac<PREC, 256>(g, p, a, b, st, chunk);
	case MMSPX_GG:
		if (g.dim == 3) return t.opt == 12 ? launch_march2<PREC, 1>(g, p, a, b, st, t, nsteps)
		                                   : launch_march<PREC, 1>(g, p, a, b, st, t);
		if (t.opt >= 9 && t.opt <= 11 && t.ty == 1 && t.tz == kChWaveOut)
			return launch_split2<PREC>(g, p, a, b, st, chunk, t.opt == 11 ? 1 : t.opt == 10 ? 4 : 2, nsteps);
		// ty > 1: 2-D tiles of ty (8 or 16) rows x 64 columns; body 2 with 64
		// columns: wave strips; otherwise LDS-row marching strips
		if ((t.opt == 4 || t.opt == 5) && t.ty == 1 && t.tz == kWaveCols)
			return launch_split<PREC>(g, p, a, b, st, chunk, t.opt == 5 ? 4 : 2);
		if (t.opt == 2 && t.ty == 1 && t.tz == kWaveCols) return launch_wave<PREC>(g, p, a, b, st, chunk);
		if (t.ty > 1) return t.ty == 16 ? launch_gg_tile<PREC, 16>(g, p, a, b, st) : launch_gg_tile<PREC, 8>(g, p, a, b, st);
		if (t.tz == 64) return launch_gg_tz<PREC, 64>(g, p, a, b, st, chunk);
		if (t.tz == 256 && PREC != MMSPX_F64) return launch_gg_tz<PREC, 256>(g, p, a, b, st, chunk);
		return launch_gg_tz<PREC, 128>(g, p, a, b, st, chunk);
	}
	return fail(MMSPX_EINVAL, "unknown model");
}

bool fused_needs_plane(const Geom& g, int model, const TuneChoice& t)
{
	return model == MMSPX_OSTWALD && g.dim == 3 && t.opt != 12;
}

int fused_steps_per_launch(int model, const TuneChoice& t)
{
	if (model == MMSPX_AC && t.tz == kTbCols) return std::max(1, std::min(t.chunk, kTbMaxSteps));
	if ((model == MMSPX_CH || model == MMSPX_CH_DW) && (t.opt == 6 || t.opt == 7)) return 2;
	if ((model == MMSPX_GG || model == MMSPX_ZENER) && t.opt >= 9 && t.opt <= 11) return 2;
	if ((model == MMSPX_AC || model == MMSPX_GG || model == MMSPX_ZENER) && t.opt == 12) return 2;
	return 1;
}

int launch_fused_step(const Geom& g, int prec, const ModelParams& p, const void* a, void* b, hipStream_t st,
                      const TuneChoice& t, int nsteps, int64_t* launches, void* aux)
{
	int rc;
	switch (prec) {
	case MMSPX_F64: rc = launch_prec<MMSPX_F64>(g, p, a, b, st, t, nsteps, aux); break;
	case MMSPX_F32: rc = launch_prec<MMSPX_F32>(g, p, a, b, st, t, nsteps, aux); break;
	case MMSPX_F32REF: rc = launch_prec<MMSPX_F32REF>(g, p, a, b, st, t, nsteps, aux); break;
	default: return fail(MMSPX_EINVAL, "unknown precision");
	}
	if (rc == 0) *launches += 1;
	return rc;
}

}  // namespace mmspx
