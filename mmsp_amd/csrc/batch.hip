// batch.hip -- many independent grids advanced together (batched solves).
//
// The reference advances one grid per process (MMSP.main.hpp:278-350); serving
// many small solves that way leaves an MI355X idle. Here every member's tiles
// go into one device work list and one persistent launch per step walks it
// (workgroups take items blockIdx.x, blockIdx.x + gridDim.x, ...), whatever the
// members' extents. Multi-step runs are captured once into a hipGraph (one
// graph per starting parity, since each launch's arguments name the source
// buffer) and replayed. The tile bodies are those of the single-grid kernels
// (tiles.hpp), so each member's bits equal a single-grid run's.
//
// Members share model and precision. Supported: Allen-Cahn and grain growth
// on 2-D grids (any boundary condition), Cahn-Hilliard on periodic 2-D and 3-D
// grids (2-D and 3-D members go to two work lists, one launch each).
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "batch.hpp"

namespace mmspx {

template <int PREC, int TZ>
__global__ void __launch_bounds__(TZ) ac_batch(const BMember* __restrict__ mem, const BItem* __restrict__ items,
                                               int nitems, int parity)
{
	using S = typename Form<PREC>::S;
	for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
		const BItem w = items[it];
		const BMember& m = mem[w.m];
		const int src = (parity + m.phase) & 1;
		ac_strip<PREC, TZ>((const S*)m.buf[src], (S*)m.buf[src ^ 1], m.g, m.p, w.y0, w.xb, w.xe);
	}
}

// Allen-Cahn by runs of consecutive nodes (item: member m, first node y0,
// count z0), one workgroup per item, one thread per node as the generic
// kernel: neighbours through check_boundary's mapping, same arithmetic.
template <int PREC>
__global__ void __launch_bounds__(256) ac_batch_cells(const BMember* __restrict__ mem, const BItem* __restrict__ items,
                                                      int parity)
{
	using F = Form<PREC>;
	using S = typename F::S;
	const BItem w = items[blockIdx.x];
	const BMember& m = mem[w.m];
	const int src = (parity + m.phase) & 1;
	const S* __restrict__ a = (const S*)m.buf[src];
	S* __restrict__ b = (S*)m.buf[src ^ 1];
	const int n0 = m.g.n[0], n1 = m.g.n[1];
	const int end = w.y0 + w.z0;
	for (int i = w.y0 + (int)threadIdx.x; i < end; i += 256) {
		const int x = i / n1, y = i - x * n1;
		int xh = x + 1, xl = x - 1, yh = y + 1, yl = y - 1;
		if (xh >= n0) xh = bc_map(xh, n0, m.g.b0[0], m.g.b1[0]);
		if (xl < 0) xl = bc_map(xl, n0, m.g.b0[0], m.g.b1[0]);
		if (yh >= n1) yh = bc_map(yh, n1, m.g.b0[1], m.g.b1[1]);
		if (yl < 0) yl = bc_map(yl, n1, m.g.b0[1], m.g.b1[1]);
		const S c = a[i];
		typename F::L lap = F::lap_zero();
		F::lap_add(lap, m.g.w[0], m.g.wf[0], a[xh * n1 + y], c, a[xl * n1 + y]);
		F::lap_add(lap, m.g.w[1], m.g.wf[1], a[x * n1 + yh], c, a[x * n1 + yl]);
		b[i] = F::ac(c, lap, m.p);
	}
}

// Temporal blocks (periodic Allen-Cahn): item = member m, tile origin (y0
// rows, z0 columns) of a kTbBatchRows x kTbCols block; `steps` steps per pass.
constexpr int kTbBatchRows = 16;

template <int PREC>
__global__ void __launch_bounds__(256) ac_batch_tb(const BMember* __restrict__ mem, const BItem* __restrict__ items,
                                                   int nitems, int parity, int steps)
{
	using S = typename Form<PREC>::S;
	for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
		const BItem w = items[it];
		const BMember& m = mem[w.m];
		const int src = (parity + m.phase) & 1;
		ac_tblock<PREC, kTbBatchRows, kTbMaxSteps, 256>((const S*)m.buf[src], (S*)m.buf[src ^ 1], m.g, m.p, w.y0, w.z0,
		                                                steps);
	}
}

template <int PREC, int Q, int TZ>
__global__ void __launch_bounds__(TZ) gg_batch(const BMember* __restrict__ mem, const BItem* __restrict__ items,
                                               int nitems, int parity)
{
	using S = typename Form<PREC>::S;
	for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
		const BItem w = items[it];
		const BMember& m = mem[w.m];
		const int src = (parity + m.phase) & 1;
		gg_strip<PREC, Q, TZ>((const S*)m.buf[src], (S*)m.buf[src ^ 1], m.g, m.p, w.y0, w.xb, w.xe);
	}
}

template <int PREC, bool DW, int DIM, int TY, int TZ, int NT, bool RQ>
__global__ void __launch_bounds__(NT) ch_batch(const BMember* __restrict__ mem, const BItem* __restrict__ items,
                                               int nitems, int parity)
{
	using S = typename Form<PREC>::S;
	for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
		const BItem w = items[it];
		const BMember& m = mem[w.m];
		const int src = (parity + m.phase) & 1;
		if (RQ)
			ch_tile_rq<PREC, DW, DIM, TY, TZ, NT>((const S*)m.buf[src], (S*)m.buf[src ^ 1], m.g, m.p, w.y0, w.z0, w.xb,
			                                      w.xe);
		else
			ch_tile<PREC, DW, DIM, TY, TZ, NT>((const S*)m.buf[src], (S*)m.buf[src ^ 1], m.g, m.p, w.y0, w.z0, w.xb, w.xe);
	}
}

// 2-D Cahn-Hilliard by wave strips: one wave per item (member m, direction
// y0: 1 marches downward, first column z0, rows [xb, xe)); four waves per
// workgroup walk the list independently.
template <int PREC, bool DW>
__global__ void __launch_bounds__(256) ch_batch_wave(const BMember* __restrict__ mem, const BItem* __restrict__ items,
                                                     int nitems, int parity)
{
	using S = typename Form<PREC>::S;
	for (int it = blockIdx.x * 4 + (int)(threadIdx.x >> 6); it < nitems; it += gridDim.x * 4) {
		const BItem w = items[it];
		const BMember& m = mem[w.m];
		const int src = (parity + m.phase) & 1;
		ch_wave<PREC, DW>((const S*)m.buf[src], (S*)m.buf[src ^ 1], m.g, m.p, w.z0, w.xb, w.xe, w.y0 != 0);
	}
}

// tile shapes of the batched Cahn-Hilliard kernels (fixed; the strips' width
// and chunk depth are what the tuner varies). 3-D members run the
// register-queue body, the faster one at 16 x 64 in both precisions.
constexpr int kCh3TY = 16, kCh3TZ = 64, kCh3NT = 256;

bool batch_model_ok(int model, const Geom& g, std::string* why)
{
	auto no = [&](const char* m) {
		if (why) *why = m;
		return false;
	};
	switch (model) {
	case MMSPX_AC:
		if (g.dim != 2) return no("batched Allen-Cahn needs 2-D members");
		return true;
	case MMSPX_GG:
		if (g.dim != 2) return no("batched grain growth needs 2-D members");
		if (!(g.fields == 2 || g.fields == 4 || g.fields == 8 || g.fields == 16 || g.fields == 32))
			return no("batched grain growth needs 2, 4, 8, 16 or 32 fields");
		return true;
	case MMSPX_CH:
	case MMSPX_CH_DW:
		if (g.dim < 2) return no("batched Cahn-Hilliard needs 2-D or 3-D members");
		for (int k = 0; k < g.dim; k++)
			if (g.b0[k] != MMSPX_PERIODIC || g.b1[k] != MMSPX_PERIODIC)
				return no("batched Cahn-Hilliard needs periodic members");
		return true;
	case MMSPX_AC_NOISE:
	case MMSPX_CH_NOISE:
		return no("batches do not run the stochastic forms (each member would need its own step count)");
	case MMSPX_ZENER:
	case MMSPX_OSTWALD:
	case MMSPX_GG_ANISO:
		return no("batches do not run Zener pinning, Ostwald ripening or multi-phase grain growth");
	}
	return no("unknown model");
}

// Work items of one member: strips/tiles of the last axes times chunks of axis 0.
void member_items(int m, const Geom& g, int model, const BatchTile& t, std::vector<BItem>& out2,
                  std::vector<BItem>& out3)
{
	const int nx = g.n[0];
	if ((model == MMSPX_CH || model == MMSPX_CH_DW) && g.dim == 3) {
		const int chunk = t.chunk3;
		for (int xb = 0; xb < nx; xb += chunk)
			for (int y0 = 0; y0 < g.n[1]; y0 += kCh3TY)
				for (int z0 = 0; z0 < g.n[2]; z0 += kCh3TZ)
					out3.push_back(BItem{m, y0, z0, xb, std::min(xb + chunk, nx)});
		return;
	}
	// 2-D: strips of the contiguous axis (axis 1). The strip bodies take the
	// strip origin as y0; the 2-D Cahn-Hilliard tile names that axis z.
	const bool ch2 = model == MMSPX_CH || model == MMSPX_CH_DW;
	if (model == MMSPX_AC && t.tb > 0) {
		for (int x0 = 0; x0 < nx; x0 += kTbBatchRows)
			for (int c0 = 0; c0 < g.n[1]; c0 += kTbCols) out2.push_back(BItem{m, x0, c0, 0, 0});
		return;
	}
	if (model == MMSPX_AC && t.cells > 0) {
		for (int f = 0; f < g.nodes; f += t.cells) out2.push_back(BItem{m, f, std::min(t.cells, g.nodes - f), 0, 0});
		return;
	}
	if (ch2 && t.tz == kChWaveOut) {
		for (int xb = 0, k = 0; xb < nx; xb += t.chunk, k++)
			for (int c0 = 0; c0 < g.n[1]; c0 += kChWaveOut)
				out2.push_back(BItem{m, k & 1, c0, xb, std::min(xb + t.chunk, nx)});
		return;
	}
	for (int xb = 0; xb < nx; xb += t.chunk)
		for (int c0 = 0; c0 < g.n[1]; c0 += t.tz) {
			const int xe = std::min(xb + t.chunk, nx);
			out2.push_back(ch2 ? BItem{m, 0, c0, xb, xe} : BItem{m, c0, 0, xb, xe});
		}
}

template <int PREC, int TZ>
static void launch_strips(int model, int fields, const BMember* mem, const BItem* items, int n, int parity, int blocks,
                          hipStream_t st)
{
	if (model == MMSPX_AC) {
		hipLaunchKernelGGL((ac_batch<PREC, TZ>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity);
	} else if (model == MMSPX_GG) {
		switch (fields) {
		case 2: hipLaunchKernelGGL((gg_batch<PREC, 2, TZ>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity); break;
		case 4: hipLaunchKernelGGL((gg_batch<PREC, 4, TZ>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity); break;
		case 8: hipLaunchKernelGGL((gg_batch<PREC, 8, TZ>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity); break;
		case 16: hipLaunchKernelGGL((gg_batch<PREC, 16, TZ>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity); break;
		default: hipLaunchKernelGGL((gg_batch<PREC, 32, TZ>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity); break;
		}
	} else {
		const bool dw = model == MMSPX_CH_DW;
		if (dw)
			hipLaunchKernelGGL((ch_batch<PREC, true, 2, 1, TZ, TZ, false>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity);
		else
			hipLaunchKernelGGL((ch_batch<PREC, false, 2, 1, TZ, TZ, false>), dim3(blocks), dim3(TZ), 0, st, mem, items, n, parity);
	}
}

template <int PREC>
static void launch_ch2_waves(int model, const BMember* mem, const BItem* items, int n, int parity, int blocks,
                             hipStream_t st)
{
	if (model == MMSPX_CH_DW)
		hipLaunchKernelGGL((ch_batch_wave<PREC, true>), dim3(blocks), dim3(256), 0, st, mem, items, n, parity);
	else
		hipLaunchKernelGGL((ch_batch_wave<PREC, false>), dim3(blocks), dim3(256), 0, st, mem, items, n, parity);
}

template <int PREC>
static void launch_ch3(bool dw, const BMember* mem, const BItem* items, int n, int parity, int blocks, hipStream_t st)
{
	if (dw)
		hipLaunchKernelGGL((ch_batch<PREC, true, 3, kCh3TY, kCh3TZ, kCh3NT, true>), dim3(blocks), dim3(kCh3NT), 0, st, mem, items,
		                   n, parity);
	else
		hipLaunchKernelGGL((ch_batch<PREC, false, 3, kCh3TY, kCh3TZ, kCh3NT, true>), dim3(blocks), dim3(kCh3NT), 0, st, mem,
		                   items, n, parity);
}

template <int PREC>
static void launch_prec(int model, int fields, const BatchTile& t, const BMember* mem, const BItem* l2, int n2,
                        const BItem* l3, int n3, int parity, int cus, hipStream_t st, int nsteps)
{
	if (n2 > 0 && model == MMSPX_AC && t.tb > 0) {
		const int blocks = std::max(1, std::min(n2, cus * t.bpc));
		hipLaunchKernelGGL((ac_batch_tb<PREC>), dim3(blocks), dim3(256), 0, st, mem, l2, n2, parity, nsteps);
	} else if (n2 > 0 && model == MMSPX_AC && t.cells > 0) {
		hipLaunchKernelGGL((ac_batch_cells<PREC>), dim3(n2), dim3(256), 0, st, mem, l2, parity);
	} else if (n2 > 0 && (model == MMSPX_CH || model == MMSPX_CH_DW) && t.tz == kChWaveOut) {
		const int blocks = std::max(1, std::min((n2 + 3) / 4, cus * t.bpc));
		launch_ch2_waves<PREC>(model, mem, l2, n2, parity, blocks, st);
	} else if (n2 > 0) {
		const int blocks = std::max(1, std::min(n2, cus * t.bpc));
		if (t.tz == 64)
			launch_strips<PREC, 64>(model, fields, mem, l2, n2, parity, blocks, st);
		else if (t.tz == 128)
			launch_strips<PREC, 128>(model, fields, mem, l2, n2, parity, blocks, st);
		else
			launch_strips<PREC, 256>(model, fields, mem, l2, n2, parity, blocks, st);
	}
	if (n3 > 0) {
		const int blocks = std::max(1, std::min(n3, cus * t.bpc));
		launch_ch3<PREC>(model == MMSPX_CH_DW, mem, l3, n3, parity, blocks, st);
	}
}

int launch_batch_step(int model, int prec, int fields, const BatchTile& t, const BMember* mem, const BItem* l2, int n2,
                      const BItem* l3, int n3, int parity, int cus, hipStream_t st, int64_t* launches, int nsteps)
{
	if (nsteps < 1 || nsteps > std::max(1, std::min(t.tb, kTbMaxSteps)))
		return fail(MMSPX_EINVAL, "steps per pass exceed what the batch tile takes");
	switch (prec) {
	case MMSPX_F64: launch_prec<MMSPX_F64>(model, fields, t, mem, l2, n2, l3, n3, parity, cus, st, nsteps); break;
	case MMSPX_F32: launch_prec<MMSPX_F32>(model, fields, t, mem, l2, n2, l3, n3, parity, cus, st, nsteps); break;
	case MMSPX_F32REF: launch_prec<MMSPX_F32REF>(model, fields, t, mem, l2, n2, l3, n3, parity, cus, st, nsteps); break;
	default: return fail(MMSPX_EINVAL, "unknown precision");
	}
	MMSPX_HIP(hipGetLastError());
	*launches += (n2 > 0) + (n3 > 0);
	return 0;
}

std::vector<BItem> xcd_order(const std::vector<BItem>& in)
{
	const size_t n = in.size();
	std::vector<BItem> out;
	out.reserve(n);
	size_t start[9];
	for (int k = 0; k <= 8; k++) start[k] = n * k / 8;
	const size_t longest = start[1] - start[0] + 1;
	for (size_t j = 0; j < longest; j++)
		for (int k = 0; k < 8; k++)
			if (start[k] + j < start[k + 1]) out.push_back(in[start[k] + j]);
	return out;
}

int device_cus()
{
	int dev = 0, cus = 0;
	if (hipGetDevice(&dev) != hipSuccess) return 256;
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 256;
	return cus;
}

}  // namespace mmspx

using namespace mmspx;

struct mmspx_batch {
	int count = 0, model = 0, prec = 0, fields = 1;
	std::vector<mmspx_desc> desc;
	std::vector<Geom> geom;
	std::vector<size_t> off;   // byte offset of each member's first buffer in the arena
	std::vector<size_t> bytes; // one buffer
	char* arena = nullptr;
	size_t arena_bytes = 0;
	std::vector<ModelParams> params;
	bool params_set = false;
	BMember* d_mem = nullptr;
	BItem* d_items = nullptr;
	int n2 = 0, n3 = 0, items_cap = 0;
	BatchTile tile{256, 32, 16, 8, 0};
	bool tuned = false;
	int cus = 256;
	int64_t steps_done = 0, launches = 0;
	int64_t passes_done = 0;  // launches per item list so far: the buffer parity
	hipStream_t st = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hipGraphExec_t graph[2] = {nullptr, nullptr};
	int graph_steps = 0;
	double* part = nullptr;
	double* part_host = nullptr;
	void* staging = nullptr;
	size_t staging_bytes = 0;
};

static void drop_graphs(mmspx_batch* b)
{
	for (auto& gx : b->graph)
		if (gx) {
			(void)hipGraphExecDestroy(gx);
			gx = nullptr;
		}
	b->graph_steps = 0;
}

// (Re)builds the device member table and work lists for the current tile.
static int upload_plan(mmspx_batch* b)
{
	std::vector<BItem> l2, l3;
	std::vector<BMember> mem(b->count);
	for (int m = 0; m < b->count; m++) {
		mem[m].buf[0] = b->arena + b->off[m];
		mem[m].buf[1] = b->arena + b->off[m] + b->bytes[m];
		mem[m].g = b->geom[m];
		mem[m].p = b->params[m];
		mem[m].phase = 0;
		member_items(m, b->geom[m], b->model, b->tile, l2, l3);
	}
	const int need = (int)(l2.size() + l3.size());
	if (need > b->items_cap) {
		if (b->d_items) (void)hipFree(b->d_items);
		b->d_items = nullptr;
		if (hipMalloc((void**)&b->d_items, sizeof(BItem) * (size_t)need) != hipSuccess)
			return fail(MMSPX_ENOMEM, "device allocation failed (work list)");
		b->items_cap = need;
	}
	std::vector<BItem> all = xcd_order(l2);
	const std::vector<BItem> o3 = xcd_order(l3);
	all.insert(all.end(), o3.begin(), o3.end());
	MMSPX_HIP(hipMemcpyAsync(b->d_items, all.data(), sizeof(BItem) * all.size(), hipMemcpyHostToDevice, b->st));
	MMSPX_HIP(hipMemcpyAsync(b->d_mem, mem.data(), sizeof(BMember) * mem.size(), hipMemcpyHostToDevice, b->st));
	MMSPX_HIP(hipStreamSynchronize(b->st));
	b->n2 = (int)l2.size();
	b->n3 = (int)l3.size();
	drop_graphs(b);
	return 0;
}

static int steps_per_pass(const mmspx_batch* b) { return b->tile.tb > 0 ? std::min(b->tile.tb, kTbMaxSteps) : 1; }

static int enqueue_steps(mmspx_batch* b, int steps)
{
	const int spp = steps_per_pass(b);
	for (int left = steps; left > 0;) {
		const int n = std::min(left, spp);
		const int parity = (int)(b->passes_done & 1);
		int rc = launch_batch_step(b->model, b->prec, b->fields, b->tile, b->d_mem, b->d_items, b->n2,
		                           b->d_items + b->n2, b->n3, parity, b->cus, b->st, &b->launches, n);
		if (rc) return rc;
		b->passes_done += 1;
		b->steps_done += n;
		left -= n;
	}
	return 0;
}

static int graph_passes(const mmspx_batch* b, int gsteps) { return (gsteps + steps_per_pass(b) - 1) / steps_per_pass(b); }

static int build_graphs(mmspx_batch* b, int gsteps)
{
	drop_graphs(b);
	const int spp = steps_per_pass(b);
	for (int start = 0; start < 2; start++) {
		hipGraph_t graph = nullptr;
		MMSPX_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
		int64_t dummy = 0;
		int rc = 0;
		int k = 0;
		for (int left = gsteps; left > 0 && rc == 0; k++) {
			const int n = std::min(left, spp);
			rc = launch_batch_step(b->model, b->prec, b->fields, b->tile, b->d_mem, b->d_items, b->n2, b->d_items + b->n2,
			                       b->n3, (start + k) & 1, b->cus, b->st, &dummy, n);
			left -= n;
		}
		hipError_t e = hipStreamEndCapture(b->st, &graph);
		if (rc) {
			if (graph) (void)hipGraphDestroy(graph);
			return rc;
		}
		if (e != hipSuccess) return fail(MMSPX_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
		e = hipGraphInstantiate(&b->graph[start], graph, nullptr, nullptr, 0);
		(void)hipGraphDestroy(graph);
		if (e != hipSuccess) return fail(MMSPX_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
	}
	b->graph_steps = gsteps;
	return 0;
}

static int batch_step_impl(mmspx_batch* b, int steps, int gsteps)
{
	if (!b || steps < 0 || gsteps < 0) return fail(MMSPX_EINVAL, "bad arguments");
	if (!b->params_set) return fail(MMSPX_EINVAL, "mmspx_batch_set_params has not been called");
	if (gsteps > 0 && steps >= gsteps) {
		if (b->graph_steps != gsteps) {
			int rc = build_graphs(b, gsteps);
			if (rc) return rc;
		}
		const int reps = steps / gsteps;
		const int gp = graph_passes(b, gsteps);
		for (int r = 0; r < reps; r++) {
			MMSPX_HIP(hipGraphLaunch(b->graph[b->passes_done & 1], b->st));
			b->steps_done += gsteps;
			b->passes_done += gp;
			b->launches += (int64_t)gp * ((b->n2 > 0) + (b->n3 > 0));
		}
		steps -= reps * gsteps;
	}
	return enqueue_steps(b, steps);
}

extern "C" {

void mmspx_batch_destroy(mmspx_batch* b)
{
	if (!b) return;
	if (b->st) (void)hipStreamSynchronize(b->st);
	drop_graphs(b);
	if (b->arena) (void)hipFree(b->arena);
	if (b->d_mem) (void)hipFree(b->d_mem);
	if (b->d_items) (void)hipFree(b->d_items);
	if (b->part) (void)hipFree(b->part);
	if (b->part_host) (void)hipHostFree(b->part_host);
	if (b->staging) (void)hipFree(b->staging);
	if (b->e0) (void)hipEventDestroy(b->e0);
	if (b->e1) (void)hipEventDestroy(b->e1);
	if (b->st) (void)hipStreamDestroy(b->st);
	delete b;
}

int mmspx_batch_create(const mmspx_desc* descs, int count, int model, mmspx_batch** out)
{
	if (!out) return fail(MMSPX_EINVAL, "null output pointer");
	*out = nullptr;
	if (!descs || count < 1 || count > (1 << 20)) return fail(MMSPX_EINVAL, "bad member count");
	auto* b = new (std::nothrow) mmspx_batch();
	if (!b) return fail(MMSPX_ENOMEM, "host allocation failed");
	b->count = count;
	b->model = model;
	b->prec = descs[0].precision;
	size_t total = 0;
	for (int m = 0; m < count; m++) {
		int rc = validate_desc(&descs[m]);
		if (rc) {
			delete b;
			return rc;
		}
		if (descs[m].precision != b->prec) {
			delete b;
			return fail(MMSPX_EINVAL, "batch members must share one precision");
		}
		Geom g = make_geom(descs[m]);
		std::string why;
		if (!batch_model_ok(model, g, &why)) {
			delete b;
			return fail(MMSPX_EUNSUPPORTED, why);
		}
		if (m > 0 && g.fields != b->fields) {
			delete b;
			return fail(MMSPX_EINVAL, "batch members must share one field count");
		}
		b->fields = g.fields;
		b->desc.push_back(descs[m]);
		b->geom.push_back(g);
		const size_t bytes = ((size_t)g.nodes * g.fields * (b->prec == MMSPX_F64 ? 8 : 4) + 255) / 256 * 256;
		b->off.push_back(total);
		b->bytes.push_back(bytes);
		total += 2 * bytes;
	}
	b->params.assign(count, ModelParams{});
	// Cahn-Hilliard: 2-D members by wave strips (faster than the LDS row strips)
	if (model == MMSPX_CH || model == MMSPX_CH_DW) b->tile = BatchTile{kChWaveOut, 8, 16, 8, 0};
	b->cus = device_cus();
	b->arena_bytes = total;
	const size_t red = (size_t)b->fields * 2 * kRedBlocks * sizeof(double);
	int rc = 0;
	if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&b->e0) != hipSuccess ||
	    hipEventCreate(&b->e1) != hipSuccess)
		rc = fail(MMSPX_EDEVICE, "HIP stream/event setup failed");
	if (rc == 0 && (hipMalloc((void**)&b->arena, total) != hipSuccess ||
	                hipMalloc((void**)&b->d_mem, sizeof(BMember) * (size_t)count) != hipSuccess ||
	                hipMalloc((void**)&b->part, red) != hipSuccess ||
	                hipHostMalloc((void**)&b->part_host, red, 0) != hipSuccess))
		rc = fail(MMSPX_ENOMEM, "device allocation failed");
	if (rc == 0) rc = upload_plan(b);
	if (rc) {
		mmspx_batch_destroy(b);
		return rc;
	}
	*out = b;
	return 0;
}

static void* cur_buf(mmspx_batch* b, int m) { return b->arena + b->off[m] + ((b->passes_done & 1) ? b->bytes[m] : 0); }

int mmspx_batch_upload(mmspx_batch* b, int m, const void* host, int layout)
{
	if (!b || !host || m < 0 || m >= b->count) return fail(MMSPX_EINVAL, "bad member");
	const Geom& g = b->geom[m];
	const size_t n = (size_t)g.nodes * g.fields * (b->prec == MMSPX_F64 ? 8 : 4);
	void* dst = cur_buf(b, m);
	if (g.fields == 1 || layout == MMSPX_FIELD_MAJOR) {
		MMSPX_HIP(hipMemcpyAsync(dst, host, n, hipMemcpyHostToDevice, b->st));
	} else {
		void* spare = b->arena + b->off[m] + ((b->passes_done & 1) ? 0 : b->bytes[m]);
		MMSPX_HIP(hipMemcpyAsync(spare, host, n, hipMemcpyHostToDevice, b->st));
		int rc = launch_transpose(spare, dst, g.nodes, g.fields, b->prec == MMSPX_F64 ? 8 : 4, true, b->st);
		if (rc) return rc;
	}
	MMSPX_HIP(hipStreamSynchronize(b->st));
	return 0;
}

int mmspx_batch_download(mmspx_batch* b, int m, void* host, int layout)
{
	if (!b || !host || m < 0 || m >= b->count) return fail(MMSPX_EINVAL, "bad member");
	const Geom& g = b->geom[m];
	const size_t n = (size_t)g.nodes * g.fields * (b->prec == MMSPX_F64 ? 8 : 4);
	void* src = cur_buf(b, m);
	if (g.fields == 1 || layout == MMSPX_FIELD_MAJOR) {
		MMSPX_HIP(hipMemcpyAsync(host, src, n, hipMemcpyDeviceToHost, b->st));
	} else {
		void* spare = b->arena + b->off[m] + ((b->passes_done & 1) ? 0 : b->bytes[m]);
		int rc = launch_transpose(src, spare, g.nodes, g.fields, b->prec == MMSPX_F64 ? 8 : 4, false, b->st);
		if (rc) return rc;
		MMSPX_HIP(hipMemcpyAsync(host, spare, n, hipMemcpyDeviceToHost, b->st));
	}
	MMSPX_HIP(hipStreamSynchronize(b->st));
	return 0;
}

int mmspx_batch_set_params(mmspx_batch* b, const double* params, int np)
{
	if (!b || !params) return fail(MMSPX_EINVAL, "null argument");
	for (int m = 0; m < b->count; m++) {
		int rc = make_params(b->model, params + (size_t)m * np, np, &b->params[m]);
		if (rc) return rc;
	}
	b->params_set = true;
	return upload_plan(b);
}

int mmspx_batch_step(mmspx_batch* b, int steps, int graph_steps) { return batch_step_impl(b, steps, graph_steps); }

int mmspx_batch_step_timed(mmspx_batch* b, int steps, int graph_steps, double* seconds)
{
	if (!b) return fail(MMSPX_EINVAL, "null batch");
	MMSPX_HIP(hipEventRecord(b->e0, b->st));
	int rc = batch_step_impl(b, steps, graph_steps);
	if (rc) return rc;
	MMSPX_HIP(hipEventRecord(b->e1, b->st));
	MMSPX_HIP(hipEventSynchronize(b->e1));
	float ms = 0.0f;
	MMSPX_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
	if (seconds) *seconds = ms * 1e-3;
	return 0;
}

int mmspx_batch_reduce(mmspx_batch* b, int m, int what, double* out, int nout)
{
	if (!b || !out || m < 0 || m >= b->count) return fail(MMSPX_EINVAL, "bad member");
	const Geom& g = b->geom[m];
	if (what == MMSPX_MASS && nout < g.fields) return fail(MMSPX_EINVAL, "mass needs one output per field");
	if (what == MMSPX_ENERGY && nout < 1) return fail(MMSPX_EINVAL, "energy needs one output");
	if (what != MMSPX_MASS && what != MMSPX_ENERGY) return fail(MMSPX_EINVAL, "unknown reduction");
	int rc = launch_reduce(g, b->prec, what, b->params[m], cur_buf(b, m), b->part, b->st);
	if (rc) return rc;
	const int sets = what == MMSPX_MASS ? g.fields : 1;
	MMSPX_HIP(hipMemcpyAsync(b->part_host, b->part, (size_t)sets * 2 * kRedBlocks * sizeof(double),
	                         hipMemcpyDeviceToHost, b->st));
	MMSPX_HIP(hipStreamSynchronize(b->st));
	for (int j = 0; j < sets; j++) out[j] = sum_partials(b->part_host + (size_t)j * 2 * kRedBlocks, kRedBlocks);
	if (what == MMSPX_ENERGY) out[0] *= g.vol;
	return 0;
}

static bool all_members_periodic(const mmspx_batch* b)
{
	for (const Geom& g : b->geom)
		for (int k = 0; k < g.dim; k++)
			if (g.b0[k] != MMSPX_PERIODIC || g.b1[k] != MMSPX_PERIODIC) return false;
	return true;
}

int mmspx_batch_set_tile(mmspx_batch* b, int tz, int chunk, int blocks_per_cu, int cells)
{
	if (!b) return fail(MMSPX_EINVAL, "null batch");
	if (tz == kTbCols) {  // temporal blocks: chunk = steps per pass
		if (b->model != MMSPX_AC || !all_members_periodic(b) || chunk < 1 || chunk > kTbMaxSteps || blocks_per_cu < 1)
			return fail(MMSPX_EINVAL, "temporal blocks: periodic Allen-Cahn members, 1..8 steps per pass");
		BatchTile t{tz, 32, b->tile.chunk3, blocks_per_cu, 0};
		t.tb = chunk;
		b->tile = t;
		return upload_plan(b);
	}
	const bool ch = b->model == MMSPX_CH || b->model == MMSPX_CH_DW;
	if ((tz != 64 && tz != 128 && tz != 256 && !(ch && tz == kChWaveOut)) || chunk < 1 || blocks_per_cu < 1 ||
	    cells < 0 || (cells > 0 && b->model != MMSPX_AC))
		return fail(MMSPX_EINVAL, "strip width 64/128/256 (60: Cahn-Hilliard wave strips), positive rows and "
		                          "workgroups; cell runs for Allen-Cahn only");
	b->tile = BatchTile{tz, chunk, b->tile.chunk3, blocks_per_cu, cells};
	return upload_plan(b);
}

int mmspx_batch_get_tile(const mmspx_batch* b, int* tz, int* chunk, int* blocks_per_cu)
{
	if (!b) return fail(MMSPX_EINVAL, "null batch");
	const BatchTile& t = b->tile;
	if (tz) *tz = t.tb > 0 ? kTbCols : t.cells > 0 ? 0 : t.tz;
	if (chunk) *chunk = t.tb > 0 ? t.tb : t.cells > 0 ? t.cells : t.chunk;
	if (blocks_per_cu) *blocks_per_cu = t.cells > 0 ? 0 : t.bpc;
	return b->tuned ? 1 : 0;
}

int mmspx_batch_autotune(mmspx_batch* b, int trial_steps, char* report, size_t report_len)
{
	if (!b || trial_steps < 1) return fail(MMSPX_EINVAL, "bad arguments");
	if (!b->params_set) return fail(MMSPX_EINVAL, "mmspx_batch_set_params has not been called");
	// every trial starts from the same state, restored afterwards
	if (b->staging_bytes < b->arena_bytes) {
		if (b->staging) (void)hipFree(b->staging);
		b->staging = nullptr;
		if (hipMalloc(&b->staging, b->arena_bytes) != hipSuccess) return fail(MMSPX_ENOMEM, "device allocation failed");
		b->staging_bytes = b->arena_bytes;
	}
	MMSPX_HIP(hipMemcpyAsync(b->staging, b->arena, b->arena_bytes, hipMemcpyDeviceToDevice, b->st));
	const int64_t steps0 = b->steps_done, passes0 = b->passes_done;
	const BatchTile keep = b->tile;
	std::vector<BatchTile> cand;
	for (int tz : {64, 128, 256})
		for (int chunk : {16, 32, 64, 128})
			for (int bpc : {4, 8, 16}) cand.push_back(BatchTile{tz, chunk, 16, bpc, 0});
	if (b->model == MMSPX_CH || b->model == MMSPX_CH_DW)
		for (int chunk : {8, 16, 32})
			for (int bpc : {4, 8}) cand.push_back(BatchTile{kChWaveOut, chunk, 16, bpc, 0});
	if (b->model == MMSPX_AC)
		for (int cells : {256, 512, 1024, 2048, 4096}) cand.push_back(BatchTile{256, 32, 16, 8, cells});
	if (b->model == MMSPX_AC && all_members_periodic(b))
		for (int tb : {4, 8})
			for (int bpc : {4, 8, 16}) {
				BatchTile t{kTbCols, 32, 16, bpc, 0};
				t.tb = tb;
				cand.push_back(t);
			}
	std::string rep;
	double best = 1e30;
	BatchTile best_t = keep;
	int rc = 0;
	for (const BatchTile& t : cand) {
		b->tile = t;
		rc = upload_plan(b);
		// temporal blocks are timed over whole passes
		const int steps = t.tb > 0 ? std::max(trial_steps, 2 * t.tb) / t.tb * t.tb : trial_steps;
		if (rc == 0) rc = enqueue_steps(b, 2);
		double sec = 0.0;
		if (rc == 0) rc = mmspx_batch_step_timed(b, steps, 0, &sec);
		if (rc) break;
		sec = sec / steps * trial_steps;
		char line[160];
		std::snprintf(line, sizeof line, "tz=%d chunk=%d blocks_per_cu=%d cells=%d steps_per_pass=%d us_per_step=%.2f\n",
		              t.tz, t.chunk, t.bpc, t.cells, t.tb, sec / trial_steps * 1e6);
		rep += line;
		if (sec < best) {
			best = sec;
			best_t = t;
		}
	}
	b->tile = rc ? keep : best_t;
	b->steps_done = steps0;
	b->passes_done = passes0;
	int rc2 = upload_plan(b);
	if (hipMemcpyAsync(b->arena, b->staging, b->arena_bytes, hipMemcpyDeviceToDevice, b->st) != hipSuccess ||
	    hipStreamSynchronize(b->st) != hipSuccess)
		return fail(MMSPX_EDEVICE, "restoring the batch after tuning failed");
	if (rc) return rc;
	if (rc2) return rc2;
	b->tuned = true;
	if (report && report_len) std::snprintf(report, report_len, "%s", rep.c_str());
	return 0;
}

}  // extern "C"
