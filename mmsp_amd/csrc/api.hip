// api.hip -- the grid-state half of the C ABI (include/mmsp_amd.h).
//
// A state is the device-resident replacement of grid<dim,T> storage
// (MMSP.grid.cpp:141-327) together with the newGrid/temp grids that each
// reference update() copy-constructs (allen-cahn.cpp:56, cahn-hilliard.cpp:56-57):
// two state buffers swapped by pointer after every step (grid::swap,
// MMSP.grid.cpp:1913-1991) plus, for Cahn-Hilliard, a chemical-potential
// buffer. Vector grids are stored field-major (one plane per field).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <vector>

#include "core.hpp"

namespace mmspx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
	set_error(msg);
	return code;
}

int validate_desc(const mmspx_desc* d)
{
	if (!d) return fail(MMSPX_EINVAL, "null descriptor");
	if (d->dim < 1 || d->dim > 3) return fail(MMSPX_EINVAL, "dim must be 1, 2 or 3");
	if (d->precision != MMSPX_F64 && d->precision != MMSPX_F32 && d->precision != MMSPX_F32REF)
		return fail(MMSPX_EINVAL, "unknown precision");
	const int q = d->vector ? d->fields : 1;
	if (q < 1 || q > 4096) return fail(MMSPX_EINVAL, "fields must be in [1, 4096] for vector grids");
	int64_t nodes = 1;
	for (int k = 0; k < d->dim; k++) {
		if (d->n[k] < 1) return fail(MMSPX_EINVAL, "every extent must be >= 1");
		for (int side = 0; side < 2; side++) {
			int b = side ? d->b1[k] : d->b0[k];
			if (b != MMSPX_MIRROR && b != MMSPX_NEUMANN && b != MMSPX_PERIODIC && b != MMSPX_DIRICHLET)
				return fail(MMSPX_EINVAL, "boundary condition must be mirror, Neumann, periodic or Dirichlet "
				                          "(the serial reference has no 'parallel' faces)");
			if (b == MMSPX_MIRROR && d->n[k] < 2) return fail(MMSPX_EINVAL, "mirror boundary needs extent >= 2");
		}
		if (!(d->dx[k] > 0.0) || !std::isfinite(d->dx[k])) return fail(MMSPX_EINVAL, "dx must be positive");
		nodes *= d->n[k];
	}
	if (nodes * q > 0x7fffffffLL) return fail(MMSPX_EINVAL, "nodes * fields must stay below 2^31");
	return 0;
}

int validate_model(const mmspx_desc& d, const Geom& g, int model)
{
	if (vector_model(model)) {
		if (!d.vector) return fail(MMSPX_EINVAL, "grain growth, Zener pinning and Ostwald ripening need a vector grid");
		if (model == MMSPX_OSTWALD && g.fields < 2)
			return fail(MMSPX_EINVAL, "Ostwald ripening needs a concentration field and at least one order parameter");
		if ((model == MMSPX_ZENER || model == MMSPX_OSTWALD || model == MMSPX_GG_ANISO) &&
		    d.precision == MMSPX_F32REF)
			return fail(MMSPX_EUNSUPPORTED, "this model runs the MMSPX_F64 and MMSPX_F32 forms");
		if (model == MMSPX_GG_ANISO && g.fields > kAnisoMaxFields)
			return fail(MMSPX_EUNSUPPORTED, "anisotropic grain growth takes at most 32 fields");
	} else if (d.vector) {
		return fail(MMSPX_EINVAL, "Allen-Cahn and Cahn-Hilliard need a scalar grid");
	}
	return 0;
}

Geom make_geom(const mmspx_desc& d)
{
	Geom g;
	std::memset(&g, 0, sizeof(g));
	g.dim = d.dim;
	g.nodes = 1;
	g.vol = 1.0;
	for (int k = 0; k < 3; k++) {
		g.n[k] = k < d.dim ? d.n[k] : 1;
		g.b0[k] = k < d.dim ? d.b0[k] : MMSPX_PERIODIC;
		g.b1[k] = k < d.dim ? d.b1[k] : MMSPX_PERIODIC;
		g.dx[k] = k < d.dim ? d.dx[k] : 1.0;
		g.w[k] = 1.0 / (g.dx[k] * g.dx[k]);
		g.wf[k] = (float)g.w[k];
	}
	for (int k = 0; k < d.dim; k++) {
		g.nodes *= g.n[k];
		g.vol *= g.dx[k];
		int s = 1;
		for (int j = k + 1; j < d.dim; j++) s *= g.n[j];
		g.s[k] = s;
	}
	for (int k = d.dim; k < 3; k++) g.s[k] = 0;
	g.fields = d.vector ? d.fields : 1;
	return g;
}

int model_nparams(int model)
{
	switch (model) {
	case MMSPX_AC: return 5;
	case MMSPX_CH: return 5;
	case MMSPX_CH_DW: return 3;
	case MMSPX_GG: return 1;
	case MMSPX_AC_NOISE: return 7;
	case MMSPX_CH_NOISE: return 7;
	case MMSPX_ZENER: return 1;
	case MMSPX_OSTWALD: return 9;
	case MMSPX_GG_ANISO: return 4;
	}
	return -1;
}

int make_params(int model, const double* p, int np, ModelParams* out)
{
	const int need = model_nparams(model);
	if (need < 0) return fail(MMSPX_EINVAL, "unknown model");
	if (!p || np != need) return fail(MMSPX_EINVAL, "wrong number of model parameters");
	for (int k = 0; k < np; k++)
		if (!std::isfinite(p[k])) return fail(MMSPX_EINVAL, "model parameters must be finite");
	ModelParams m;
	std::memset(&m, 0, sizeof(m));
	m.model = model;
	if (noisy_model(model)) {
		// p = {r, u, K, M, dt, sigma, seed}
		if (!(p[5] >= 0.0)) return fail(MMSPX_EINVAL, "the noise amplitude sigma must be >= 0");
		if (!(p[6] >= 0.0 && p[6] < 9007199254740992.0 && p[6] == std::floor(p[6])))
			return fail(MMSPX_EINVAL, "the noise seed must be an integer in [0, 2^53)");
		m.sigma = p[5];
		m.sigmaf = (float)p[5];
		m.seed = (uint64_t)p[6];
	}
	if (model == MMSPX_AC || model == MMSPX_CH || noisy_model(model)) {
		// p = {r, u, K, M, dt}; the reference writes dt*M*(...) and -r*phi
		m.nr = -p[0];
		m.u = p[1];
		m.K = p[2];
		m.dtM = p[4] * p[3];
		m.nrf = -(float)p[0];
		m.uf = (float)p[1];
		m.Kf = (float)p[2];
		m.dtMf = (float)p[4] * (float)p[3];
	} else if (model == MMSPX_CH_DW) {
		// p = {kappa, M, dt}
		m.kappa = p[0];
		m.dtM = p[2] * p[1];
		m.kappaf = (float)p[0];
		m.dtMf = (float)p[2] * (float)p[1];
	} else if (model == MMSPX_OSTWALD) {
		// p = {dt, L, D, Calpha, Cbeta, A, delta, epsilon, kappa}; the derived
		// constants as ostwald.cpp:68-82 computes them (pow(x, 2) included)
		const double dt = p[0], L = p[1], D = p[2], Ca = p[3], Cb = p[4], A = p[5], delta = p[6];
		const double Cm = 0.5 * (Ca + Cb);
		const double gamma = 2.0 / std::pow(Cb - Ca, 2);
		m.ow[OW_A] = A;
		m.ow[OW_B] = A / std::pow(Cb - Cm, 2);
		m.ow[OW_G] = gamma;
		m.ow[OW_DA] = gamma / std::pow(delta, 2);
		m.ow[OW_DB] = gamma / std::pow(delta, 2);
		m.ow[OW_DELTA] = delta;
		m.ow[OW_EPS] = p[7];
		m.ow[OW_K] = p[8];
		m.ow[OW_CA] = Ca;
		m.ow[OW_CB] = Cb;
		m.ow[OW_CM] = Cm;
		m.ow[OW_DTD] = dt * D;
		m.ow[OW_DTL] = dt * L;
		for (int k = 0; k < 13; k++) m.owf[k] = (float)m.ow[k];
	} else if (model == MMSPX_GG_ANISO) {
		// p = {dt, width, gamma, mu}; energy() and mobility() return float
		// (anisotropy.hpp), then eps and w as graingrowth.cpp:146-148 form them
		const double dt = p[0], width = p[1];
		const double gamma = (double)(float)p[2], mu = (double)(float)p[3];
		if (!(width > 0.0) || !(gamma >= 0.0)) return fail(MMSPX_EINVAL, "width must be > 0 and gamma >= 0");
		const double eps = 4.0 / std::acos(-1.0) * std::sqrt(0.5 * gamma * width);
		m.an[AN_C2] = 0.5 * eps * eps;
		m.an[AN_W] = 4.0 * gamma / width;
		m.an[AN_MU] = mu;
		m.an[AN_DT] = dt;
		for (int k = 0; k < 4; k++) m.anf[k] = (float)m.an[k];
	} else {
		m.dt = p[0];
		m.dtf = (float)p[0];
	}
	*out = m;
	return 0;
}

}  // namespace mmspx

using namespace mmspx;

extern "C" {

const char* mmspx_strerror(int code)
{
	switch (code) {
	case MMSPX_OK: return "ok";
	case MMSPX_EINVAL: return "invalid argument";
	case MMSPX_ENOMEM: return "out of memory";
	case MMSPX_EDEVICE: return "HIP runtime error";
	case MMSPX_EIO: return "file I/O error";
	case MMSPX_EFORMAT: return "bad .dat content";
	case MMSPX_EBUSY: return "admission refused";
	case MMSPX_ENOTFOUND: return "unknown request";
	case MMSPX_EUNSUPPORTED: return "unsupported";
	}
	return "unknown error";
}

const char* mmspx_last_error(void) { return g_last_error.c_str(); }

int mmspx_abi_version(void) { return MMSPX_ABI_VERSION; }

size_t mmspx_footprint(const mmspx_desc* d, int model)
{
	if (validate_desc(d)) return 0;
	const size_t elem = d->precision == MMSPX_F64 ? 8 : 4;
	size_t nodes = 1;
	for (int k = 0; k < d->dim; k++) nodes *= (size_t)d->n[k];
	const size_t q = d->vector ? (size_t)d->fields : 1;
	size_t bufs = ch_family(model) ? 3 : 2;
	// Ostwald ripening: one scalar plane for w (the state-sized auxiliary buffer)
	if (model == MMSPX_OSTWALD) bufs = 3;
	return bufs * nodes * q * elem + q * 2 * kRedBlocks * sizeof(double);
}

void mmspx_destroy(mmspx_state* s)
{
	if (!s) return;
	// teardown: errors here cannot be reported to anyone, so they are dropped
	if (s->stream) (void)hipStreamSynchronize(s->stream);
	if (s->graph) (void)hipGraphExecDestroy(s->graph);
	if (s->cur) (void)hipFree(s->cur);
	if (s->nxt) (void)hipFree(s->nxt);
	if (s->aux) (void)hipFree(s->aux);
	if (s->part) (void)hipFree(s->part);
	if (s->part_host) (void)hipHostFree(s->part_host);
	if (s->ev0) (void)hipEventDestroy(s->ev0);
	if (s->ev1) (void)hipEventDestroy(s->ev1);
	if (s->stream) (void)hipStreamDestroy(s->stream);
	delete s;
}

int mmspx_create(const mmspx_desc* d, mmspx_state** out)
{
	if (!out) return fail(MMSPX_EINVAL, "null output pointer");
	*out = nullptr;
	int rc = validate_desc(d);
	if (rc) return rc;
	mmspx_state* s = new (std::nothrow) mmspx_state();
	if (!s) return fail(MMSPX_ENOMEM, "host allocation failed");
	s->desc = *d;
	s->geom = make_geom(*d);
	s->elem = d->precision == MMSPX_F64 ? 8 : 4;
	s->bytes = (size_t)s->geom.nodes * s->geom.fields * s->elem;
	hipError_t e = hipGetDevice(&s->device);
	if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
	if (e == hipSuccess) e = hipEventCreate(&s->ev0);
	if (e == hipSuccess) e = hipEventCreate(&s->ev1);
	if (e != hipSuccess) {
		mmspx_destroy(s);
		return fail(MMSPX_EDEVICE, std::string("HIP setup failed: ") + hipGetErrorString(e));
	}
	const size_t red = (size_t)s->geom.fields * 2 * kRedBlocks * sizeof(double);
	if (hipMalloc(&s->cur, s->bytes) != hipSuccess || hipMalloc(&s->nxt, s->bytes) != hipSuccess ||
	    hipMalloc((void**)&s->part, red) != hipSuccess || hipHostMalloc((void**)&s->part_host, red, 0) != hipSuccess) {
		mmspx_destroy(s);
		return fail(MMSPX_ENOMEM, "device allocation failed");
	}
	if (hipMemsetAsync(s->cur, 0, s->bytes, s->stream) != hipSuccess) {
		mmspx_destroy(s);
		return fail(MMSPX_EDEVICE, "hipMemsetAsync failed");
	}
	*out = s;
	return 0;
}

int mmspx_get_desc(const mmspx_state* s, mmspx_desc* out)
{
	if (!s || !out) return fail(MMSPX_EINVAL, "null argument");
	*out = s->desc;
	return 0;
}

int mmspx_upload(mmspx_state* s, const void* host, int layout)
{
	if (!s || !host) return fail(MMSPX_EINVAL, "null argument");
	if (layout != MMSPX_NODE_MAJOR && layout != MMSPX_FIELD_MAJOR) return fail(MMSPX_EINVAL, "unknown layout");
	if (s->geom.fields == 1 || layout == MMSPX_FIELD_MAJOR) {
		MMSPX_HIP(hipMemcpyAsync(s->cur, host, s->bytes, hipMemcpyHostToDevice, s->stream));
	} else {
		// stage node-major in the spare buffer, transpose into the state
		MMSPX_HIP(hipMemcpyAsync(s->nxt, host, s->bytes, hipMemcpyHostToDevice, s->stream));
		int rc = launch_transpose(s->nxt, s->cur, s->geom.nodes, s->geom.fields, s->elem, true, s->stream);
		if (rc) return rc;
	}
	MMSPX_HIP(hipStreamSynchronize(s->stream));
	return 0;
}

int mmspx_download(mmspx_state* s, void* host, int layout)
{
	if (!s || !host) return fail(MMSPX_EINVAL, "null argument");
	if (layout != MMSPX_NODE_MAJOR && layout != MMSPX_FIELD_MAJOR) return fail(MMSPX_EINVAL, "unknown layout");
	if (s->geom.fields == 1 || layout == MMSPX_FIELD_MAJOR) {
		MMSPX_HIP(hipMemcpyAsync(host, s->cur, s->bytes, hipMemcpyDeviceToHost, s->stream));
	} else {
		int rc = launch_transpose(s->cur, s->nxt, s->geom.nodes, s->geom.fields, s->elem, false, s->stream);
		if (rc) return rc;
		MMSPX_HIP(hipMemcpyAsync(host, s->nxt, s->bytes, hipMemcpyDeviceToHost, s->stream));
	}
	MMSPX_HIP(hipStreamSynchronize(s->stream));
	return 0;
}

int mmspx_set_variant(mmspx_state* s, int variant)
{
	if (!s || variant < 0 || variant > 2) return fail(MMSPX_EINVAL, "variant must be 0, 1 or 2");
	s->variant = variant;
	return 0;
}

// Splits n steps into passes of at most spl steps; with `even`, into an even
// number of passes (a captured graph must leave the buffers where it found
// them). Splitting the largest pass keeps every pass within spl.
static std::vector<int> plan_passes(int n, int spl, bool even)
{
	std::vector<int> v;
	for (int left = n; left > 0; left -= std::min(left, spl)) v.push_back(std::min(left, spl));
	if (even && (v.size() & 1)) {
		auto it = std::max_element(v.begin(), v.end());
		if (*it >= 2) {
			const int h = *it / 2;
			*it -= h;
			v.insert(it, h);
		}
	}
	return v;
}

// Every constant a launch bakes into its arguments (all but the per-launch
// step count): a captured graph is replayed only when all of them match.
static bool same_baked_params(const ModelParams& a, const ModelParams& b)
{
	if (a.model != b.model || a.nr != b.nr || a.u != b.u || a.K != b.K || a.dtM != b.dtM || a.kappa != b.kappa ||
	    a.dt != b.dt || a.nrf != b.nrf || a.uf != b.uf || a.Kf != b.Kf || a.dtMf != b.dtMf || a.kappaf != b.kappaf ||
	    a.dtf != b.dtf || a.sigma != b.sigma || a.sigmaf != b.sigmaf || a.seed != b.seed)
		return false;
	for (int k = 0; k < 13; k++)
		if (a.ow[k] != b.ow[k] || a.owf[k] != b.owf[k]) return false;
	for (int k = 0; k < 4; k++)
		if (a.an[k] != b.an[k] || a.anf[k] != b.anf[k]) return false;
	return true;
}

static int step_impl(mmspx_state* s, int model, const double* params, int np, int steps)
{
	if (!s) return fail(MMSPX_EINVAL, "null state");
	if (steps < 0) return fail(MMSPX_EINVAL, "steps must be >= 0");
	ModelParams p;
	int rc = make_params(model, params, np, &p);
	if (rc) return rc;
	rc = validate_model(s->desc, s->geom, model);
	if (rc) return rc;
	const Geom& g = s->geom;
	bool fused = false;
	TuneChoice tc{};
	if (s->variant != 1 && fused_supported(g, s->desc.precision, model)) {
		tc = s->tuned ? s->tile : default_choice(g, s->desc.precision, model);
		if (s->variant == 2 && tc.variant != 2) tc = default_choice(g, s->desc.precision, model);
		// where one thread per node is the default: the 3-D plane march, the 2-D strips
		if (s->variant == 2 && tc.variant != 2)
			tc = g.dim == 3 ? TuneChoice{2, 8, 64, 16, 8} : TuneChoice{2, 1, 128, 16, 0};
		fused = s->variant == 2 || tc.variant == 2;
	} else if (s->variant == 2) {
		return fail(MMSPX_EUNSUPPORTED, "no tiled/fused kernel for this model, shape and boundary condition");
	}
	s->last = fused ? tc : TuneChoice{1, 0, 0, 0, 0};
	s->has_last = true;
	const int spl = fused ? fused_steps_per_launch(model, tc) : 1;
	if ((((ch_family(model) || model == MMSPX_OSTWALD) && !fused) || (fused && fused_needs_plane(g, model, tc))) &&
	    !s->aux) {
		if (hipMalloc(&s->aux, s->bytes) != hipSuccess) return fail(MMSPX_ENOMEM, "device allocation failed");
	}
	// Small grids are launch-bound: replay an even number of launches from a
	// captured graph (the buffers are back in place after each replay). The
	// graph's step count is a multiple of two launches of spl steps (50 for
	// one step per launch, 48 for two or eight), so no launch in it is split.
	// Not the stochastic forms: each launch carries its step number.
	const size_t small_bytes = (size_t)64 << 20;
	const int graph_steps = std::max(2 * spl, kGraphSteps / (2 * spl) * (2 * spl));
	if (steps >= 2 * kGraphSteps && steps >= graph_steps && s->bytes <= small_bytes && !noisy_model(model)) {
		const bool same_params = same_baked_params(s->graph_params, p);
		const bool same_tile = s->graph_tile.variant == tc.variant && s->graph_tile.ty == tc.ty &&
		                       s->graph_tile.tz == tc.tz && s->graph_tile.chunk == tc.chunk &&
		                       s->graph_tile.opt == tc.opt;
		const bool hit = s->graph && s->graph_model == model && s->graph_fused == fused && s->graph_cur == s->cur &&
		                 s->graph_parity == (int)(s->steps_run & 1) && same_params && (!fused || same_tile);
		if (!hit) {
			if (s->graph) (void)hipGraphExecDestroy(s->graph);
			s->graph = nullptr;
			hipGraph_t graph = nullptr;
			MMSPX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
			int64_t nl = 0;
			void *a = s->cur, *b = s->nxt;
			for (int n : plan_passes(graph_steps, spl, true)) {
				rc = fused ? launch_fused_step(g, s->desc.precision, p, a, b, s->stream, tc, n, &nl, s->aux)
				           : launch_generic_step(g, s->desc.precision, p, a, b, s->aux, s->stream, &nl);
				if (rc) break;
				std::swap(a, b);
			}
			hipError_t e = hipStreamEndCapture(s->stream, &graph);
			if (rc) {
				if (graph) (void)hipGraphDestroy(graph);
				return rc;
			}
			if (e != hipSuccess) return fail(MMSPX_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
			e = hipGraphInstantiate(&s->graph, graph, nullptr, nullptr, 0);
			(void)hipGraphDestroy(graph);
			if (e != hipSuccess) return fail(MMSPX_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
			s->graph_model = model;
			s->graph_parity = (int)(s->steps_run & 1);
			s->graph_fused = fused;
			s->graph_cur = s->cur;
			s->graph_params = p;
			s->graph_tile = tc;
			s->graph_launches = (int)nl;
		}
		while (steps >= graph_steps) {
			MMSPX_HIP(hipGraphLaunch(s->graph, s->stream));
			s->launches += s->graph_launches;
			s->steps_run += graph_steps;
			steps -= graph_steps;
		}
	}
	while (steps > 0) {
		const int n = std::min(steps, spl);
		p.step = (uint64_t)s->steps_run;
		rc = fused ? launch_fused_step(g, s->desc.precision, p, s->cur, s->nxt, s->stream, tc, n, &s->launches, s->aux)
		           : launch_generic_step(g, s->desc.precision, p, s->cur, s->nxt, s->aux, s->stream, &s->launches);
		if (rc) return rc;
		std::swap(s->cur, s->nxt);
		s->steps_run += n;
		steps -= n;
	}
	return 0;
}

int mmspx_step(mmspx_state* s, int model, const double* params, int np, int steps)
{
	return step_impl(s, model, params, np, steps);
}

int mmspx_set_tile(mmspx_state* s, int ty, int tz, int chunk)
{
	if (!s) return fail(MMSPX_EINVAL, "null state");
	// on a 2-D grid ty > 1 names a grain-growth 2-D tile (body 1) or, with
	// tz == kTbCols, Allen-Cahn temporal blocks (body 3); otherwise the body
	// chosen by mmspx_set_tile_options (register queues, wave strips) is kept
	int opt = (s->tuned && (s->tile.opt == 1 || s->tile.opt == 2 || (s->tile.opt == 3 && s->geom.dim == 3) ||
	                        (s->tile.opt >= 4 && s->geom.dim == 2)))
	              ? s->tile.opt
	              : 0;
	if (s->geom.dim == 2 && ty > 1 && tz == kTbCols) opt = 3;
	else if (s->geom.dim == 2 && ty > 1) opt = 1;
	else if (s->geom.dim == 2 && s->tuned && s->tile.ty > 1) opt = 0;
	// 3-D two-step bodies: 32 x 32 tiles run body 7 only; body 6 runs 32 x 64
	if (s->geom.dim == 3 && ty == 32 && tz == 32) opt = 7;
	else if (s->geom.dim == 3 && s->tuned && s->tile.opt == 6 && ty == 32 && tz == 64) opt = 6;
	else if (s->geom.dim == 3 && s->tuned && (s->tile.opt == 8 || s->tile.opt == 12)) opt = s->tile.opt;
	TuneChoice t{2, ty, tz, chunk, opt};
	if (!tile_known(s->geom, t)) return fail(MMSPX_EINVAL, "no compiled tile of that shape for this grid");
	s->tile = t;
	s->tuned = true;
	return 0;
}

int mmspx_set_tile_options(mmspx_state* s, int body)
{
	if (!s) return fail(MMSPX_EINVAL, "null state");
	if (body < 0 || body > 12 || (body == 3 && s->geom.dim != 3) ||
	    ((body == 4 || body == 5 || (body >= 9 && body <= 11)) && s->geom.dim != 2) ||
	    (((body >= 6 && body <= 8) || body == 12) && s->geom.dim != 3))
		return fail(MMSPX_EINVAL, "unknown tile body");
	if (!s->tuned) {
		s->tile = default_choice(s->geom, s->desc.precision, MMSPX_CH_DW);
		s->tuned = true;
	}
	s->tile.opt = body;
	// 2-D: the split strips run 1 x 64 strips (bodies 4, 5), two steps per
	// pass 1 x kChWaveOut (bodies 9, 10)
	if (s->geom.dim == 2 && s->tile.variant == 2) {
		if (body >= 9 && body <= 11) {
			s->tile.ty = 1;
			s->tile.tz = kChWaveOut;
		} else if (body == 4 || body == 5) {
			s->tile.ty = 1;
			s->tile.tz = kWaveCols;
		}
	}
	// 3-D: keep a tile the body runs (32 x 32 tiles run body 7 only, body 6
	// runs 32 x 64 tiles, the plane march 8 x 64 for every model, the
	// two-step march 16 x 32 for every model)
	if (s->geom.dim == 3 && s->tile.variant == 2) {
		const int ty = s->tile.ty, tz = s->tile.tz;
		if (body == 12) {
			if (!((ty == 16 && tz == 32) || (ty == 16 && tz == 64) || (ty == 32 && tz == 64) || (ty == 8 && tz == 64) ||
			      (ty == 8 && tz == 32))) {
				s->tile.ty = 16;
				s->tile.tz = 32;
			}
		} else if (body == 8) {
			if (!((ty == 8 && tz == 64) || (ty == 16 && tz == 32) || (ty == 16 && tz == 64) || (ty == 32 && tz == 64))) {
				s->tile.ty = 8;
				s->tile.tz = 64;
			}
		} else if (body == 7) {
			s->tile.ty = 32;
			s->tile.tz = 32;
		} else if (body == 6 || (s->tile.ty == 32 && s->tile.tz == 32)) {
			s->tile.ty = 32;
			s->tile.tz = 64;
		}
	}
	return 0;
}

int mmspx_get_tile(const mmspx_state* s, int* ty, int* tz, int* chunk)
{
	if (!s) return fail(MMSPX_EINVAL, "null state");
	TuneChoice t = s->tuned ? s->tile : s->has_last ? s->last : default_choice(s->geom, s->desc.precision, MMSPX_CH_DW);
	if (ty) *ty = t.ty;
	if (tz) *tz = t.tz;
	if (chunk) *chunk = t.chunk;
	return t.variant + 16 * t.opt;
}

// Steps one launch of the state's current kernel choice advances (the
// temporal-block bodies take several steps per launch).
static int steps_per_launch_now(const mmspx_state* s, int model)
{
	if (s->variant == 1 || !fused_supported(s->geom, s->desc.precision, model)) return 1;
	TuneChoice tc = s->tuned ? s->tile : default_choice(s->geom, s->desc.precision, model);
	if (s->variant == 2 && tc.variant != 2) tc = default_choice(s->geom, s->desc.precision, model);
	return tc.variant == 2 ? fused_steps_per_launch(model, tc) : 1;
}

int mmspx_kernel_time(mmspx_state* s, int model, const double* params, int np, int reps, double* mean_s,
                      double* min_s)
{
	if (!s || reps < 1 || reps > 100000) return fail(MMSPX_EINVAL, "bad arguments");
	const int spl = steps_per_launch_now(s, model);
	std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
	int rc = 0;
	for (auto& e : ev) {
		if (hipEventCreate(&e) != hipSuccess) {
			rc = fail(MMSPX_EDEVICE, "hipEventCreate failed");
			break;
		}
	}
	for (int k = 0; k < reps && rc == 0; k++) {
		if (hipEventRecord(ev[2 * k], s->stream) != hipSuccess) rc = fail(MMSPX_EDEVICE, "hipEventRecord failed");
		if (rc == 0) rc = step_impl(s, model, params, np, spl);
		if (rc == 0 && hipEventRecord(ev[2 * k + 1], s->stream) != hipSuccess)
			rc = fail(MMSPX_EDEVICE, "hipEventRecord failed");
	}
	if (rc == 0 && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(MMSPX_EDEVICE, "synchronize failed");
	double sum = 0.0, mn = 1e30;
	for (int k = 0; k < reps && rc == 0; k++) {
		float ms = 0.0f;
		if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess) {
			rc = fail(MMSPX_EDEVICE, "hipEventElapsedTime failed");
			break;
		}
		sum += ms * 1e-3;
		mn = std::min(mn, ms * 1e-3);
	}
	for (auto& e : ev)
		if (e) (void)hipEventDestroy(e);
	if (rc) return rc;
	if (mean_s) *mean_s = sum / reps / spl;
	if (min_s) *min_s = mn / spl;
	return 0;
}

int mmspx_autotune(mmspx_state* s, int model, const double* params, int np, int trial_steps, char* report,
                   size_t report_len)
{
	if (!s || trial_steps < 1) return fail(MMSPX_EINVAL, "bad arguments");
	ModelParams p;
	int rc = make_params(model, params, np, &p);
	if (rc) return rc;
	std::string rep;
	TuneChoice cand[64];
	int nc = fused_supported(s->geom, s->desc.precision, model) ? tile_candidates(s->geom, s->desc.precision, model, cand, 64) : 0;
	// keep the state: every trial starts from it and it is restored at the end
	void* saved = nullptr;
	if (hipMalloc(&saved, s->bytes) != hipSuccess) return fail(MMSPX_ENOMEM, "device allocation failed");
	MMSPX_HIP(hipMemcpyAsync(saved, s->cur, s->bytes, hipMemcpyDeviceToDevice, s->stream));
	const int old_variant = s->variant;
	const bool old_tuned = s->tuned;
	const TuneChoice old_tile = s->tile;
	const int64_t old_steps = s->steps_run;  // the stochastic forms' noise continues from here
	double best = 1e30;
	TuneChoice best_t{1, 0, 0, 0, 0};
	for (int k = -1; k < nc && rc == 0; k++) {
		if (k < 0) {
			s->variant = 1;
		} else {
			s->variant = 2;
			s->tile = cand[k];
			s->tuned = true;
		}
		double mean = 0.0, mn = 0.0;
		rc = step_impl(s, model, params, np, 4);  // warm-up (and lazy buffers)
		if (rc == 0) rc = mmspx_kernel_time(s, model, params, np, trial_steps, &mean, &mn);
		if (rc == 0 && hipMemcpyAsync(s->cur, saved, s->bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
			rc = fail(MMSPX_EDEVICE, "restore copy failed");
		if (rc) break;
		char line[160];
		const TuneChoice t = k < 0 ? TuneChoice{1, 0, 0, 0, 0} : cand[k];
		std::snprintf(line, sizeof line, "variant=%d ty=%d tz=%d chunk=%d opt=%d mean_us=%.2f min_us=%.2f\n", t.variant,
		              t.ty, t.tz, t.chunk, t.opt, mean * 1e6, mn * 1e6);
		rep += line;
		if (mean < best) {
			best = mean;
			best_t = t;
		}
	}
	(void)hipStreamSynchronize(s->stream);
	(void)hipFree(saved);
	s->variant = old_variant;
	s->steps_run = old_steps;
	if (rc) {
		s->tuned = old_tuned;
		s->tile = old_tile;
		return rc;
	}
	s->tile = best_t;
	s->tuned = true;
	if (report && report_len) {
		std::snprintf(report, report_len, "%s", rep.c_str());
	}
	return 0;
}

int mmspx_copy_bandwidth(size_t bytes, int reps, double* bytes_per_s)
{
	if (bytes < 4096 || reps < 1) return fail(MMSPX_EINVAL, "bad arguments");
	bytes &= ~(size_t)255;
	void *a = nullptr, *b = nullptr;
	hipStream_t st = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	int rc = 0;
	if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = fail(MMSPX_ENOMEM, "alloc");
	if (rc == 0 && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
	                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
		rc = fail(MMSPX_EDEVICE, "stream/event setup failed");
	if (rc == 0 && hipMemsetAsync(a, 1, bytes, st) != hipSuccess) rc = fail(MMSPX_EDEVICE, "memset failed");
	double best = 1e30;
	for (int k = 0; k < reps + 1 && rc == 0; k++) {
		(void)hipEventRecord(e0, st);
		rc = launch_copy(a, b, bytes, st);
		(void)hipEventRecord(e1, st);
		if (rc == 0 && hipEventSynchronize(e1) != hipSuccess) rc = fail(MMSPX_EDEVICE, "sync failed");
		float ms = 0.0f;
		if (rc == 0) (void)hipEventElapsedTime(&ms, e0, e1);
		if (k > 0) best = std::min(best, (double)ms * 1e-3);
	}
	if (rc == 0 && bytes_per_s) *bytes_per_s = 2.0 * (double)bytes / best;
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	if (st) (void)hipStreamDestroy(st);
	if (a) (void)hipFree(a);
	if (b) (void)hipFree(b);
	return rc;
}

int mmspx_sync(mmspx_state* s)
{
	if (!s) return fail(MMSPX_EINVAL, "null state");
	MMSPX_HIP(hipStreamSynchronize(s->stream));
	return 0;
}

int mmspx_step_timed(mmspx_state* s, int model, const double* params, int np, int steps, double* seconds,
                     int64_t* launches)
{
	if (!s) return fail(MMSPX_EINVAL, "null state");
	const int64_t l0 = s->launches;
	MMSPX_HIP(hipEventRecord(s->ev0, s->stream));
	int rc = step_impl(s, model, params, np, steps);
	if (rc) return rc;
	MMSPX_HIP(hipEventRecord(s->ev1, s->stream));
	MMSPX_HIP(hipEventSynchronize(s->ev1));
	float ms = 0.0f;
	MMSPX_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
	if (seconds) *seconds = ms * 1e-3;
	if (launches) *launches = s->launches - l0;
	return 0;
}

int mmspx_reduce(mmspx_state* s, int what, int model, const double* params, int np, double* out, int nout)
{
	if (!s || !out) return fail(MMSPX_EINVAL, "null argument");
	ModelParams p;
	std::memset(&p, 0, sizeof(p));
	if (what == MMSPX_ENERGY) {
		int rc = make_params(model, params, np, &p);
		if (rc) return rc;
		if (nout < 1) return fail(MMSPX_EINVAL, "energy needs one output");
	} else if (what == MMSPX_MASS) {
		if (nout < s->geom.fields) return fail(MMSPX_EINVAL, "mass needs one output per field");
	} else {
		return fail(MMSPX_EINVAL, "unknown reduction");
	}
	int rc = launch_reduce(s->geom, s->desc.precision, what, p, s->cur, s->part, s->stream);
	if (rc) return rc;
	const int sets = what == MMSPX_MASS ? s->geom.fields : 1;
	const size_t bytes = (size_t)sets * 2 * kRedBlocks * sizeof(double);
	MMSPX_HIP(hipMemcpyAsync(s->part_host, s->part, bytes, hipMemcpyDeviceToHost, s->stream));
	MMSPX_HIP(hipStreamSynchronize(s->stream));
	for (int j = 0; j < sets; j++) out[j] = sum_partials(s->part_host + (size_t)j * 2 * kRedBlocks, kRedBlocks);
	if (what == MMSPX_ENERGY) out[0] *= s->geom.vol;
	return 0;
}

int mmspx_digest(mmspx_state* s, uint64_t* out)
{
	if (!s || !out) return fail(MMSPX_EINVAL, "null argument");
	uint64_t* part = (uint64_t*)s->part;  // kRedBlocks words: the partials' buffer holds 2 * kRedBlocks doubles
	int rc = launch_digest(s->geom, s->desc.precision, s->cur, part, s->stream);
	if (rc) return rc;
	MMSPX_HIP(hipMemcpyAsync(s->part_host, part, kRedBlocks * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
	MMSPX_HIP(hipStreamSynchronize(s->stream));
	const uint64_t* h = (const uint64_t*)s->part_host;
	uint64_t total = 0;
	for (int k = 0; k < kRedBlocks; k++) total += h[k];
	*out = total;
	return 0;
}

int mmspx_host_digest(const void* node_major, int64_t elements, int elem_bytes, uint64_t* out)
{
	if (!out || elements < 0 || (elements > 0 && !node_major) || (elem_bytes != 4 && elem_bytes != 8))
		return fail(MMSPX_EINVAL, "bad arguments");
	*out = host_digest(node_major, elements, elem_bytes);
	return 0;
}

int mmspx_write_dat(mmspx_state* s, const char* filename, int compress)
{
	if (!s || !filename) return fail(MMSPX_EINVAL, "null argument");
	std::vector<char> host(s->bytes);
	int rc = mmspx_download(s, host.data(), MMSPX_NODE_MAJOR);
	if (rc) return rc;
	return dat_write(filename, s->desc, host.data(), s->elem, compress);
}

int mmspx_peek_dat(const char* filename, mmspx_desc* d)
{
	if (!filename || !d) return fail(MMSPX_EINVAL, "null argument");
	int f64 = 0;
	return dat_read_header(filename, d, &f64);
}

int mmspx_read_dat(const char* filename, int precision, mmspx_state** out)
{
	if (!filename || !out) return fail(MMSPX_EINVAL, "null argument");
	mmspx_desc d;
	int f64 = 0;
	int rc = dat_read_header(filename, &d, &f64);
	if (rc) return rc;
	if (precision >= 0) d.precision = precision;
	rc = validate_desc(&d);
	if (rc) return rc;
	const Geom g = make_geom(d);
	const size_t count = (size_t)g.nodes * g.fields;
	std::vector<char> host(count * (f64 ? 8 : 4));
	rc = dat_read(filename, d, f64, host.data());
	if (rc) return rc;
	const bool want64 = d.precision == MMSPX_F64;
	std::vector<char> conv;
	const void* src = host.data();
	if (want64 != (bool)f64) {
		conv.resize(count * (want64 ? 8 : 4));
		for (size_t i = 0; i < count; i++) {
			if (want64)
				((double*)conv.data())[i] = ((const float*)host.data())[i];
			else
				((float*)conv.data())[i] = (float)((const double*)host.data())[i];
		}
		src = conv.data();
	}
	mmspx_state* s = nullptr;
	rc = mmspx_create(&d, &s);
	if (rc) return rc;
	rc = mmspx_upload(s, src, MMSPX_NODE_MAJOR);
	if (rc) {
		mmspx_destroy(s);
		return rc;
	}
	*out = s;
	return 0;
}

void mmspx_fill_uniform(double* out, int64_t n, uint64_t seed, double lo, double hi)
{
	for (int64_t i = 0; i < n; i++) {
		uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		z ^= z >> 31;
		out[i] = lo + (hi - lo) * ((double)(z >> 11) * 0x1.0p-53);
	}
}

}  // extern "C"
