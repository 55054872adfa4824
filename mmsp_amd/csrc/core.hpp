// core.hpp -- internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/mmsp_amd.h"
#include "dat.hpp"

namespace mmspx {

// ---- error plumbing ---------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define MMSPX_HIP(call)                                                                        \
	do {                                                                                       \
		hipError_t e_ = (call);                                                                \
		if (e_ != hipSuccess)                                                                  \
			return ::mmspx::fail(MMSPX_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
	} while (0)

// ---- geometry passed to kernels by value -----------------------------------
// Axis 0 is the slowest, axis dim-1 the contiguous one (MMSP.grid.cpp:306-313).
struct Geom {
	int dim;
	int n[3];      // extents (unused axes 1)
	int b0[3];
	int b1[3];
	int s[3];      // node strides per axis
	int nodes;     // product of extents
	int fields;
	double w[3];   // 1/(dx*dx), as MMSP.grid.cpp:2062
	float wf[3];   // the same weight rounded to float (FP32 form)
	double vol;    // product of dx (energy density -> energy)
	double dx[3];
};

Geom make_geom(const mmspx_desc& d);
int validate_desc(const mmspx_desc* d);

// ---- model parameters, pre-folded the way the reference's expressions fold
struct ModelParams {
	int model;
	double nr, u, K, dtM, kappa; // nr = -r; dtM = dt*M (left-to-right as in the reference)
	double dt;                   // grain growth
	float nrf, uf, Kf, dtMf, kappaf, dtf;
	// stochastic forms: noise amplitude, Philox key and the state's step
	// count at the launch (set per launch)
	double sigma;
	float sigmaf;
	uint64_t seed, step;
	// Ostwald ripening constants, derived as ostwald.cpp:68-82 derives them
	double ow[13];
	float owf[13];
	// anisotropic grain growth: 0.5*eps*eps, w = 4 gamma/width, mu, dt
	// (graingrowth.cpp:146-148, 165, 105)
	double an[4];
	float anf[4];
};
enum { AN_C2, AN_W, AN_MU, AN_DT };
constexpr int kAnisoMaxFields = 32;
enum { OW_A, OW_B, OW_G, OW_DA, OW_DB, OW_DELTA, OW_EPS, OW_K, OW_CA, OW_CB, OW_CM, OW_DTD, OW_DTL };

inline bool vector_model(int model)
{
	return model == MMSPX_GG || model == MMSPX_ZENER || model == MMSPX_OSTWALD || model == MMSPX_GG_ANISO;
}

inline bool noisy_model(int model) { return model == MMSPX_AC_NOISE || model == MMSPX_CH_NOISE; }
inline bool ch_family(int model) { return model == MMSPX_CH || model == MMSPX_CH_DW || model == MMSPX_CH_NOISE; }
inline bool ac_family(int model) { return model == MMSPX_AC || model == MMSPX_AC_NOISE; }
int make_params(int model, const double* p, int np, ModelParams* out);
int model_nparams(int model);
// the model runs on a grid of this kind (scalar/vector, fields, precision form)
int validate_model(const mmspx_desc& d, const Geom& g, int model);

// ---- the grid state ------------------------------------------------------------
struct TuneChoice {
	int variant;  // 1 generic, 2 tiled/fused
	int ty, tz, chunk;
	int opt;      // fused CH body: 0 LDS rings, 1 register queues
};

struct mmspx_state_impl;

}  // namespace mmspx

struct mmspx_state {
	mmspx_desc desc;
	mmspx::Geom geom;
	size_t elem;           // 4 or 8
	size_t bytes;          // one state buffer
	void* cur = nullptr;   // current state (field-major)
	void* nxt = nullptr;   // next state
	void* aux = nullptr;   // chemical potential / staging
	double* part = nullptr;  // reduction partials (device)
	double* part_host = nullptr;
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	int device = 0;
	int variant = 0;
	bool tuned = false;      // a tile chosen by mmspx_set_tile / mmspx_autotune
	mmspx::TuneChoice tile{};
	mmspx::TuneChoice last{};  // what the last step ran (reported by mmspx_get_tile when not tuned)
	bool has_last = false;
	int64_t launches = 0;
	int64_t steps_run = 0;
	// captured multi-step graph for small grids (launch-bound), keyed by what
	// its launches baked in
	hipGraphExec_t graph = nullptr;
	int graph_model = -1, graph_launches = 0, graph_parity = 0;
	bool graph_fused = false;
	mmspx::TuneChoice graph_tile{};
	mmspx::ModelParams graph_params{};
	void* graph_cur = nullptr;
};

namespace mmspx {

constexpr int kRedBlocks = 512;
constexpr int kGraphSteps = 50;  // steps per captured graph (even: buffers return to place)
constexpr int kRedThreads = 256;
// columns of one wave strip (2-D Allen-Cahn and grain growth): one per lane
constexpr int kWaveCols = 64;
// output columns of one Cahn-Hilliard wave strip (two halo lanes each side)
constexpr int kChWaveOut = 60;
// temporal blocks (2-D periodic Allen-Cahn): tile columns, and the most steps
// one pass takes (tile plus halo fits the 64 lanes of a wave)
constexpr int kTbCols = 48;
constexpr int kTbMaxSteps = 8;

// ---- kernel launchers (kernels_generic.hip, kernels_fused.hip, reduce.hip) ----
// Each returns 0 or a negative code; all enqueue on `st`.
int launch_generic_step(const Geom& g, int prec, const ModelParams& p, const void* a, void* b, void* aux,
                        hipStream_t st, int64_t* launches);
bool fused_supported(const Geom& g, int prec, int model);
// One launch of the tiled kernels; it advances `nsteps` steps, which must not
// exceed fused_steps_per_launch (1 except for temporal blocks).
int launch_fused_step(const Geom& g, int prec, const ModelParams& p, const void* a, void* b, hipStream_t st,
                      const TuneChoice& t, int nsteps, int64_t* launches, void* aux = nullptr);
// device scratch the fused kernels of `model` on tile choice t need beside
// the two state buffers: one scalar plane for 3-D Ostwald ripening on the two
// marches (w), none otherwise (the one-pass march, body 12, keeps w in LDS)
bool fused_needs_plane(const Geom& g, int model, const TuneChoice& t);
int fused_steps_per_launch(int model, const TuneChoice& t);
TuneChoice default_choice(const Geom& g, int prec, int model);
int launch_reduce(const Geom& g, int prec, int what, const ModelParams& p, const void* a, double* part,
                  hipStream_t st);
int launch_transpose(const void* src, void* dst, int nodes, int fields, size_t elem, bool to_field_major,
                     hipStream_t st);
int launch_convert(const void* src, int src_f64, void* dst, int dst_f64, int64_t n, hipStream_t st);
int launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st);
bool tile_known(const Geom& g, const TuneChoice& t);
int tile_candidates(const Geom& g, int prec, int model, TuneChoice* out, int max);

// host-side compensated sum of the (sum, comp) partial pairs
double sum_partials(const double* part, int count);
// state digests (reduce.hip): per-block partial sums of the device state, and
// the same digest of a node-major host array
int launch_digest(const Geom& g, int prec, const void* a, uint64_t* part, hipStream_t st);
uint64_t host_digest(const void* data, int64_t count, int elem_bytes);

}  // namespace mmspx
