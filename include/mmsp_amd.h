/* mmsp_amd.h -- C ABI of the MI355X (gfx950) explicit finite-difference
 * update library.
 *
 * The library replaces the per-node loops of the reference's example
 * update(grid&, int steps) plug-ins and the MMSP::laplacian they call:
 *
 *   MMSP::update(grid<dim,T>&, int)            allen-cahn.cpp:47-75
 *   MMSP::update(grid<dim,T>&, int)            cahn-hilliard/explicit/cahn-hilliard.cpp:47-81
 *   MMSP::update(grid<dim,vector<T> >&, int)   grain_growth/isotropic/phase_field/graingrowth.cpp:73-111
 *   MMSP::laplacian(grid, x)                   include/MMSP.grid.cpp:2049-2088
 *   check_boundary / ghostswap                 include/MMSP.utility.cpp:42-57, MMSP.grid.cpp:414-515
 *   grid::swap                                 include/MMSP.grid.cpp:1913-1991
 *   grid::output / grid::input (.dat layout)   include/MMSP.grid.cpp:1183-1402, 1818-1909
 *
 * (paths relative to the reference tree). The reference has no C ABI: its
 * plug-in boundary is the C++ template update(grid&, steps) called by
 * MMSP.main.hpp:279,305,331. include/mmsp_gpu.hpp maps that template onto
 * these calls.
 *
 * Conventions: every function returns 0 on success or a negative MMSPX_E*
 * code (never throws, never exits); mmspx_strerror() names the code and
 * mmspx_last_error() gives the message of the calling thread's last failure.
 * Host pointers are borrowed for the duration of the call. Device state is
 * owned by the library and persists across calls. Work is enqueued on the
 * state's own HIP stream; calls that return data to the host synchronise it.
 * One state must not be used from two host threads at once; distinct states
 * may be.
 */
#ifndef MMSP_AMD_H
#define MMSP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMSPX_ABI_VERSION 2

/* ---- error codes ---- */
enum {
	MMSPX_OK = 0,
	MMSPX_EINVAL = -1,    /* bad argument (shape, model, parameter count ...) */
	MMSPX_ENOMEM = -2,    /* device or host allocation failed */
	MMSPX_EDEVICE = -3,   /* a HIP runtime call failed */
	MMSPX_EIO = -4,       /* .dat read/write failure */
	MMSPX_EFORMAT = -5,   /* .dat content does not match the expected grid */
	MMSPX_EBUSY = -6,     /* admission refused (serving budget) */
	MMSPX_ENOTFOUND = -7, /* unknown request id */
	MMSPX_EUNSUPPORTED = -8
};

/* ---- boundary conditions: the reference's enum (MMSP.utility.h:22-29) ---- */
enum { MMSPX_MIRROR = 0, MMSPX_NEUMANN = 1, MMSPX_PERIODIC = 2, MMSPX_PARALLEL = 3, MMSPX_DIRICHLET = 4 };

/* ---- storage / arithmetic forms ---- */
enum {
	MMSPX_F64 = 0,    /* double storage and arithmetic, the reference's order (grid<dim,double>) */
	MMSPX_F32 = 1,    /* float storage and float arithmetic, same operation order */
	MMSPX_F32REF = 2  /* float storage, the reference's mixed float/double semantics (grid<dim,float>) */
};

/* ---- models and their parameter vectors ---- */
enum {
	MMSPX_AC = 0,     /* Allen-Cahn, allen-cahn.cpp:58-71.    p = {r, u, K, M, dt}
	                     new = phi - dt*M*(-r*phi + u*phi^3 - K*lap(phi)) */
	MMSPX_CH = 1,     /* Cahn-Hilliard, cahn-hilliard.cpp:59-79. p = {r, u, K, M, dt}
	                     mu = -r*c + u*c^3 - K*lap(c);  new = c + dt*M*lap(mu) */
	MMSPX_CH_DW = 2,  /* Cahn-Hilliard, double-well mu.      p = {kappa, M, dt}
	                     mu = 2c(1-c)(1-2c) - kappa*lap(c); new = c + dt*M*lap(mu) */
	MMSPX_GG = 3,     /* grain growth, graingrowth.cpp:84-105. p = {dt}
	                     new_j = phi_j - dt*(-phi_j - phi_j^3 + 2*(phi_j*sum_k phi_k^2 - lap phi_j)) */
	/* Stochastic forms: a Gaussian noise term xi = sigma*g(seed, step, node)
	 * added where the reference adds gaussian(0, std). p = {r, u, K, M, dt,
	 * sigma, seed}; seed an integer in [0, 2^53). */
	MMSPX_AC_NOISE = 4, /* model A, model_A.cpp:74-105 (sigma = sqrt(2kT/(dt dV)))
	                       new = phi - dt*M*(-r*phi + u*phi^3 - K*lap(phi) + xi) */
	MMSPX_CH_NOISE = 5, /* model B, model_B.cpp:74-111 (sigma = sqrt(2kT/(dt dV)));
	                       spinodal.cpp:74-107 is r = u = K = 1, M = 1/(2 dV), sigma = sqrt(eps dt/dV)
	                       mu = -r*c + u*c^3 - K*lap(c) + xi;  new = c + dt*M*lap(mu) */
	/* Vector-laplacian models (vector grids; forms MMSPX_F64 and MMSPX_F32): */
	MMSPX_ZENER = 6,    /* grain growth with Zener pinning, zener_pinning/isotropic/phase_field/zener.cpp:104-141.
	                       p = {dt}; field 0 (the particles) keeps its values, fields j >= 1 as MMSPX_GG */
	MMSPX_OSTWALD = 7,  /* Ostwald ripening, ostwald_ripening/isotropic/phase_field/ostwald.cpp:56-123.
	                       p = {dt, L, D, Calpha, Cbeta, A, delta, epsilon, kappa}; field 0 is the
	                       concentration C, fields j >= 1 the order parameters; Cmatrix = (Calpha+Cbeta)/2,
	                       B = A/(Cbeta-Cmatrix)^2, gamma = 2/(Cbeta-Calpha)^2, Dalpha = Dbeta = gamma/delta^2;
	                       w = -A(C-Cm) + B(C-Cm)^3 + Dalpha(C-Calpha)^3 + Dbeta(C-Cbeta)^3
	                           - gamma(C-Calpha) sum_j phi_j^2 - kappa lap C;  C' = C + dt*D*lap(w);
	                       phi_j' = phi_j - dt*L*(-gamma(C-Calpha)^2 + delta phi_j^3
	                                + epsilon phi_j (sum - phi_j^2) - kappa lap phi_j) */
	MMSPX_GG_ANISO = 8  /* multi-phase grain growth, grain_growth/anisotropic/phase_field/graingrowth.cpp:94-191.
	                       p = {dt, width, gamma, mu}: every pair of grains with boundary energy gamma and
	                       mobility mu (the example's anisotropy.hpp gives 1 and 1, as float). Per node: the
	                       fields present in the node's 2*dim+1 neighbourhood; if fewer than two, the node is
	                       copied; else pairwise variational derivatives and rates, phi' = phi + dt*(2/S)*dphi/dt
	                       clamped to [0, 1], then divided by the sum (Gibbs simplex). At most 32 fields. */
};
/* The noise: g(seed, step, node) is a standard normal deviate (float). For a
 * node at axis-0 index x and index r within its x-plane (the node's position
 * in the reference's order is x * plane + r), one Philox4x32-10 block
 * (Salmon et al., SC'11) with key {seed low, seed high} and counter {r, x/2,
 * step low, step high} gives, by the Box-Muller transform of its first two
 * words, a pair of independent deviates: the cosine one for even x, the sine
 * one for odd x. `step` counts the state's steps since its creation
 * (mmspx_step calls continue it; mmspx_autotune leaves it as it was). The
 * noise does not depend on the kernel, its tiles or the launch: every path
 * gives the same bits. The reference's gaussian() (model_A.cpp:13-39) draws
 * from rand() in node order by the polar method, so trajectories agree in
 * distribution only. Every precision form; single grids only (not batches
 * or servers). */

/* ---- host array layouts for upload/download ---- */
enum {
	MMSPX_NODE_MAJOR = 0,  /* node i's fields contiguous (the reference's node order; scalar grids: plain array) */
	MMSPX_FIELD_MAJOR = 1  /* one contiguous plane per field (the device layout) */
};

/* ---- reductions ---- */
enum {
	MMSPX_MASS = 0,    /* per field: sum over nodes (FP64, compensated) -> out[fields] */
	MMSPX_ENERGY = 1   /* total free energy of the model (FP64, compensated) -> out[0] */
};

typedef struct mmspx_desc {
	int dim;          /* 1, 2 or 3 */
	int g0[3];        /* lower global coordinate per axis (written to .dat) */
	int n[3];         /* extent per axis; axis 0 slowest, axis dim-1 contiguous (MMSP.grid.cpp:306-313) */
	int fields;       /* 1 for scalar grids, Q for grid<dim,vector<T> > */
	int vector;       /* 1 when the grid's node type is vector<T> (affects .dat type name and node layout) */
	int precision;    /* MMSPX_F64 / MMSPX_F32 / MMSPX_F32REF */
	int b0[3], b1[3]; /* boundary conditions per axis */
	double dx[3];     /* grid spacing per axis */
} mmspx_desc;

typedef struct mmspx_state mmspx_state;

const char* mmspx_strerror(int code);
const char* mmspx_last_error(void);
int mmspx_abi_version(void);

/* Grid state: creation allocates the device buffers (two state buffers; CH
 * adds one chemical-potential buffer on first use). Replaces grid::setup
 * (MMSP.grid.cpp:141-327) and the copy-constructed newGrid/temp of update(). */
int mmspx_create(const mmspx_desc* desc, mmspx_state** out);
void mmspx_destroy(mmspx_state* s);
int mmspx_get_desc(const mmspx_state* s, mmspx_desc* out);
/* Bytes of device memory the state holds for `model` (admission accounting). */
size_t mmspx_footprint(const mmspx_desc* desc, int model);

int mmspx_upload(mmspx_state* s, const void* host, int layout);
int mmspx_download(mmspx_state* s, void* host, int layout);

/* Advance `steps` explicit steps. Enqueued; returns without waiting. */
int mmspx_step(mmspx_state* s, int model, const double* params, int nparams, int steps);
int mmspx_sync(mmspx_state* s);

/* Explicit steps on caller-owned device buffers, allocating nothing.
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. The same kernels as mmspx_step with its default choice
 * (no tuned tile, no captured graph), enqueued on `stream` (a hipStream_t;
 * NULL: the null stream); the call returns without waiting.
 *   buf0 holds the state in the device layout (scalar grids: the nodes in the
 *   reference's order; vector grids: one contiguous plane per field), buf1 is
 *   a second buffer of the same size; both belong to the caller, on the
 *   current device. On return *result_buf (0 or 1) names the buffer that will
 *   hold the state after `steps` steps; the other holds intermediate data.
 *   first_step numbers the first step (the stochastic forms' noise counter).
 * The two-pass Cahn-Hilliard and Ostwald kernels (1-D Cahn-Hilliard, an axis
 * periodic on one side only, 1-D and small 3-D Ostwald ripening) and the 3-D
 * Ostwald marches (w between their two passes; not the one-pass march of
 * periodic grids with 2 fields) keep one scalar plane (nodes x
 * element size) in a fixed device array of MMSPX_NOALLOC_SCRATCH_BYTES that
 * each GPU holds until the process ends; larger planes are refused
 * (MMSPX_EINVAL).
 * Calls on one GPU share that array: calls whose work overlaps on one GPU
 * race. The first call on each device looks the array's address up (steps = 0
 * is enough); make that call before capturing later ones into a graph. */
#define MMSPX_NOALLOC_SCRATCH_BYTES (32u << 20)
int mmspx_step_noalloc(const mmspx_desc* desc, void* buf0, void* buf1, int model, const double* params,
                       int nparams, int steps, int64_t first_step, void* stream, int* result_buf);

/* Same as mmspx_step, bracketed by HIP events on the state's stream; waits and
 * returns the elapsed seconds and the kernel launches enqueued. */
int mmspx_step_timed(mmspx_state* s, int model, const double* params, int nparams, int steps,
                     double* seconds, int64_t* launches);

/* On-device reductions (FP64 with Neumaier compensation). `params` as for
 * mmspx_step (ENERGY uses the model's gradient coefficient and bulk terms). */
int mmspx_reduce(mmspx_state* s, int what, int model, const double* params, int nparams, double* out, int nout);

/* A 64-bit digest of the state, in the order of the node-major host array
 * that mmspx_download(s, host, MMSPX_NODE_MAJOR) writes: the sum, modulo 2^64,
 * over every element of a mix of its bits and its index (node * fields +
 * field). mmspx_host_digest gives the same value for a host array (elem_bytes
 * 8 for double, 4 for float), so a caller can tell whether a host copy still
 * holds what the device holds without uploading it. No counterpart in the
 * reference: include/mmsp_gpu.hpp uses it to keep a grid's device state
 * across update() calls (MMSP.main.hpp:304-325). */
int mmspx_digest(mmspx_state* s, uint64_t* out);
int mmspx_host_digest(const void* node_major, int64_t elements, int elem_bytes, uint64_t* out);

/* .dat output/input in the reference's serial layout. compress=1 writes a
 * zlib level-9 block (the reference's default), 0 a raw block (readable by the
 * reference: MMSP.grid.cpp:1306-1330). */
int mmspx_write_dat(mmspx_state* s, const char* filename, int compress);
/* Reads a .dat file into a new state with the given precision (-1: from the
 * file's type). Multi-block files are assembled by their block limits. */
int mmspx_read_dat(const char* filename, int precision, mmspx_state** out);
/* Reads just the header of a .dat file into *desc (no device work). */
int mmspx_peek_dat(const char* filename, mmspx_desc* desc);

/* ---- checkpoints that overlap stepping ----
 * Replaces grid::output (MMSP.grid.cpp:1358-1402, write_buffer :1818-1909) in
 * the driver loop of MMSP.main.hpp:278-350, which deflates the whole grid on
 * one core between update() calls. The file has the serial layout (one block)
 * and reads back through grid::input and the reference's converters.
 *
 * mmspx_checkpoint_begin snapshots the state on the state's stream (one extra
 * state-sized device allocation, stream-ordered) and returns; later
 * mmspx_step calls on the state run while host threads download and write the
 * snapshot. `threads` host threads deflate it (0: one per host core): with
 * one, the zlib stream is compress2 level 9's, byte for byte the reference's
 * file; with more, the payload is deflated in chunks that form one zlib
 * stream (identical payload, different compressed bytes). compress=0 writes
 * a raw block. mmspx_checkpoint_wait joins the writer, returns its status and
 * the seconds from begin to the file's close, and frees the handle; every
 * begun checkpoint must be waited for. The state may be stepped, read or
 * destroyed while its checkpoints are in flight. */
typedef struct mmspx_checkpoint mmspx_checkpoint;
int mmspx_checkpoint_begin(mmspx_state* s, const char* filename, int compress, int threads, mmspx_checkpoint** out);
int mmspx_checkpoint_wait(mmspx_checkpoint* c, double* seconds);
/* The same writer over a host array of nodes in row-major order, each node's
 * fields contiguous (no device work). */
int mmspx_write_dat_host(const mmspx_desc* desc, const void* node_major, const char* filename, int compress,
                         int threads);

/* Seeded inputs: element i = lo + (hi-lo) * u_i, u_i = top 53 bits of the
 * (i+1)-th SplitMix64 output from state `seed`, over [0,1). */
void mmspx_fill_uniform(double* out, int64_t n, uint64_t seed, double lo, double hi);

/* Kernel variant selection: 0 auto (tuned tile if any, else the default),
 * 1 the generic one-thread-per-node kernels, 2 the tiled/fused kernels
 * (MMSPX_EUNSUPPORTED where none exists for the model/shape/boundaries). */
int mmspx_set_variant(mmspx_state* s, int variant);
/* Tile of the fused kernels: ty x tz cells of the (axis 1, last axis) plane,
 * marching `chunk` planes along axis 0. 2-D grids: ty = 1 selects strips of
 * tz columns (tz = 60: the wave strips of Cahn-Hilliard); ty = 8 or 16 with
 * tz = 64 the 2-D grain-growth tiles; ty = 16 or 32 with tz = 48 the
 * temporal blocks of periodic Allen-Cahn, which take `chunk` (at most 8)
 * steps per pass over HBM. */
int mmspx_set_tile(mmspx_state* s, int ty, int tz, int chunk);
/* Tile body, all with the same bits. Cahn-Hilliard: 0 keeps every neighbour
 * in LDS rings; 1 keeps the axis-0 neighbours in per-thread register queues;
 * 3-D 32 x 64 tiles only: 2 and 3 are the register queues with 768 and 1024
 * threads per tile (fewer points, fewer registers each). 2-D Allen-Cahn and grain growth with
 * 1 x 64 strips: 2 selects the wave strips (one 64-column strip per wave,
 * neighbours across lanes by DPP, no LDS); 2-D grain growth and Zener pinning
 * with 8, 16 or 32 fields: 4 and 5 select the split strips (two or four
 * waves per strip, each holding its share of the fields), 9 and 10 the same
 * two steps per pass on periodic grids (1 x 60 strips); 11 two steps per pass
 * on one wave per strip for 2, 4 or 8 fields (periodic). 3-D periodic
 * Cahn-Hilliard (deterministic chemical potentials): 6 and 7 take two steps
 * per pass over HBM (6 on 32 x 64 tiles, 7 on 32 x 32 tiles; an odd last
 * step runs body 2). On a 3-D grid the tile follows the body: 7 takes 32 x
 * 32 tiles, 6 takes 32 x 64, and the other bodies leave 32 x 32 for 32 x 64.
 * 3-D Allen-Cahn, model A, grain growth, Zener pinning and multi-phase grain growth
 * (2, 4 or 8 fields), any boundary: 8 marches the tile along axis 0 (8, 16,
 * 32 x 64 and 16 x 32 tiles for Allen-Cahn; 8 x 64 and 16 x 32 for the
 * vector models). 12 is that march two steps per pass over HBM on periodic
 * grids (Allen-Cahn on 16 x 32, 16 x 64 and 32 x 64 tiles; grain growth and
 * Zener pinning with 2, 4 or 8 fields on 8 x 64, 16 x 32 and 8 x 32 tiles,
 * the FP64 form with 8 fields on 8 x 32 only; an odd last step runs body 8)
 * and, for Ostwald ripening (2, 3 or 4 fields, 8 x 64 and 16 x 32 tiles),
 * its two passes in one march with w kept in LDS; the tile follows the body
 * (16 x 32 where the current tile is not one of those). */
int mmspx_set_tile_options(mmspx_state* s, int body);
/* Returns variant + 16 * tile body (1: register queues, or the 2-D
 * grain-growth tile; 2: wave strips, or 3-D register queues on 768 threads;
 * 3: temporal blocks, or 3-D register queues on 1024 threads; 4 and 5: the
 * split strips on two and four waves). Before any tile is set, reports what the last step ran
 * (the Cahn-Hilliard default before any step). */
int mmspx_get_tile(const mmspx_state* s, int* ty, int* tz, int* chunk);
/* Project autotuner: times every candidate tile for `model` on this state's
 * shape (trial_steps each, state restored afterwards), keeps the fastest and
 * writes a one-line-per-candidate report. */
int mmspx_autotune(mmspx_state* s, int model, const double* params, int nparams, int trial_steps, char* report,
                   size_t report_len);
/* Per-step device time: `reps` launches of the current kernel choice (one
 * step, or the steps one launch of a multi-step body advances), each bracketed
 * by its own HIP event pair on the state's stream; mean and minimum seconds
 * per step. */
int mmspx_kernel_time(mmspx_state* s, int model, const double* params, int nparams, int reps, double* mean_s,
                      double* min_s);
/* Device-to-device copy kernel over `bytes` (read + write), best of `reps`;
 * gives the attainable HBM rate and the counter calibration kernel. */
int mmspx_copy_bandwidth(size_t bytes, int reps, double* bytes_per_s);

/* ---- batches: many independent grids advanced together ---------------------
 * All members share model, precision and field count and advance by the same
 * number of steps; extents and parameters may differ. One persistent kernel
 * launch per step walks a device work list holding every member's tiles
 * (2-D Allen-Cahn / grain growth, periodic 2-D and 3-D Cahn-Hilliard; 3-D
 * members take a second launch); multi-step runs replay a captured hipGraph.
 * No counterpart in the reference, which advances one grid per process. */
typedef struct mmspx_batch mmspx_batch;

int mmspx_batch_create(const mmspx_desc* descs, int count, int model, mmspx_batch** out);
void mmspx_batch_destroy(mmspx_batch* b);
int mmspx_batch_upload(mmspx_batch* b, int member, const void* host, int layout);
int mmspx_batch_download(mmspx_batch* b, int member, void* host, int layout);
/* params: count * nparams doubles, member-major. */
int mmspx_batch_set_params(mmspx_batch* b, const double* params, int nparams);
/* Advance every member by `steps`; graph_steps > 0 captures that many steps
 * per hipGraph and replays it. */
int mmspx_batch_step(mmspx_batch* b, int steps, int graph_steps);
int mmspx_batch_step_timed(mmspx_batch* b, int steps, int graph_steps, double* seconds);
int mmspx_batch_reduce(mmspx_batch* b, int member, int what, double* out, int nout);
/* Project autotuner: times every candidate (strip width, rows per work item,
 * persistent workgroups per CU) on the batch itself, keeps the fastest and
 * restores the members' state. */
int mmspx_batch_autotune(mmspx_batch* b, int trial_steps, char* report, size_t report_len);
/* Sets the work decomposition directly: strips of strip_width (64, 128, 256)
 * columns by rows_per_item rows with blocks_per_cu persistent workgroups per
 * CU, or, for Allen-Cahn with cells > 0, runs of `cells` consecutive nodes.
 * strip_width 48 selects temporal blocks (periodic Allen-Cahn members only):
 * 16 x 48-node blocks taking rows_per_item (1..8) steps per pass over HBM. */
int mmspx_batch_set_tile(mmspx_batch* b, int strip_width, int rows_per_item, int blocks_per_cu, int cells);
/* Returns 1 when the tile came from the autotuner, 0 for the default. For
 * Allen-Cahn batches run by runs of consecutive nodes, strip_width and
 * blocks_per_cu are 0 and rows_per_item holds the nodes per workgroup; for
 * temporal blocks strip_width is 48 and rows_per_item the steps per pass. */
int mmspx_batch_get_tile(const mmspx_batch* b, int* strip_width, int* rows_per_item, int* blocks_per_cu);

/* ---- serving: continuous batching of solves ------------------------------ */
typedef struct mmspx_server mmspx_server;

typedef struct mmspx_server_config {
	size_t hbm_budget;   /* bytes of device memory the server may hold for request state */
	int max_slots;       /* most requests resident at once */
	int chunk_steps;     /* most steps in one round; a round advances the resident requests with the
	                        least remaining work (nodes x steps left), as many as fill the GPU */
	int model;           /* MMSPX_CH_DW, MMSPX_CH or MMSPX_AC */
	int precision;       /* MMSPX_F32 or MMSPX_F64 */
	int fp64_above_steps; /* MMSPX_F32 servers: requests of more steps run in MMSPX_F64 (0: every request in
	                         `precision`). The FP32 error of a Cahn-Hilliard solve against the FP64 form grows
	                         with its length (the spinodal instability amplifies the rounding of the stored
	                         state); 1000 keeps the served results of the benchmark stream within 1e-5 */
} mmspx_server_config;

typedef struct mmspx_request_info {
	int64_t id;
	int status;           /* 0 queued, 1 running, 2 done, 3 rejected, 4 cancelled */
	int steps_done;
	double t_submit;      /* seconds on the server clock */
	double t_admit;
	double t_done;
	double mass0, mass1;  /* FP64 compensated mass at admission and at completion */
	int precision;        /* the form the request runs in (MMSPX_F32 or MMSPX_F64) */
} mmspx_request_info;

int mmspx_server_create(const mmspx_server_config* cfg, mmspx_server** out);
/* Stops the server's threads (the rounds already queued finish on the device)
 * and frees every request, finished or not. */
void mmspx_server_destroy(mmspx_server* sv);
/* Queue a solve. `init` (node-major doubles, converted to the server's
 * precision) is copied before return. params as for mmspx_step. */
int mmspx_server_submit(mmspx_server* sv, const mmspx_desc* desc, const double* init, const double* params,
                        int nparams, int steps, int64_t* id);
int mmspx_server_cancel(mmspx_server* sv, int64_t id);
/* Wait up to timeout_s (negative: forever) for a request to finish. */
int mmspx_server_wait(mmspx_server* sv, int64_t id, double timeout_s, mmspx_request_info* info);
/* Copy a finished request's final state (node-major doubles) and release it. */
int mmspx_server_fetch(mmspx_server* sv, int64_t id, double* out, size_t count);
int mmspx_server_info(mmspx_server* sv, int64_t id, mmspx_request_info* info);
double mmspx_server_clock(mmspx_server* sv);
/* Counters: [0] admitted, [1] completed, [2] rejected, [3] rounds, [4] peak bytes,
 * [5] peak requests resident at once, [6] kernel launches, [7] requests run in
 * MMSPX_F64 by an MMSPX_F32 server (fp64_above_steps). */
int mmspx_server_stats(mmspx_server* sv, int64_t* out, int nout);

#ifdef __cplusplus
}
#endif

#endif /* MMSP_AMD_H */
