// mmsp_gpu.hpp -- drop-in GPU bodies for MMSP's update(grid&, steps) plug-ins.
//
// The reference's plug-in boundary is the C++ template each example defines in
// namespace MMSP and MMSP.main.hpp calls (MMSP.main.hpp:279, 305, 331):
//
//   template <int dim, typename T> void update(grid<dim,T>& oldGrid, int steps);   // allen-cahn.cpp:47,
//                                                                                  // cahn-hilliard.cpp:47
//   template <int dim, typename T> void update(grid<dim,vector<T> >&, int steps);  // graingrowth.cpp:73
//
// This header keeps that boundary. An example replaces its per-node loop with
// one call, and main(), output() and the .dat layout stay the reference's:
//
//   #include "MMSP.hpp"
//   #include "mmsp_gpu.hpp"
//   namespace MMSP {
//   template <int dim, typename T> void update(grid<dim,T>& g, int steps)
//   { mmsp_gpu::allen_cahn(g, steps, 1.0, 1.0, 1.0, 1.0, 0.01); }   // r, u, K, M, dt of allen-cahn.cpp:58-62
//   }
//   #include "MMSP.main.hpp"
//
// Link with -lmmsp_amd (include/mmsp_amd.h is the C ABI underneath).
//
// What a call does: all `steps` explicit steps run on the device (the
// MMSP::laplacian, check_boundary, swap and ghostswap of the loop become
// device work), and the result comes back into the same grid, so the caller
// sees the same grid the reference's loop leaves (the reference swaps a
// locally constructed grid into the caller's, MMSP.grid.cpp:1913-1991; values
// are what matter to callers). The device state stays resident between calls:
// MMSP.main.hpp:304-325 calls update(grid, increment) in a loop with only
// output() in between, so each grid (keyed by the grid object, its node data
// and its shape) keeps one library state, and a call uploads the host data
// only when it no longer matches what the device holds: a 64-bit digest of the
// device state is taken before each download (mmspx_digest) and compared with
// the host data's (mmspx_host_digest) at the next call. A grid the caller
// changed in between is therefore uploaded again; an unchanged one is not.
// mmsp_gpu::release() frees the resident states (otherwise they live until
// the process ends). The progress bar of MMSP::print_progress
// (MMSP.utility.cpp:138-178) is printed with the reference's marks; the steps
// between marks are batched into runs of at least 100 steps (a mark is
// printed once the device has passed it), so small grids still replay the
// library's captured graphs. MMSP_GPU_STATS=1 in the environment prints the
// calls, uploads, downloads and their seconds to stderr at exit.
// grid<dim,double> runs the FP64 form
// (bitwise the reference's arithmetic except the cube, see the library);
// grid<dim,float> and vector<float> run the reference's mixed float/double
// semantics unless set_native_float(true) selects pure float arithmetic.
//
// Errors follow the reference's convention: a message on std::cerr, then
// MMSP::Abort(-1) (MMSP.utility.cpp:30-36).
//
// Serial builds only: under MPI the grid holds ghost cells and one block per
// rank (MMSP.grid.cpp:141-327), which this header does not map.
#ifndef MMSP_GPU_HPP
#define MMSP_GPU_HPP

#ifdef MPI_VERSION
#error "mmsp_gpu.hpp drives one serial grid per process; MPI builds of MMSP are not supported"
#endif

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "mmsp_amd.h"

namespace mmsp_gpu {

inline bool& native_float_flag()
{
	static bool v = false;
	return v;
}
// Pure float arithmetic for float grids (MMSPX_F32) instead of the reference's
// float-storage/double-expression semantics (MMSPX_F32REF).
inline void set_native_float(bool on) { native_float_flag() = on; }

inline void check(int rc, const char* what)
{
	if (rc != 0) {
		std::cerr << "mmsp_gpu: " << what << " failed: " << mmspx_strerror(rc) << ": " << mmspx_last_error()
		          << std::endl;
		MMSP::Abort(-1);
	}
}

template <typename T> struct precision_of;
template <> struct precision_of<double> {
	static int get() { return MMSPX_F64; }
};
template <> struct precision_of<float> {
	static int get() { return native_float_flag() ? MMSPX_F32 : MMSPX_F32REF; }
};

// Descriptor of a serial MMSP grid through its public friends
// (MMSP.grid.h:268-337): extents x0/x1, spacing dx, boundary conditions b0/b1.
template <int dim, typename G> mmspx_desc describe(const G& grid, int fields, int is_vector, int precision)
{
	mmspx_desc d;
	std::memset(&d, 0, sizeof d);
	d.dim = dim;
	for (int k = 0; k < 3; k++) {
		d.n[k] = 1;
		d.dx[k] = 1.0;
		d.b0[k] = d.b1[k] = MMSPX_PERIODIC;
	}
	// the accessors are friends defined inside grid: found by argument-dependent lookup
	for (int k = 0; k < dim; k++) {
		d.g0[k] = x0(grid, k);
		d.n[k] = x1(grid, k) - x0(grid, k);
		d.dx[k] = dx(grid, k);
		d.b0[k] = b0(grid, k);
		d.b1[k] = b1(grid, k);
	}
	d.fields = fields;
	d.vector = is_vector;
	d.precision = precision;
	return d;
}

// ---- resident device states -------------------------------------------------
struct resident {
	const void* grid = nullptr;  // the grid object
	const void* data = nullptr;  // its node data (scalar grids: &grid(0))
	mmspx_desc desc{};
	mmspx_state* s = nullptr;
	uint64_t digest = 0;  // of the device state when it was last downloaded
	bool has_digest = false;
};

struct counters {
	long calls = 0, uploads = 0, downloads = 0, steps = 0;
	double upload_s = 0, download_s = 0, step_s = 0, digest_s = 0;
};

inline counters& stats()
{
	static counters c;
	return c;
}

inline void print_stats()
{
	const counters& c = stats();
	std::fprintf(stderr,
	             "mmsp_gpu: update() calls %ld, steps %ld, uploads %ld, downloads %ld; seconds: steps %.6f, "
	             "uploads %.6f, downloads %.6f, digests %.6f\n",
	             c.calls, c.steps, c.uploads, c.downloads, c.step_s, c.upload_s, c.download_s, c.digest_s);
}

inline double now_s()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline std::vector<resident>& residents()
{
	static std::vector<resident> v;
	static bool hooked = false;
	if (!hooked) {
		hooked = true;
		const char* e = std::getenv("MMSP_GPU_STATS");
		if (e && *e && std::strcmp(e, "0") != 0) std::atexit(print_stats);
	}
	return v;
}

// Frees every resident device state (the next update() uploads again).
inline void release()
{
	for (resident& r : residents())
		if (r.s) mmspx_destroy(r.s);
	residents().clear();
}

inline bool same_desc(const mmspx_desc& a, const mmspx_desc& b) { return std::memcmp(&a, &b, sizeof a) == 0; }

// The resident state of this grid, created (or created again when the grid's
// data or shape changed); at most eight grids stay resident, the oldest leaves.
inline resident& resident_for(const void* grid, const void* data, const mmspx_desc& d)
{
	const size_t kMaxResident = 8;
	std::vector<resident>& v = residents();
	for (size_t k = 0; k < v.size(); k++) {
		if (v[k].grid != grid) continue;
		if (v[k].data == data && same_desc(v[k].desc, d)) return v[k];
		mmspx_destroy(v[k].s);
		v.erase(v.begin() + (long)k);
		break;
	}
	if (v.size() >= kMaxResident) {
		mmspx_destroy(v.front().s);
		v.erase(v.begin());
	}
	resident r;
	r.grid = grid;
	r.data = data;
	r.desc = d;
	check(mmspx_create(&d, &r.s), "mmspx_create");
	v.push_back(r);
	return v.back();
}

// Uploads `host` (node-major) unless the device already holds it.
inline void upload_if_changed(resident& r, const void* host, int64_t elements, int elem_bytes)
{
	if (r.has_digest) {
		const double t0 = now_s();
		uint64_t h = 0;
		check(mmspx_host_digest(host, elements, elem_bytes, &h), "mmspx_host_digest");
		stats().digest_s += now_s() - t0;
		if (h == r.digest) return;
	}
	const double t0 = now_s();
	check(mmspx_upload(r.s, host, MMSPX_NODE_MAJOR), "mmspx_upload");
	stats().upload_s += now_s() - t0;
	stats().uploads += 1;
	r.has_digest = false;
}

// Downloads the state into `host` (node-major), keeping its digest.
inline void download_keep_digest(resident& r, void* host)
{
	double t0 = now_s();
	check(mmspx_digest(r.s, &r.digest), "mmspx_digest");
	stats().digest_s += now_s() - t0;
	t0 = now_s();
	check(mmspx_download(r.s, host, MMSPX_NODE_MAJOR), "mmspx_download");
	stats().download_s += now_s() - t0;
	stats().downloads += 1;
	r.has_digest = true;
}

// Runs `steps` on the state with the reference's progress bar: print_progress
// marks step 0, every step with (20 * step) % steps == 0 and the last step,
// each printed once that many steps are done. The device advances in runs of
// at least 100 steps between host synchronisations.
inline void run_with_progress(mmspx_state* s, int model, const double* p, int np, int steps)
{
	const int kMinRun = 100;  // at least two of the library's captured graphs (48 or 50 steps)
	const double t0 = now_s();
	int done = 0;
	std::vector<int> pending;
	auto advance = [&](int to) {
		if (to > done) {
			check(mmspx_step(s, model, p, np, to - done), "mmspx_step");
			check(mmspx_sync(s), "mmspx_sync");
			done = to;
		}
		size_t k = 0;
		for (; k < pending.size() && pending[k] <= done; k++) MMSP::print_progress(pending[k], steps);
		pending.erase(pending.begin(), pending.begin() + (long)k);
	};
	for (int step = 0; step < steps; step++) {
		const bool prints = step == 0 || step == steps - 1 || (20 * step) % steps == 0;
		if (!prints) continue;
		pending.push_back(step);
		if (step == 0 || step - done >= kMinRun) advance(step);
	}
	advance(steps);
	stats().step_s += now_s() - t0;
	stats().steps += steps;
}

// Scalar grids: grid<dim,double> / grid<dim,float>; in a serial build the
// nodes are one contiguous array starting at &grid(0) (MMSP.grid.cpp:320, 385-400).
template <int dim, typename T> void run_scalar(MMSP::grid<dim, T>& grid, int model, const double* p, int np, int steps)
{
	if (nodes(grid) == 0 || steps <= 0) return;
	stats().calls += 1;
	mmspx_desc d = describe<dim>(grid, fields(grid), 0, precision_of<T>::get());
	T* data = &grid(0);
	resident& r = resident_for(&grid, data, d);
	upload_if_changed(r, data, nodes(grid), (int)sizeof(T));
	run_with_progress(r.s, model, p, np, steps);
	download_keep_digest(r, data);
}

// Allen-Cahn, allen-cahn.cpp:47-75: new = phi - dt*M*(-r*phi + u*phi^3 - K*lap(phi)).
template <int dim, typename T>
void allen_cahn(MMSP::grid<dim, T>& grid, int steps, double r = 1.0, double u = 1.0, double K = 1.0, double M = 1.0,
                double dt = 0.01)
{
	const double p[5] = {r, u, K, M, dt};
	run_scalar(grid, MMSPX_AC, p, 5, steps);
}

// Cahn-Hilliard with the reference's chemical potential, cahn-hilliard.cpp:47-81:
// mu = -r*phi + u*phi^3 - K*lap(phi); new = phi + dt*M*lap(mu).
template <int dim, typename T>
void cahn_hilliard(MMSP::grid<dim, T>& grid, int steps, double r = 1.0, double u = 1.0, double K = 1.0, double M = 1.0,
                   double dt = 0.01)
{
	const double p[5] = {r, u, K, M, dt};
	run_scalar(grid, MMSPX_CH, p, 5, steps);
}

// Cahn-Hilliard with the double-well chemical potential mu = 2c(1-c)(1-2c) - kappa*lap(c).
template <int dim, typename T>
void cahn_hilliard_double_well(MMSP::grid<dim, T>& grid, int steps, double kappa = 1.0, double M = 1.0,
                               double dt = 0.005)
{
	const double p[3] = {kappa, M, dt};
	run_scalar(grid, MMSPX_CH_DW, p, 3, steps);
}

// Noise seed of one update() call, drawn from rand() -- the generator the
// reference's gaussian() uses, seeded once by MMSP.main.hpp -- so successive
// update() calls get fresh noise and srand() fixes a run. The state made for
// the call counts its steps from 0; the seed keeps calls apart.
inline double next_noise_seed()
{
	const uint64_t a = (uint64_t)std::rand(), b = (uint64_t)std::rand();
	return (double)(((a << 31) ^ b) & ((uint64_t(1) << 53) - 1));
}

// Model A, model_A.cpp:74-105: new = phi - dt*M*(-r*phi + u*phi^3 - K*lap(phi) + noise),
// noise = gaussian(0, sqrt(2 kT/(dt dV))).
template <int dim, typename T>
void model_a(MMSP::grid<dim, T>& grid, int steps, double r = 1.0, double u = 1.0, double K = 1.0, double M = 1.0,
             double dt = 0.01, double kT = 0.01, double dV = 1.0)
{
	const double p[7] = {r, u, K, M, dt, std::sqrt(2.0 * kT / (dt * dV)), next_noise_seed()};
	run_scalar(grid, MMSPX_AC_NOISE, p, 7, steps);
}

// Model B, model_B.cpp:74-111: mu = -r*phi + u*phi^3 - K*lap(phi) + noise;
// new = phi + dt*M*lap(mu), noise = gaussian(0, sqrt(2 kT/(dt dV))).
template <int dim, typename T>
void model_b(MMSP::grid<dim, T>& grid, int steps, double r = 1.0, double u = 1.0, double K = 1.0, double M = 1.0,
             double dt = 0.01, double kT = 0.01, double dV = 1.0)
{
	const double p[7] = {r, u, K, M, dt, std::sqrt(2.0 * kT / (dt * dV)), next_noise_seed()};
	run_scalar(grid, MMSPX_CH_NOISE, p, 7, steps);
}

// Spinodal decomposition, spinodal.cpp:74-107: model B with r = u = K = 1 and
// dt/(2 dV) in place of dt*M (the same bits when dV is a power of two),
// noise = gaussian(0, sqrt(epsilon dt/dV)).
template <int dim, typename T>
void spinodal(MMSP::grid<dim, T>& grid, int steps, double dt = 0.01, double dV = 1.0, double epsilon = 0.05)
{
	const double p[7] = {1.0, 1.0, 1.0, 1.0 / (2.0 * dV), dt, std::sqrt(epsilon * dt / dV), next_noise_seed()};
	run_scalar(grid, MMSPX_CH_NOISE, p, 7, steps);
}

// Vector grids grid<dim,vector<T> >: the per-node vectors are packed
// node-major (the reference's node order) for the upload and unpacked after.
template <int dim, typename T>
void run_vector(MMSP::grid<dim, MMSP::vector<T> >& grid, int model, const double* p, int np, int steps)
{
	const int n = nodes(grid);
	const int q = fields(grid);
	if (n == 0 || steps <= 0) return;
	stats().calls += 1;
	mmspx_desc d = describe<dim>(grid, q, 1, precision_of<T>::get());
	std::vector<T> host((size_t)n * q);
	for (int i = 0; i < n; i++) {
		const MMSP::vector<T>& v = grid(i);
		if (MMSP::length(v) != q) {
			std::cerr << "mmsp_gpu: node " << i << " has " << MMSP::length(v) << " fields, expected " << q << std::endl;
			MMSP::Abort(-1);
		}
		for (int j = 0; j < q; j++) host[(size_t)i * q + j] = v[j];
	}
	// the per-node vectors live on the heap: the grid object keys the state
	resident& r = resident_for(&grid, &grid, d);
	upload_if_changed(r, host.data(), (int64_t)n * q, (int)sizeof(T));
	run_with_progress(r.s, model, p, np, steps);
	download_keep_digest(r, host.data());
	for (int i = 0; i < n; i++) {
		MMSP::vector<T>& v = grid(i);
		for (int j = 0; j < q; j++) v[j] = host[(size_t)i * q + j];
	}
}

// Grain growth on grid<dim,vector<T> >, graingrowth.cpp:73-111.
template <int dim, typename T> void grain_growth(MMSP::grid<dim, MMSP::vector<T> >& grid, int steps, double dt = 0.01)
{
	const double p[1] = {dt};
	run_vector(grid, MMSPX_GG, p, 1, steps);
}

// Grain growth with Zener pinning, zener_pinning/isotropic/phase_field/zener.cpp:104-141:
// field 0 (the particles) has zero mobility.
template <int dim, typename T> void zener(MMSP::grid<dim, MMSP::vector<T> >& grid, int steps, double dt = 0.01)
{
	const double p[1] = {dt};
	run_vector(grid, MMSPX_ZENER, p, 1, steps);
}

// Multi-phase grain growth, grain_growth/anisotropic/phase_field/graingrowth.cpp:94-191;
// gamma and mu: every pair's boundary energy and mobility (anisotropy.hpp gives 1).
template <int dim, typename T>
void grain_growth_multiphase(MMSP::grid<dim, MMSP::vector<T> >& grid, int steps, double dt = 0.01, double width = 8.0,
                             double gamma = 1.0, double mu = 1.0)
{
	const double p[4] = {dt, width, gamma, mu};
	run_vector(grid, MMSPX_GG_ANISO, p, 4, steps);
}

// Ostwald ripening, ostwald_ripening/isotropic/phase_field/ostwald.cpp:56-123;
// field 0 is the concentration. Defaults: the example's constants (:66-84).
template <int dim, typename T>
void ostwald(MMSP::grid<dim, MMSP::vector<T> >& grid, int steps, double dt = 0.001, double L = 1.0, double D = 1.0,
             double Calpha = 0.05, double Cbeta = 0.95, double A = 2.0, double delta = 1.0, double epsilon = 3.0,
             double kappa = 2.0)
{
	const double p[9] = {dt, L, D, Calpha, Cbeta, A, delta, epsilon, kappa};
	run_vector(grid, MMSPX_OSTWALD, p, 9, steps);
}

}  // namespace mmsp_gpu

#endif  // MMSP_GPU_HPP
