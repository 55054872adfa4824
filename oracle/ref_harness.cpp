// ref_harness.cpp -- TEST INFRASTRUCTURE, not product code.
//
// Drives the reference MMSP library (header-only, compiled from the sources
// where they lie) so that its own update() loops produce golden vectors and a
// CPU timing baseline. Built by oracle/Makefile into oracle/_ref/ only.
//
// One binary per model, selected at compile time:
//   -DHARNESS_AC  includes examples/phase_transitions/allen-cahn/allen-cahn.cpp
//   -DHARNESS_CH  includes examples/phase_transitions/cahn-hilliard/explicit/cahn-hilliard.cpp
//   -DHARNESS_GG  includes examples/coarsening/grain_growth/isotropic/phase_field/graingrowth.cpp
//   -DHARNESS_MA  includes examples/phase_transitions/model_A/model_A.cpp (stock only)
//   -DHARNESS_MB  includes examples/phase_transitions/model_B/model_B.cpp (stock only)
//   -DHARNESS_ZN  includes examples/coarsening/zener_pinning/isotropic/phase_field/zener.cpp (stock only)
//   -DHARNESS_OW  includes examples/coarsening/ostwald_ripening/isotropic/phase_field/ostwald.cpp (stock only)
//   -DHARNESS_GA  includes examples/coarsening/grain_growth/anisotropic/phase_field/graingrowth.cpp (stock only)
// MMSP_MAIN is predefined so that the reference's main() (MMSP.main.hpp:38) is
// skipped; this file supplies a small main() instead.
//
// Usage:
//   ref_xx run <in.dat> <out.dat|-> <steps> [stock|config] [p0 p1 ...]
// "stock" calls the example's own MMSP::update (hard-coded parameters).
// "config" calls the loops below, which keep the example's structure and
// evaluation order but take parameters (and, for CH, the double-well
// chemical potential mu = 2c(1-c)(1-2c) - kappa*lap(c)).
// The element type is taken from the file header (grid:double, grid:float,
// grid:vector:double, grid:vector:float) so the same template code runs for
// each. Prints one line "REFTIME <seconds> <nodes> <steps>" (update() only).

#define MMSP_MAIN
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#if defined(HARNESS_AC)
#include "allen-cahn.cpp"
#elif defined(HARNESS_CH)
#include "cahn-hilliard.cpp"
#elif defined(HARNESS_GG)
#include "graingrowth.cpp"
#elif defined(HARNESS_MA)
#include "model_A.cpp"
#elif defined(HARNESS_MB)
#include "model_B.cpp"
#elif defined(HARNESS_ZN)
#include "zener.cpp"
#elif defined(HARNESS_OW)
#include "ostwald.cpp"
#elif defined(HARNESS_GA)
#include "graingrowth.cpp"
#else
#error "define one of HARNESS_AC, HARNESS_CH, HARNESS_GG, HARNESS_MA, HARNESS_MB, HARNESS_ZN, HARNESS_OW, HARNESS_GA"
#endif

namespace harness {

using MMSP::grid;

static std::vector<double> P;  // config parameters from the command line
[[maybe_unused]] static double par(size_t i, double dflt) { return i < P.size() ? P[i] : dflt; }

#if defined(HARNESS_AC)
// Allen-Cahn with parameters; same expression and order as allen-cahn.cpp:70.
template <int dim, typename T> void update_config(grid<dim, T>& oldGrid, int steps)
{
	MMSP::ghostswap(oldGrid);
	grid<dim, T> newGrid(oldGrid);
	double r = par(0, 1.0), u = par(1, 1.0), K = par(2, 1.0), M = par(3, 1.0), dt = par(4, 0.01);
	for (int step = 0; step < steps; step++) {
		for (int i = 0; i < nodes(oldGrid); i++) {
			T phi = oldGrid(i);
			newGrid(i) = phi - dt * M * (-r * phi + u * pow(phi, 3) - K * MMSP::laplacian(oldGrid, i));
		}
		MMSP::swap(oldGrid, newGrid);
		MMSP::ghostswap(oldGrid);
	}
}
#endif

#if defined(HARNESS_CH)
// Cahn-Hilliard, two stencil passes as cahn-hilliard.cpp:69-79, with the
// double-well chemical potential mu = 2c(1-c)(1-2c) - kappa*lap(c).
template <int dim, typename T> void update_config(grid<dim, T>& oldGrid, int steps)
{
	MMSP::ghostswap(oldGrid);
	grid<dim, T> newGrid(oldGrid);
	grid<dim, T> temp(oldGrid);
	double kappa = par(0, 1.0), M = par(1, 1.0), dt = par(2, 0.005);
	for (int step = 0; step < steps; step++) {
		for (int i = 0; i < nodes(oldGrid); i++) {
			T c = oldGrid(i);
			temp(i) = 2.0 * c * (1.0 - c) * (1.0 - 2.0 * c) - kappa * MMSP::laplacian(oldGrid, i);
		}
		MMSP::ghostswap(temp);
		for (int i = 0; i < nodes(oldGrid); i++)
			newGrid(i) = oldGrid(i) + dt * M * MMSP::laplacian(temp, i);
		MMSP::swap(oldGrid, newGrid);
		MMSP::ghostswap(oldGrid);
	}
}
#endif

#if defined(HARNESS_GG)
// Grain growth exactly as graingrowth.cpp:86-110 with dt as a parameter.
template <int dim, typename T> void update_config(grid<dim, MMSP::vector<T> >& oldGrid, int steps)
{
	MMSP::ghostswap(oldGrid);
	grid<dim, MMSP::vector<T> > newGrid(oldGrid);
	double dt = par(0, 0.01);
	for (int step = 0; step < steps; step++) {
		for (int i = 0; i < nodes(oldGrid); i++) {
			const MMSP::vector<T>& phi = oldGrid(i);
			MMSP::vector<T> lap = MMSP::laplacian(oldGrid, i);
			T sum = 0.0;
			for (int j = 0; j < fields(oldGrid); j++)
				sum += phi[j] * phi[j];
			for (int j = 0; j < fields(oldGrid); j++)
				newGrid(i)[j] = phi[j] - dt * (-phi[j] - pow(phi[j], 3) + 2.0 * (phi[j] * sum - lap[j]));
		}
		MMSP::swap(oldGrid, newGrid);
		MMSP::ghostswap(oldGrid);
	}
}
#endif

#if defined(HARNESS_MA) || defined(HARNESS_MB) || defined(HARNESS_ZN) || defined(HARNESS_OW) || defined(HARNESS_GA)
// model A / model B, Zener pinning, Ostwald ripening: only the stock update
// (hard-coded parameters); model A/B's noise comes from rand(), seeded by
// srand(HARNESS_SEED) when that is set
template <class G> void update_config(G&, int)
{
	std::fprintf(stderr, "ref harness: this example runs the stock update only\n");
	std::exit(2);
}
#endif

template <class G> int run(const char* in, const char* out, int steps, bool stock)
{
	G g(in);
	if (const char* sd = std::getenv("HARNESS_SEED")) std::srand((unsigned)std::atoi(sd));
	auto t0 = std::chrono::steady_clock::now();
	if (stock)
		MMSP::update(g, steps);
	else
		update_config(g, steps);
	auto t1 = std::chrono::steady_clock::now();
	double sec = std::chrono::duration<double>(t1 - t0).count();
	std::printf("\nREFTIME %.9f %d %d\n", sec, nodes(g), steps);
	std::fflush(stdout);
	if (std::string(out) != "-")
		MMSP::output(g, out);
	return 0;
}

template <int dim> int dispatch_type(const std::string& type, const char* in, const char* out, int steps, bool stock)
{
#if defined(HARNESS_GG) || defined(HARNESS_ZN) || defined(HARNESS_OW) || defined(HARNESS_GA)
	if (type == "grid:vector:double") return run<grid<dim, MMSP::vector<double> > >(in, out, steps, stock);
	if (type == "grid:vector:float") return run<grid<dim, MMSP::vector<float> > >(in, out, steps, stock);
#else
	if (type == "grid:double") return run<grid<dim, double> >(in, out, steps, stock);
	if (type == "grid:float") return run<grid<dim, float> >(in, out, steps, stock);
#endif
	std::fprintf(stderr, "ref harness: unsupported grid type %s\n", type.c_str());
	return 2;
}

} // namespace harness

int main(int argc, char* argv[])
{
	if (argc < 5 || std::string(argv[1]) != "run") {
		std::fprintf(stderr, "usage: %s run <in.dat> <out.dat|-> <steps> [stock|config] [params...]\n", argv[0]);
		return 2;
	}
	const char* in = argv[2];
	const char* out = argv[3];
	int steps = std::atoi(argv[4]);
	bool stock = true;
	if (argc > 5) stock = std::string(argv[5]) != "config";
	for (int i = 6; i < argc; i++) harness::P.push_back(std::atof(argv[i]));

	std::ifstream f(in);
	if (!f) {
		std::fprintf(stderr, "ref harness: cannot open %s\n", in);
		return 2;
	}
	std::string type;
	int dim = 0;
	std::getline(f, type, '\n');
	f >> dim;
	f.close();
	if (dim == 1) return harness::dispatch_type<1>(type, in, out, steps, stock);
	if (dim == 2) return harness::dispatch_type<2>(type, in, out, steps, stock);
	if (dim == 3) return harness::dispatch_type<3>(type, in, out, steps, stock);
	std::fprintf(stderr, "ref harness: bad dimension %d\n", dim);
	return 2;
}
