// cahn_hilliard_dw_gpu.cpp -- the reference's explicit Cahn-Hilliard program
// (cahn-hilliard.cpp:47-81: two passes per step, mu into a temporary grid,
// then c + dt*M*lap(mu)) with the benchmark's double-well chemical potential
// mu = 2c(1-c)(1-2c) - kappa*lap(c), kappa = M = 1, dt = 0.005, run by the GPU
// drop-in; main(), I/O and naming are the reference's own headers, unchanged.
#include "MMSP.hpp"
#include "cahn-hilliard.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use a seeded .dat input" << std::endl;
	MMSP::Abort(-1);
}

template <int dim, typename T> void update(grid<dim, T>& oldGrid, int steps)
{
	mmsp_gpu::cahn_hilliard_double_well(oldGrid, steps, 1.0, 1.0, 0.005);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
