// cahn_hilliard_gpu.cpp -- the reference's explicit Cahn-Hilliard program with
// its two-pass update() (cahn-hilliard.cpp:47-81) replaced by the GPU drop-in;
// main(), I/O and naming are the reference's own headers, unchanged.
#include "MMSP.hpp"
#include "cahn-hilliard.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// cahn-hilliard.cpp:59-63: r = u = K = M = 1, dt = 0.01
template <int dim, typename T> void update(grid<dim, T>& oldGrid, int steps)
{
	mmsp_gpu::cahn_hilliard(oldGrid, steps, 1.0, 1.0, 1.0, 1.0, 0.01);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
