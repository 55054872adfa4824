// ostwald_gpu.cpp -- the reference's Ostwald-ripening program with its update() loop
// (ostwald_ripening/isotropic/phase_field/ostwald.cpp:56-123) replaced by the GPU drop-in. main(), the
// command line, the .dat reader/writer and checkpoint naming are the
// reference's own MMSP.main.hpp and headers, compiled unchanged.
#include "MMSP.hpp"
#include "ostwald.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// ostwald.cpp:68-82: dt = 0.001, L = D = 1, Calpha = 0.05, Cbeta = 0.95, A = 2,
// delta = 1, epsilon = 3, kappa = 2
template <int dim, typename T> void update(grid<dim, vector<T> >& oldGrid, int steps)
{
	mmsp_gpu::ostwald(oldGrid, steps, 0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
