// zener_gpu.cpp -- the reference's Zener-pinning program with its update() loop
// (zener_pinning/isotropic/phase_field/zener.cpp:104-141) replaced by the GPU drop-in. main(), the
// command line, the .dat reader/writer and checkpoint naming are the
// reference's own MMSP.main.hpp and headers, compiled unchanged.
#include "MMSP.hpp"
#include "zener.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// zener.cpp:115: dt = 0.01
template <int dim, typename T> void update(grid<dim, vector<T> >& oldGrid, int steps)
{
	mmsp_gpu::zener(oldGrid, steps, 0.01);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
