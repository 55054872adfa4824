// graingrowth_aniso_gpu.cpp -- the reference's multi-phase (anisotropic) grain
// growth program with its update() loop
// (grain_growth/anisotropic/phase_field/graingrowth.cpp:94-191) replaced by the
// GPU drop-in. main(), the command line, the .dat reader/writer and checkpoint
// naming are the reference's own MMSP.main.hpp and headers.
#include "MMSP.hpp"
#include "graingrowth.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// graingrowth.cpp:105-106: dt = 0.01, width = 8; anisotropy.hpp: energy and mobility 1
template <int dim, typename T> void update(grid<dim, vector<T> >& oldGrid, int steps)
{
	mmsp_gpu::grain_growth_multiphase(oldGrid, steps, 0.01, 8.0, 1.0, 1.0);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
