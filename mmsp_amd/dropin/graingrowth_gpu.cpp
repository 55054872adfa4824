// graingrowth_gpu.cpp -- the reference's isotropic phase-field grain growth
// program with its update() (graingrowth.cpp:73-111) replaced by the GPU
// drop-in; main(), I/O and naming are the reference's own headers, unchanged.
#include "MMSP.hpp"
#include "graingrowth.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// graingrowth.cpp:84: dt = 0.01
template <int dim, typename T> void update(grid<dim, vector<T> >& oldGrid, int steps)
{
	mmsp_gpu::grain_growth(oldGrid, steps, 0.01);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
