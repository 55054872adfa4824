// model_a_gpu.cpp -- the reference's model A program with its update() loop (model_A.cpp:74-105)
// replaced by the GPU drop-in. main(), the command line, the .dat reader/writer
// and checkpoint naming are the reference's own MMSP.main.hpp and headers. The
// noise comes from the library's counter-based generator, seeded per update()
// call from rand() (which MMSP.main.hpp seeds once), so runs agree with the
// stock program in distribution, not bit for bit.
#include "MMSP.hpp"
#include "model_A.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// model_A.cpp:85-91: r = u = K = M = 1, dt = 0.01, kT = 0.01, dV = 1
template <int dim, typename T> void update(grid<dim, T>& oldGrid, int steps)
{
	mmsp_gpu::model_a(oldGrid, steps, 1.0, 1.0, 1.0, 1.0, 0.01, 0.01, 1.0);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
