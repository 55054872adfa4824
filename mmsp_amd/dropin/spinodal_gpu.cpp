// spinodal_gpu.cpp -- the reference's spinodal-decomposition program with its update() loop (spinodal.cpp:74-107)
// replaced by the GPU drop-in. main(), the command line, the .dat reader/writer
// and checkpoint naming are the reference's own MMSP.main.hpp and headers. The
// noise comes from the library's counter-based generator, seeded per update()
// call from rand() (which MMSP.main.hpp seeds once), so runs agree with the
// stock program in distribution, not bit for bit.
#include "MMSP.hpp"
#include "spinodal.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// spinodal.cpp:86-88: dt = 0.01, dV = 1, epsilon = 0.05
template <int dim, typename T> void update(grid<dim, T>& oldGrid, int steps)
{
	mmsp_gpu::spinodal(oldGrid, steps, 0.01, 1.0, 0.05);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
