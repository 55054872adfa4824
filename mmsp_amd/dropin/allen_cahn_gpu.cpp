// allen_cahn_gpu.cpp -- the reference's Allen-Cahn program with its update()
// loop (allen-cahn.cpp:47-75) replaced by the GPU drop-in. Everything else --
// main(), the command line, the .dat reader/writer, checkpoint naming -- is
// the reference's own MMSP.main.hpp and MMSP headers, compiled unchanged.
// generate() is not part of the hot path; run the stock program for --example.
#include "MMSP.hpp"
#include "allen-cahn.hpp"
#include "mmsp_gpu.hpp"

namespace MMSP {

void generate(int, const char*)
{
	std::cerr << PROGRAM << " (GPU update): use the stock program to generate example grids" << std::endl;
	MMSP::Abort(-1);
}

// allen-cahn.cpp:58-62: r = u = K = M = 1, dt = 0.01
template <int dim, typename T> void update(grid<dim, T>& oldGrid, int steps)
{
	mmsp_gpu::allen_cahn(oldGrid, steps, 1.0, 1.0, 1.0, 1.0, 0.01);
}

}  // namespace MMSP

#include "MMSP.main.hpp"
