// batch.hpp -- device work lists shared by the batched solver and the server.
#pragma once

#include <string>
#include <vector>

#include "tiles.hpp"

namespace mmspx {

// One member grid: its two state buffers, geometry and folded parameters.
// The source buffer of a launch is buf[(parity + phase) & 1], so members that
// joined at different global steps can share one launch.
struct BMember {
	void* buf[2];
	Geom g;
	ModelParams p;
	int phase;
};

// One unit of work: member m, tile origin (y0, z0) of the last axes, planes [xb, xe).
struct BItem {
	int m, y0, z0, xb, xe;
};

// Strip width and depth of the 2-D kernels, planes per item of the 3-D
// Cahn-Hilliard kernel, and the persistent grid size in workgroups per CU.
struct BatchTile {
	int tz, chunk, chunk3, bpc;
	int cells;   // > 0: 2-D Allen-Cahn by runs of `cells` consecutive nodes, one workgroup each
	int tb = 0;  // > 0: periodic 2-D Allen-Cahn by temporal blocks, up to `tb` steps per pass
};

bool batch_model_ok(int model, const Geom& g, std::string* why);
void member_items(int m, const Geom& g, int model, const BatchTile& t, std::vector<BItem>& out2,
                  std::vector<BItem>& out3);
// One pass over every item: list l2 (2-D strips/rows/blocks) and l3 (3-D
// tiles). A pass is one step, or with temporal blocks `nsteps` (<= t.tb) steps.
int launch_batch_step(int model, int prec, int fields, const BatchTile& t, const BMember* mem, const BItem* l2, int n2,
                      const BItem* l3, int n3, int parity, int cus, hipStream_t st, int64_t* launches, int nsteps = 1);
int device_cus();
// Reorders a spatially ordered work list so that the persistent launch (item
// p runs on workgroup p mod gridDim, and workgroup ids are dealt round-robin
// over the 8 XCDs) gives each XCD one contiguous run of it: neighbouring
// tiles then share their halo reads in one L2.
std::vector<BItem> xcd_order(const std::vector<BItem>& in);

}  // namespace mmspx
