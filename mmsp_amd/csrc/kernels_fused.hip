// kernels_fused.hip -- tiled / fused kernels for the hot shapes.
#include "device.hpp"

namespace mmspx {

bool fused_supported(const Geom&, int, int) { return false; }

TuneChoice default_choice(const Geom&, int, int) { return TuneChoice{1, 0, 0, 0}; }

int launch_fused_step(const Geom&, int, const ModelParams&, const void*, void*, hipStream_t, const TuneChoice&,
                      int64_t*)
{
	return fail(MMSPX_EUNSUPPORTED, "no fused kernel for this shape");
}

}  // namespace mmspx
