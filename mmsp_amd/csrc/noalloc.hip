// noalloc.hip -- mmspx_step_noalloc: explicit steps on caller-owned device
// buffers, with no device allocation by the library.
//
// mmspx_step runs the same kernels on a state the library allocated. Two of
// its paths need device memory beyond the state: the two-pass Cahn-Hilliard
// kernels keep the chemical potential in a state-sized scalar plane, and the
// two-pass Ostwald kernels keep w (one scalar plane). Here that plane is
// `scratch`, a file-scope device array of fixed size that each GPU holds from
// the library's load to the end of the process. Every other path (the fused
// and strip kernels) needs no scratch and has no size limit here.
//
// Host code never passes `scratch` itself to a kernel: in host code the name
// is a stand-in, not a device address. The first call on a device resolves
// the array's address on that device (hipGetSymbolAddress, with the device
// current) and keeps it per device; kernels take that pointer.
//
// Calls on one GPU share its array, which keeps the last call's contents;
// nothing in it has to start at zero (each pass writes the whole plane before
// the next pass reads it). Calls that overlap on one GPU (two host threads,
// or two streams whose work overlaps) race on it.
#include <mutex>

#include "core.hpp"

namespace mmspx {

static __device__ __align__(256) unsigned char scratch[MMSPX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The array's address on the current device (resolved once per device).
static int scratch_on_current_device(void** out)
{
	int dev = -1;
	MMSPX_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices) return fail(MMSPX_EUNSUPPORTED, "device index beyond the library's table");
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		void* ptr = nullptr;
		hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr)
			return fail(MMSPX_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
		g_scratch_ptr[dev] = ptr;
	}
	*out = g_scratch_ptr[dev];
	return 0;
}

}  // namespace mmspx

using namespace mmspx;

extern "C" int mmspx_step_noalloc(const mmspx_desc* desc, void* buf0, void* buf1, int model, const double* params,
                                  int nparams, int steps, int64_t first_step, void* stream, int* result_buf)
{
	if (!result_buf) return fail(MMSPX_EINVAL, "null result_buf");
	*result_buf = 0;
	int rc = validate_desc(desc);
	if (rc) return rc;
	if (!buf0 || !buf1 || buf0 == buf1) return fail(MMSPX_EINVAL, "two distinct device buffers are needed");
	if (steps < 0 || first_step < 0) return fail(MMSPX_EINVAL, "steps and first_step must be >= 0");
	ModelParams p;
	rc = make_params(model, params, nparams, &p);
	if (rc) return rc;
	const Geom g = make_geom(*desc);
	rc = validate_model(*desc, g, model);
	if (rc) return rc;
	// resolved before any step, so a call with steps = 0 prepares the device
	// (for callers that capture later calls into a graph)
	void* sp = nullptr;
	rc = scratch_on_current_device(&sp);
	if (rc) return rc;
	// the kernels mmspx_step picks by default for this model and shape
	TuneChoice tc{1, 0, 0, 0, 0};
	bool fused = false;
	if (fused_supported(g, desc->precision, model)) {
		tc = default_choice(g, desc->precision, model);
		fused = tc.variant == 2;
	}
	const size_t elem = desc->precision == MMSPX_F64 ? 8 : 4;
	const bool needs_plane =
	    (!fused && (ch_family(model) || model == MMSPX_OSTWALD)) || (fused && fused_needs_plane(g, model, tc));
	if (needs_plane && (size_t)g.nodes * elem > (size_t)MMSPX_NOALLOC_SCRATCH_BYTES)
		return fail(MMSPX_EINVAL, "this grid's two-pass kernels need a " + std::to_string((size_t)g.nodes * elem) +
		                              "-byte scalar plane; mmspx_step_noalloc holds at most " +
		                              std::to_string((size_t)MMSPX_NOALLOC_SCRATCH_BYTES) +
		                              " bytes (use mmspx_step on a state)");
	hipStream_t st = (hipStream_t)stream;
	const int spl = fused ? fused_steps_per_launch(model, tc) : 1;
	void *a = buf0, *b = buf1;
	int which = 0;
	int64_t launches = 0, done = 0;
	while (steps > 0) {
		const int n = std::min(steps, spl);
		p.step = (uint64_t)(first_step + done);
		rc = fused ? launch_fused_step(g, desc->precision, p, a, b, st, tc, n, &launches, needs_plane ? sp : nullptr)
		           : launch_generic_step(g, desc->precision, p, a, b, needs_plane ? sp : nullptr, st, &launches);
		if (rc) return rc;
		std::swap(a, b);
		which ^= 1;
		done += n;
		steps -= n;
	}
	*result_buf = which;
	return 0;
}
