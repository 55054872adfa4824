// batch.hip -- placeholder until the batched solver lands.
#include "core.hpp"
using namespace mmspx;
extern "C" {
int mmspx_batch_create(const mmspx_desc*, int, int, mmspx_batch** out) { if (out) *out = nullptr; return fail(MMSPX_EUNSUPPORTED, "batch not built"); }
void mmspx_batch_destroy(mmspx_batch*) {}
int mmspx_batch_upload(mmspx_batch*, int, const void*, int) { return MMSPX_EUNSUPPORTED; }
int mmspx_batch_download(mmspx_batch*, int, void*, int) { return MMSPX_EUNSUPPORTED; }
int mmspx_batch_set_params(mmspx_batch*, const double*, int) { return MMSPX_EUNSUPPORTED; }
int mmspx_batch_step(mmspx_batch*, int, int) { return MMSPX_EUNSUPPORTED; }
int mmspx_batch_step_timed(mmspx_batch*, int, int, double*) { return MMSPX_EUNSUPPORTED; }
int mmspx_batch_reduce(mmspx_batch*, int, int, double*, int) { return MMSPX_EUNSUPPORTED; }
int mmspx_batch_autotune(mmspx_batch*, int, char*, size_t) { return MMSPX_EUNSUPPORTED; }
int mmspx_batch_get_tile(const mmspx_batch*, int*, int*, int*) { return MMSPX_EUNSUPPORTED; }
}
