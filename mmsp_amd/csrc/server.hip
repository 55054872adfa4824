// server.hip -- placeholder until the serving loop lands.
#include "core.hpp"
using namespace mmspx;
extern "C" {
int mmspx_server_create(const mmspx_server_config*, mmspx_server** out) { if (out) *out = nullptr; return fail(MMSPX_EUNSUPPORTED, "server not built"); }
void mmspx_server_destroy(mmspx_server*) {}
int mmspx_server_submit(mmspx_server*, const mmspx_desc*, const double*, const double*, int, int, int64_t*) { return MMSPX_EUNSUPPORTED; }
int mmspx_server_cancel(mmspx_server*, int64_t) { return MMSPX_EUNSUPPORTED; }
int mmspx_server_wait(mmspx_server*, int64_t, double, mmspx_request_info*) { return MMSPX_EUNSUPPORTED; }
int mmspx_server_fetch(mmspx_server*, int64_t, double*, size_t) { return MMSPX_EUNSUPPORTED; }
int mmspx_server_info(mmspx_server*, int64_t, mmspx_request_info*) { return MMSPX_EUNSUPPORTED; }
double mmspx_server_clock(mmspx_server*) { return 0.0; }
int mmspx_server_stats(mmspx_server*, int64_t*, int) { return MMSPX_EUNSUPPORTED; }
}
