// server.hip -- continuous batching of solves on one GPU.
//
// The reference runs one solve per process invocation (MMSP.main.hpp:38-354);
// here a worker thread owns the device and serves a stream of requests:
//   * submitted requests wait in a FIFO queue;
//   * at every round boundary the worker admits queued requests while slots
//     are free and their device footprint (two state buffers) fits the HBM
//     budget; a request that could never fit is rejected;
//   * admitted state stays resident on the device: every resident request
//     advances in the same launches (one persistent launch per step for the
//     2-D members, one for the 3-D members, all members' tiles in one work
//     list), so the work of requests that arrived at different times overlaps;
//   * a round runs until the next request finishes or `chunk_steps` steps,
//     whichever comes first; finished requests get their FP64 compensated mass
//     and their state copied out, and their memory is freed for admission.
// Each member's buffers alternate on the global step parity plus a phase set
// at admission, so joining mid-stream needs no copy. The tile bodies are the
// single-grid kernels' (tiles.hpp): a request's result does not depend on what
// else was resident.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"

namespace mmspx {

struct Request {
	int64_t id = 0;
	mmspx_desc desc{};
	Geom g{};
	ModelParams p{};
	int steps = 0;
	int done = 0;
	int status = 0;  // 0 queued, 1 running, 2 done, 3 rejected, 4 cancelled
	bool cancel = false;
	double t_submit = 0, t_admit = 0, t_done = 0;
	double mass0 = 0, mass1 = 0;
	size_t bytes = 0;  // one state buffer
	char* dev = nullptr;
	int phase = 0;
	std::vector<double> init;
	std::vector<double> result;
};

}  // namespace mmspx

using namespace mmspx;

struct mmspx_server {
	mmspx_server_config cfg{};
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<Request*> queue;
	std::map<int64_t, Request*> all;
	std::vector<Request*> active;
	std::thread worker;
	bool stop = false;
	int64_t next_id = 1;
	std::chrono::steady_clock::time_point t0;
	hipStream_t st = nullptr;
	int device = 0;
	int cus = 256;
	size_t in_use = 0;
	int64_t gstep = 0;
	BMember* d_mem = nullptr;
	int mem_cap = 0;
	BItem* d_items = nullptr;
	int items_cap = 0;
	double* part = nullptr;
	double* part_host = nullptr;
	void* conv = nullptr;  // staging for double <-> float conversion
	size_t conv_bytes = 0;
	int64_t stats[7] = {0, 0, 0, 0, 0, 0, 0};
	int worker_error = 0;
	// 64-wide 2-D strips: small 2-D requests would leave wider strips mostly idle
	BatchTile tile{64, 16, 16, 8, 0};

	double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static size_t elem_of(const mmspx_server* sv) { return sv->cfg.precision == MMSPX_F64 ? 8 : 4; }

static int ensure_conv(mmspx_server* sv, size_t bytes)
{
	if (sv->conv_bytes >= bytes) return 0;
	if (sv->conv) (void)hipFree(sv->conv);
	sv->conv = nullptr;
	sv->conv_bytes = 0;
	if (hipMalloc(&sv->conv, bytes) != hipSuccess) return fail(MMSPX_ENOMEM, "device allocation failed (staging)");
	sv->conv_bytes = bytes;
	return 0;
}

static double member_mass(mmspx_server* sv, Request* r, const void* state)
{
	if (launch_reduce(r->g, sv->cfg.precision, MMSPX_MASS, r->p, state, sv->part, sv->st)) return 0.0;
	if (hipMemcpyAsync(sv->part_host, sv->part, 2 * kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, sv->st) !=
	        hipSuccess ||
	    hipStreamSynchronize(sv->st) != hipSuccess)
		return 0.0;
	return sum_partials(sv->part_host, kRedBlocks);
}

// Device side of admission (worker thread, lock not held).
static int admit_upload(mmspx_server* sv, Request* r)
{
	if (hipMalloc((void**)&r->dev, 2 * r->bytes) != hipSuccess) return fail(MMSPX_ENOMEM, "device allocation failed");
	const size_t n = (size_t)r->g.nodes;
	int rc = ensure_conv(sv, n * sizeof(double));
	if (rc) return rc;
	MMSPX_HIP(hipMemcpyAsync(sv->conv, r->init.data(), n * sizeof(double), hipMemcpyHostToDevice, sv->st));
	// data goes to buf[0]; phase makes buf[(gstep + phase) & 1] == buf[0]
	rc = launch_convert(sv->conv, 1, r->dev, sv->cfg.precision == MMSPX_F64, (int64_t)n, sv->st);
	if (rc) return rc;
	r->phase = (int)(sv->gstep & 1);
	r->mass0 = member_mass(sv, r, r->dev);
	return 0;
}

static void* cur_state(const mmspx_server* sv, const Request* r)
{
	return r->dev + (((sv->gstep + r->phase) & 1) ? r->bytes : 0);
}

static int retire_download(mmspx_server* sv, Request* r)
{
	void* s = cur_state(sv, r);
	r->mass1 = member_mass(sv, r, s);
	const size_t n = (size_t)r->g.nodes;
	int rc = ensure_conv(sv, n * sizeof(double));
	if (rc) return rc;
	rc = launch_convert(s, sv->cfg.precision == MMSPX_F64, sv->conv, 1, (int64_t)n, sv->st);
	if (rc) return rc;
	r->result.resize(n);
	MMSPX_HIP(hipMemcpyAsync(r->result.data(), sv->conv, n * sizeof(double), hipMemcpyDeviceToHost, sv->st));
	MMSPX_HIP(hipStreamSynchronize(sv->st));
	return 0;
}

static void free_member(mmspx_server* sv, Request* r)
{
	if (r->dev) (void)hipFree(r->dev);
	r->dev = nullptr;
	std::lock_guard<std::mutex> lk(sv->mu);
	sv->in_use -= 2 * r->bytes;
}

static int run_round(mmspx_server* sv)
{
	std::vector<Request*> act;
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		act = sv->active;
	}
	int steps = sv->cfg.chunk_steps;
	for (Request* r : act) steps = std::min(steps, r->steps - r->done);
	if (steps <= 0) steps = 0;
	std::vector<BMember> mem(act.size());
	std::vector<BItem> l2, l3;
	for (size_t m = 0; m < act.size(); m++) {
		Request* r = act[m];
		mem[m].buf[0] = r->dev;
		mem[m].buf[1] = r->dev + r->bytes;
		mem[m].g = r->g;
		mem[m].p = r->p;
		mem[m].phase = r->phase;
		member_items((int)m, r->g, sv->cfg.model, sv->tile, l2, l3);
	}
	if ((int)mem.size() > sv->mem_cap) {
		if (sv->d_mem) (void)hipFree(sv->d_mem);
		sv->d_mem = nullptr;
		sv->mem_cap = 0;
		if (hipMalloc((void**)&sv->d_mem, sizeof(BMember) * mem.size()) != hipSuccess)
			return fail(MMSPX_ENOMEM, "device allocation failed (member table)");
		sv->mem_cap = (int)mem.size();
	}
	const int nit = (int)(l2.size() + l3.size());
	if (nit > sv->items_cap) {
		if (sv->d_items) (void)hipFree(sv->d_items);
		sv->d_items = nullptr;
		sv->items_cap = 0;
		if (hipMalloc((void**)&sv->d_items, sizeof(BItem) * (size_t)nit) != hipSuccess)
			return fail(MMSPX_ENOMEM, "device allocation failed (work list)");
		sv->items_cap = nit;
	}
	const int n2 = nit - (int)l3.size(), n3 = (int)l3.size();
	l2 = xcd_order(l2);
	{
		const std::vector<BItem> o3 = xcd_order(l3);
		l2.insert(l2.end(), o3.begin(), o3.end());
	}
	if (!mem.empty()) {
		MMSPX_HIP(hipMemcpyAsync(sv->d_mem, mem.data(), sizeof(BMember) * mem.size(), hipMemcpyHostToDevice, sv->st));
		MMSPX_HIP(hipMemcpyAsync(sv->d_items, l2.data(), sizeof(BItem) * l2.size(), hipMemcpyHostToDevice, sv->st));
	}
	int64_t launches = 0;
	for (int k = 0; k < steps && !act.empty(); k++) {
		int rc = launch_batch_step(sv->cfg.model, sv->cfg.precision, 1, sv->tile, sv->d_mem, sv->d_items, n2,
		                           sv->d_items + n2, n3, (int)((sv->gstep + k) & 1), sv->cus, sv->st, &launches);
		if (rc) return rc;
	}
	MMSPX_HIP(hipStreamSynchronize(sv->st));
	sv->gstep += steps;
	std::lock_guard<std::mutex> lk(sv->mu);
	for (Request* r : act) r->done += steps;
	sv->stats[3] += 1;
	sv->stats[6] += launches;
	return 0;
}

static void worker_main(mmspx_server* sv)
{
	(void)hipSetDevice(sv->device);
	for (;;) {
		std::vector<Request*> admit;
		{
			std::unique_lock<std::mutex> lk(sv->mu);
			if (sv->active.empty())
				sv->cv_work.wait_for(lk, std::chrono::milliseconds(50), [&] { return sv->stop || !sv->queue.empty(); });
			if (sv->stop) break;
			// admission: FIFO while slots and memory allow
			while (!sv->queue.empty() && (int)(sv->active.size() + admit.size()) < sv->cfg.max_slots) {
				Request* r = sv->queue.front();
				if (r->cancel) {
					sv->queue.pop_front();
					r->status = 4;
					r->t_done = sv->now();
					sv->cv_done.notify_all();
					continue;
				}
				const size_t need = 2 * r->bytes;
				if (need > sv->cfg.hbm_budget) {
					sv->queue.pop_front();
					r->status = 3;
					r->t_done = sv->now();
					sv->stats[2] += 1;
					sv->cv_done.notify_all();
					continue;
				}
				if (sv->in_use + need > sv->cfg.hbm_budget) break;
				sv->queue.pop_front();
				sv->in_use += need;
				sv->stats[4] = std::max<int64_t>(sv->stats[4], (int64_t)sv->in_use);
				admit.push_back(r);
			}
		}
		for (Request* r : admit) {
			int rc = admit_upload(sv, r);
			std::lock_guard<std::mutex> lk(sv->mu);
			if (rc) {
				sv->worker_error = rc;
				r->status = 3;
				r->t_done = sv->now();
				sv->stats[2] += 1;
				sv->cv_done.notify_all();
				if (r->dev) (void)hipFree(r->dev);
				r->dev = nullptr;
				sv->in_use -= 2 * r->bytes;
				continue;
			}
			r->status = 1;
			r->t_admit = sv->now();
			r->init.clear();
			r->init.shrink_to_fit();
			sv->active.push_back(r);
			sv->stats[0] += 1;
			sv->stats[5] = std::max<int64_t>(sv->stats[5], (int64_t)sv->active.size());
		}
		bool any;
		{
			std::lock_guard<std::mutex> lk(sv->mu);
			any = !sv->active.empty();
		}
		if (!any) continue;
		int rc = run_round(sv);
		if (rc) {
			std::lock_guard<std::mutex> lk(sv->mu);
			sv->worker_error = rc;
		}
		// retire finished and cancelled members
		std::vector<Request*> fin;
		{
			std::lock_guard<std::mutex> lk(sv->mu);
			std::vector<Request*> keep;
			for (Request* r : sv->active) {
				if (r->done >= r->steps || r->cancel || rc) fin.push_back(r);
				else keep.push_back(r);
			}
			sv->active.swap(keep);
		}
		for (Request* r : fin) {
			int rc2 = (r->cancel || rc) ? 0 : retire_download(sv, r);
			free_member(sv, r);
			std::lock_guard<std::mutex> lk(sv->mu);
			r->t_done = sv->now();
			r->status = (r->cancel || rc || rc2) ? (r->cancel ? 4 : 3) : 2;
			if (r->status == 2) sv->stats[1] += 1;
			sv->cv_done.notify_all();
		}
	}
}

extern "C" {

int mmspx_server_create(const mmspx_server_config* cfg, mmspx_server** out)
{
	if (!out || !cfg) return fail(MMSPX_EINVAL, "null argument");
	*out = nullptr;
	if (cfg->max_slots < 1 || cfg->chunk_steps < 1 || cfg->hbm_budget == 0)
		return fail(MMSPX_EINVAL, "max_slots, chunk_steps and hbm_budget must be positive");
	if (cfg->model != MMSPX_CH && cfg->model != MMSPX_CH_DW && cfg->model != MMSPX_AC)
		return fail(MMSPX_EINVAL, "the server runs Cahn-Hilliard or Allen-Cahn");
	if (cfg->precision != MMSPX_F32 && cfg->precision != MMSPX_F64)
		return fail(MMSPX_EINVAL, "the server runs MMSPX_F32 or MMSPX_F64");
	auto* sv = new (std::nothrow) mmspx_server();
	if (!sv) return fail(MMSPX_ENOMEM, "host allocation failed");
	sv->cfg = *cfg;
	sv->t0 = std::chrono::steady_clock::now();
	int rc = 0;
	if (hipGetDevice(&sv->device) != hipSuccess || hipStreamCreateWithFlags(&sv->st, hipStreamNonBlocking) != hipSuccess)
		rc = fail(MMSPX_EDEVICE, "HIP stream setup failed");
	if (rc == 0 && (hipMalloc((void**)&sv->part, 2 * kRedBlocks * sizeof(double)) != hipSuccess ||
	                hipHostMalloc((void**)&sv->part_host, 2 * kRedBlocks * sizeof(double), 0) != hipSuccess))
		rc = fail(MMSPX_ENOMEM, "device allocation failed");
	if (rc) {
		if (sv->part) (void)hipFree(sv->part);
		if (sv->part_host) (void)hipHostFree(sv->part_host);
		if (sv->st) (void)hipStreamDestroy(sv->st);
		delete sv;
		return rc;
	}
	sv->cus = device_cus();
	sv->worker = std::thread(worker_main, sv);
	*out = sv;
	return 0;
}

void mmspx_server_destroy(mmspx_server* sv)
{
	if (!sv) return;
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		sv->stop = true;
	}
	sv->cv_work.notify_all();
	if (sv->worker.joinable()) sv->worker.join();
	(void)hipStreamSynchronize(sv->st);
	for (auto& kv : sv->all) {
		if (kv.second->dev) (void)hipFree(kv.second->dev);
		delete kv.second;
	}
	if (sv->d_mem) (void)hipFree(sv->d_mem);
	if (sv->d_items) (void)hipFree(sv->d_items);
	if (sv->part) (void)hipFree(sv->part);
	if (sv->part_host) (void)hipHostFree(sv->part_host);
	if (sv->conv) (void)hipFree(sv->conv);
	if (sv->st) (void)hipStreamDestroy(sv->st);
	delete sv;
}

int mmspx_server_submit(mmspx_server* sv, const mmspx_desc* desc, const double* init, const double* params, int np,
                        int steps, int64_t* id)
{
	if (!sv || !desc || !init || !id || steps < 0) return fail(MMSPX_EINVAL, "bad arguments");
	int rc = validate_desc(desc);
	if (rc) return rc;
	if (desc->vector) return fail(MMSPX_EINVAL, "the server takes scalar grids");
	auto* r = new (std::nothrow) Request();
	if (!r) return fail(MMSPX_ENOMEM, "host allocation failed");
	r->desc = *desc;
	r->desc.precision = sv->cfg.precision;
	r->g = make_geom(r->desc);
	std::string why;
	if (!batch_model_ok(sv->cfg.model, r->g, &why)) {
		delete r;
		return fail(MMSPX_EUNSUPPORTED, why);
	}
	rc = make_params(sv->cfg.model, params, np, &r->p);
	if (rc) {
		delete r;
		return rc;
	}
	r->steps = steps;
	r->bytes = ((size_t)r->g.nodes * elem_of(sv) + 255) / 256 * 256;
	r->init.assign(init, init + r->g.nodes);
	std::lock_guard<std::mutex> lk(sv->mu);
	r->id = sv->next_id++;
	r->t_submit = sv->now();
	sv->all[r->id] = r;
	sv->queue.push_back(r);
	*id = r->id;
	sv->cv_work.notify_all();
	return 0;
}

static void fill_info(const Request* r, mmspx_request_info* info)
{
	info->id = r->id;
	info->status = r->status;
	info->steps_done = r->done;
	info->t_submit = r->t_submit;
	info->t_admit = r->t_admit;
	info->t_done = r->t_done;
	info->mass0 = r->mass0;
	info->mass1 = r->mass1;
}

int mmspx_server_cancel(mmspx_server* sv, int64_t id)
{
	if (!sv) return fail(MMSPX_EINVAL, "null server");
	std::lock_guard<std::mutex> lk(sv->mu);
	auto it = sv->all.find(id);
	if (it == sv->all.end()) return fail(MMSPX_ENOTFOUND, "unknown request id");
	it->second->cancel = true;
	sv->cv_work.notify_all();
	return 0;
}

int mmspx_server_wait(mmspx_server* sv, int64_t id, double timeout_s, mmspx_request_info* info)
{
	if (!sv) return fail(MMSPX_EINVAL, "null server");
	std::unique_lock<std::mutex> lk(sv->mu);
	auto it = sv->all.find(id);
	if (it == sv->all.end()) return fail(MMSPX_ENOTFOUND, "unknown request id");
	Request* r = it->second;
	auto finished = [&] { return r->status >= 2; };
	if (timeout_s < 0)
		sv->cv_done.wait(lk, finished);
	else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished))
		return fail(MMSPX_EBUSY, "request still running at timeout");
	if (info) fill_info(r, info);
	return 0;
}

int mmspx_server_info(mmspx_server* sv, int64_t id, mmspx_request_info* info)
{
	if (!sv || !info) return fail(MMSPX_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(sv->mu);
	auto it = sv->all.find(id);
	if (it == sv->all.end()) return fail(MMSPX_ENOTFOUND, "unknown request id");
	fill_info(it->second, info);
	return 0;
}

int mmspx_server_fetch(mmspx_server* sv, int64_t id, double* out, size_t count)
{
	if (!sv || !out) return fail(MMSPX_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(sv->mu);
	auto it = sv->all.find(id);
	if (it == sv->all.end()) return fail(MMSPX_ENOTFOUND, "unknown request id");
	Request* r = it->second;
	if (r->status != 2) return fail(MMSPX_EBUSY, "request has not completed");
	if (count != r->result.size()) return fail(MMSPX_EINVAL, "output size does not match the request's grid");
	std::copy(r->result.begin(), r->result.end(), out);
	sv->all.erase(it);
	delete r;
	return 0;
}

double mmspx_server_clock(mmspx_server* sv) { return sv ? sv->now() : 0.0; }

int mmspx_server_stats(mmspx_server* sv, int64_t* out, int nout)
{
	if (!sv || !out) return fail(MMSPX_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(sv->mu);
	for (int k = 0; k < nout && k < 7; k++) out[k] = sv->stats[k];
	return sv->worker_error;
}

}  // extern "C"
