// server.hip -- continuous batching of solves on one GPU.
//
// The reference runs one solve per process invocation (MMSP.main.hpp:38-354);
// here the library serves a stream of requests with three threads, each with
// its own stream:
//   * submitted requests wait in a FIFO queue, their initial state already in
//     the request's precision (the submitting thread converts it). A request
//     runs in the server's precision, except that an FP32 server runs the
//     requests of more than `fp64_above_steps` steps in FP64: the FP32
//     trajectory of a spinodal decomposition drifts from the FP64 one as the
//     instability amplifies the rounding of the stored state;
//   * the worker admits queued requests while slots are free and their device
//     footprint (two state buffers) fits the HBM budget; a request that could
//     never fit is rejected;
//   * the uploader allocates an admitted request's buffers (stream-ordered,
//     from the server's own pool), copies its state in and takes its FP64
//     compensated mass; the request is then resident;
//   * the worker advances resident requests in rounds. Each round takes the
//     requests with the least remaining work until it holds enough nodes per
//     step to fill the GPU (plus any request left out for too long); it lasts
//     until one of them finishes or `chunk_steps` steps, and is one
//     persistent launch per step over every member's tiles (2-D and 3-D
//     members in two launches). 3-D members of 2^21 nodes or more fill the
//     GPU on their own: they run the single-grid kernels instead (for
//     periodic Cahn-Hilliard two steps per pass), one launch per member and
//     pass, on a second compute stream, so they overlap the batch launches
//     (both streams at the default priority). A round is planned on the host
//     and queued on the
//     compute stream; up to two rounds are queued ahead, so the device does
//     not wait for the host between rounds;
//   * when a request's last round is queued, the downloader waits for that
//     round, takes the FP64 mass of the final state, copies it out and frees
//     its buffers; only then is its memory counted free for admission.
// Uploads and downloads therefore overlap the compute of the other requests.
// A batched member's state after s steps sits in buf[s & 1], so joining
// mid-stream or sitting out rounds needs no copy; a member on the single-grid
// kernels keeps the index of the buffer holding its state (a two-step pass
// swaps the buffers once). The tile bodies are the
// single-grid kernels' (tiles.hpp): a request's result does not depend on what
// else was resident.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "batch.hpp"

namespace mmspx {

constexpr int kSlots = 4;        // work-list buffers cycled by queued rounds
constexpr int kRoundsAhead = 2;  // rounds queued on the compute stream at most
constexpr int64_t kRoundNodes = (int64_t)1 << 25;  // nodes per step a round aims to hold
constexpr int kMaxIdleRounds = 64;                 // a resident request advances at least this often
constexpr int64_t kSoloNodes = (int64_t)1 << 21;   // 3-D members this large run the single-grid kernels

struct Request {
	int64_t id = 0;
	mmspx_desc desc{};
	Geom g{};
	ModelParams p{};
	int prec = MMSPX_F32;  // the form the request runs in
	int steps = 0;
	int done = 0;     // steps of rounds the device has finished
	int planned = 0;  // steps queued
	int status = 0;   // 0 queued, 1 running, 2 done, 3 rejected, 4 cancelled
	bool cancel = false;
	double t_submit = 0, t_admit = 0, t_done = 0;
	double mass0 = 0, mass1 = 0;
	size_t bytes = 0;     // one state buffer (request precision, padded to 256 B)
	char* dev = nullptr;  // buf[0] at dev, buf[1] at dev + bytes
	int idle_rounds = 0;             // rounds planned since it last advanced
	bool solo = false;               // runs the single-grid kernels with `tile`
	TuneChoice tile{};
	int cur = 0;                     // solo: the buffer holding the state after `planned` steps
	int in_rounds = 0;               // queued rounds that list it (under the server's mutex)
	bool fetched = false;            // fetched while rounds still listed it: the worker deletes it
	hipEvent_t ev_last = nullptr;    // recorded after its last round (or when dropped)
	std::vector<char> init;          // initial state, until uploaded
	std::vector<char> out;           // final state
};

struct Slot {
	BMember* d_mem = nullptr;
	BItem* d_items = nullptr;
	BMember* h_mem = nullptr;  // pinned mirrors, copied asynchronously
	BItem* h_items = nullptr;
	int mem_cap = 0, items_cap = 0;
	hipEvent_t done = nullptr;  // the last round that read this slot
	bool used = false;
};

struct Round {
	hipEvent_t ev = nullptr;   // the batch launches (compute stream)
	hipEvent_t ev2 = nullptr;  // the large members' launches (second compute stream), if any
	int steps = 0;
	std::vector<Request*> members;
};

// A copy-side job: a request's retirement (mass, copy out, free) or, for a
// cancelled one, only the free.
struct IoJob {
	Request* r;
	bool drop;
};

}  // namespace mmspx

using namespace mmspx;

struct mmspx_server {
	mmspx_server_config cfg{};
	std::mutex mu;
	std::condition_variable cv_work, cv_done, cv_up, cv_down;
	std::deque<Request*> queue;
	std::map<int64_t, Request*> all;
	std::thread worker, uploader, downloader;
	bool stop = false;
	int64_t next_id = 1;
	std::chrono::steady_clock::time_point t0;
	hipStream_t st = nullptr;   // compute (worker): batch launches
	hipStream_t st2 = nullptr;  // compute (worker): the large members' single-grid launches
	hipStream_t ust = nullptr;  // uploads
	hipStream_t dst = nullptr;  // downloads
	hipMemPool_t pool = nullptr;
	double* upart = nullptr;    // reduction partials of the uploader and the downloader
	double* dpart = nullptr;
	int device = 0;
	int cus = 256;
	size_t in_use = 0;
	int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	int worker_error = 0;
	// 2-D members by 60-column wave strips of 8 rows, 3-D members by 16-plane chunks
	BatchTile tile{kChWaveOut, 8, 16, 8, 0};

	// hand-offs between the threads (under mu)
	std::deque<Request*> up_jobs;   // admitted, to upload
	std::vector<Request*> ready;    // uploaded, to join the next round
	std::deque<IoJob> down_jobs;
	int uploading = 0;              // admitted and not yet ready

	// worker-only state
	std::vector<Request*> resident;
	std::deque<Round> rounds;
	int64_t gstep = 0;  // global steps queued so far
	Slot slot[kSlots];
	int next_slot = 0;

	double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static size_t elem_of(const Request* r) { return r->prec == MMSPX_F64 ? 8 : 4; }

// the compute stream that advances a request
static hipStream_t stream_of(const mmspx_server* sv, const Request* r) { return r->solo ? sv->st2 : sv->st; }

static void finish(mmspx_server* sv, Request* r, int status)
{
	// caller holds sv->mu
	r->status = status;
	r->t_done = sv->now();
	if (status == 2) sv->stats[1] += 1;
	if (status == 3) sv->stats[2] += 1;
	sv->cv_done.notify_all();
}

// ---- uploader ----------------------------------------------------------------

static int upload_one(mmspx_server* sv, Request* r, std::vector<double>& part_host)
{
	MMSPX_HIP(hipMallocFromPoolAsync((void**)&r->dev, 2 * r->bytes, sv->pool, sv->ust));
	// the data goes to buf[0], the state after 0 steps
	MMSPX_HIP(hipMemcpyAsync(r->dev, r->init.data(), r->init.size(), hipMemcpyHostToDevice, sv->ust));
	int rc = launch_reduce(r->g, r->prec, MMSPX_MASS, r->p, r->dev, sv->upart, sv->ust);
	if (rc) return rc;
	MMSPX_HIP(hipMemcpyAsync(part_host.data(), sv->upart, part_host.size() * sizeof(double), hipMemcpyDeviceToHost,
	                         sv->ust));
	MMSPX_HIP(hipStreamSynchronize(sv->ust));
	r->mass0 = sum_partials(part_host.data(), kRedBlocks);
	return 0;
}

static void uploader_main(mmspx_server* sv)
{
	(void)hipSetDevice(sv->device);
	std::vector<double> part_host(2 * kRedBlocks);
	for (;;) {
		Request* r;
		{
			std::unique_lock<std::mutex> lk(sv->mu);
			sv->cv_up.wait(lk, [&] { return sv->stop || !sv->up_jobs.empty(); });
			if (sv->stop) return;
			r = sv->up_jobs.front();
			sv->up_jobs.pop_front();
		}
		const int rc = upload_one(sv, r, part_host);
		std::lock_guard<std::mutex> lk(sv->mu);
		sv->uploading -= 1;
		if (rc) {
			if (!sv->worker_error) sv->worker_error = rc;
			if (r->dev) (void)hipFreeAsync(r->dev, sv->ust);
			r->dev = nullptr;
			sv->in_use -= 2 * r->bytes;
			finish(sv, r, 3);
		} else {
			r->init.clear();
			r->init.shrink_to_fit();
			r->status = 1;
			r->t_admit = sv->now();
			sv->stats[0] += 1;
			sv->ready.push_back(r);
		}
		sv->cv_work.notify_all();
	}
}

// ---- downloader --------------------------------------------------------------

static int download_one(mmspx_server* sv, Request* r, std::vector<double>& part_host)
{
	MMSPX_HIP(hipStreamWaitEvent(sv->dst, r->ev_last, 0));
	const int last = r->solo ? r->cur : (r->steps & 1);
	const void* s = r->dev + (last ? r->bytes : 0);
	int rc = launch_reduce(r->g, r->prec, MMSPX_MASS, r->p, s, sv->dpart, sv->dst);
	if (rc) return rc;
	MMSPX_HIP(hipMemcpyAsync(part_host.data(), sv->dpart, part_host.size() * sizeof(double), hipMemcpyDeviceToHost,
	                         sv->dst));
	r->out.resize((size_t)r->g.nodes * elem_of(r));
	MMSPX_HIP(hipMemcpyAsync(r->out.data(), s, r->out.size(), hipMemcpyDeviceToHost, sv->dst));
	MMSPX_HIP(hipStreamSynchronize(sv->dst));
	r->mass1 = sum_partials(part_host.data(), kRedBlocks);
	return 0;
}

static void downloader_main(mmspx_server* sv)
{
	(void)hipSetDevice(sv->device);
	std::vector<double> part_host(2 * kRedBlocks);
	for (;;) {
		IoJob job;
		{
			std::unique_lock<std::mutex> lk(sv->mu);
			sv->cv_down.wait(lk, [&] { return sv->stop || !sv->down_jobs.empty(); });
			if (sv->stop) return;
			job = sv->down_jobs.front();
			sv->down_jobs.pop_front();
		}
		Request* r = job.r;
		int rc = 0;
		if (job.drop)
			rc = hipStreamWaitEvent(sv->dst, r->ev_last, 0) == hipSuccess ? 0 : fail(MMSPX_EDEVICE, "hipStreamWaitEvent failed");
		else
			rc = download_one(sv, r, part_host);
		// after the wait above, the free follows every launch that used the buffers
		if (r->dev && hipFreeAsync(r->dev, sv->dst) != hipSuccess && !rc) rc = fail(MMSPX_EDEVICE, "hipFreeAsync failed");
		(void)hipEventDestroy(r->ev_last);
		std::lock_guard<std::mutex> lk(sv->mu);
		r->dev = nullptr;
		r->ev_last = nullptr;
		if (!job.drop) r->done = r->steps;
		sv->in_use -= 2 * r->bytes;
		if (rc && !sv->worker_error) sv->worker_error = rc;
		finish(sv, r, job.drop ? 4 : (rc ? 3 : 2));
		sv->cv_work.notify_all();
	}
}

// ---- worker ------------------------------------------------------------------

static int ensure_slot(Slot& sl, int nmem, int nitems)
{
	if (nmem > sl.mem_cap) {
		const int cap = std::max(nmem, 2 * sl.mem_cap + 16);
		if (sl.d_mem) (void)hipFree(sl.d_mem);
		if (sl.h_mem) (void)hipHostFree(sl.h_mem);
		sl.d_mem = nullptr;
		sl.h_mem = nullptr;
		sl.mem_cap = 0;
		if (hipMalloc((void**)&sl.d_mem, sizeof(BMember) * cap) != hipSuccess ||
		    hipHostMalloc((void**)&sl.h_mem, sizeof(BMember) * cap, 0) != hipSuccess)
			return fail(MMSPX_ENOMEM, "allocation failed (member table)");
		sl.mem_cap = cap;
	}
	if (nitems > sl.items_cap) {
		const int cap = std::max(nitems, 2 * sl.items_cap + 1024);
		if (sl.d_items) (void)hipFree(sl.d_items);
		if (sl.h_items) (void)hipHostFree(sl.h_items);
		sl.d_items = nullptr;
		sl.h_items = nullptr;
		sl.items_cap = 0;
		if (hipMalloc((void**)&sl.d_items, sizeof(BItem) * cap) != hipSuccess ||
		    hipHostMalloc((void**)&sl.h_items, sizeof(BItem) * cap, 0) != hipSuccess)
			return fail(MMSPX_ENOMEM, "allocation failed (work list)");
		sl.items_cap = cap;
	}
	return 0;
}

// Picks the members that advance in the next round: resident requests in
// order of least remaining work (nodes x steps left) until the round holds
// kRoundNodes nodes per step, enough to fill the GPU, plus any request that
// has sat out kMaxIdleRounds rounds. Short requests thus finish in a few
// rounds instead of sharing every launch with the longest ones, and each
// launch is still wide enough to keep the chip busy.
static std::vector<Request*> pick_round(mmspx_server* sv)
{
	std::vector<Request*> order = sv->resident;
	auto left = [](const Request* r) { return (double)r->g.nodes * (double)(r->steps - r->planned); };
	std::stable_sort(order.begin(), order.end(), [&](const Request* x, const Request* y) { return left(x) < left(y); });
	std::vector<Request*> pick;
	int64_t nodes = 0;
	for (Request* r : order) {
		if (nodes < kRoundNodes || r->idle_rounds >= kMaxIdleRounds) {
			pick.push_back(r);
			nodes += r->g.nodes;
		}
	}
	return pick;
}

// Plans one round over the picked members and queues it on the compute stream;
// members whose last step it holds are handed to the downloader. A member's
// state after s steps sits in buf[s & 1] (buf[0] holds the uploaded state),
// so members that sat out rounds need no copy: the phase sent with each
// round maps the launch parity onto the member's own step count.
static int enqueue_round(mmspx_server* sv)
{
	std::vector<Request*> act = pick_round(sv);
	int steps = sv->cfg.chunk_steps;
	for (Request* r : act) steps = std::min(steps, r->steps - r->planned);
	steps = std::max(steps, 0);
	Slot& sl = sv->slot[sv->next_slot];
	sv->next_slot = (sv->next_slot + 1) % kSlots;
	// the pinned mirrors are rewritten below: the copy that read them last must be done
	if (sl.used) MMSPX_HIP(hipEventSynchronize(sl.done));
	// work lists per form: group 0 the server's precision, group 1 the FP64
	// requests of an FP32 server; each group is one launch per step and kind
	std::vector<BItem> l2[2], l3[2];
	for (size_t m = 0; m < act.size(); m++) {
		if (act[m]->solo) continue;
		const int grp = act[m]->prec != sv->cfg.precision;
		member_items((int)m, act[m]->g, sv->cfg.model, sv->tile, l2[grp], l3[grp]);
	}
	int n2[2], n3[2], off[2], total = 0;
	for (int k = 0; k < 2; k++) {
		n2[k] = (int)l2[k].size();
		n3[k] = (int)l3[k].size();
		off[k] = total;
		total += n2[k] + n3[k];
	}
	int rc = ensure_slot(sl, (int)act.size(), total);
	if (rc) return rc;
	for (size_t m = 0; m < act.size(); m++) {
		const Request* r = act[m];
		BMember& bm = sl.h_mem[m];
		bm.buf[0] = r->dev;
		bm.buf[1] = r->dev + r->bytes;
		bm.g = r->g;
		bm.p = r->p;
		bm.phase = (int)((r->planned - sv->gstep) & 1);
	}
	for (int k = 0; k < 2; k++) {
		const std::vector<BItem> o2 = xcd_order(l2[k]), o3 = xcd_order(l3[k]);
		std::copy(o2.begin(), o2.end(), sl.h_items + off[k]);
		std::copy(o3.begin(), o3.end(), sl.h_items + off[k] + n2[k]);
	}
	MMSPX_HIP(hipMemcpyAsync(sl.d_mem, sl.h_mem, sizeof(BMember) * act.size(), hipMemcpyHostToDevice, sv->st));
	MMSPX_HIP(hipMemcpyAsync(sl.d_items, sl.h_items, sizeof(BItem) * (size_t)total, hipMemcpyHostToDevice, sv->st));
	int64_t launches = 0;
	const int form[2] = {sv->cfg.precision, MMSPX_F64};
	for (int k = 0; k < steps; k++) {
		for (int grp = 0; grp < 2; grp++) {
			if (n2[grp] + n3[grp] == 0) continue;
			const BItem* it = sl.d_items + off[grp];
			rc = launch_batch_step(sv->cfg.model, form[grp], 1, sv->tile, sl.d_mem, it, n2[grp], it + n2[grp], n3[grp],
			                       (int)((sv->gstep + k) & 1), sv->cus, sv->st, &launches);
			if (rc) return rc;
		}
	}
	// the large 3-D members: their own single-grid launches
	for (Request* r : act) {
		if (!r->solo) continue;
		const int spl = fused_steps_per_launch(sv->cfg.model, r->tile);
		for (int left = steps; left > 0;) {
			const int n = std::min(left, spl);
			const char* a = r->dev + (r->cur ? r->bytes : 0);
			char* b = r->dev + (r->cur ? 0 : r->bytes);
			rc = launch_fused_step(r->g, r->prec, r->p, a, b, sv->st2, r->tile, n, &launches);
			if (rc) return rc;
			r->cur ^= 1;
			left -= n;
		}
	}
	Round rd;
	rd.steps = steps;
	MMSPX_HIP(hipEventCreateWithFlags(&rd.ev, hipEventDisableTiming));
	MMSPX_HIP(hipEventRecord(rd.ev, sv->st));
	if (std::any_of(act.begin(), act.end(), [](const Request* r) { return r->solo; })) {
		MMSPX_HIP(hipEventCreateWithFlags(&rd.ev2, hipEventDisableTiming));
		MMSPX_HIP(hipEventRecord(rd.ev2, sv->st2));
	}
	MMSPX_HIP(hipEventRecord(sl.done, sv->st));
	sl.used = true;
	sv->gstep += steps;
	std::vector<Request*> keep;
	std::vector<IoJob> fin;
	for (Request* r : sv->resident) r->idle_rounds += 1;
	for (Request* r : act) {
		r->idle_rounds = 0;
		r->planned += steps;
		if (r->planned < r->steps) {
			keep.push_back(r);
			continue;
		}
		MMSPX_HIP(hipEventCreateWithFlags(&r->ev_last, hipEventDisableTiming));
		MMSPX_HIP(hipEventRecord(r->ev_last, stream_of(sv, r)));
		fin.push_back({r, false});
	}
	// a finishing member may be fetched and deleted once downloaded: the round
	// keeps only the members that stay
	rd.members = keep;
	{
		std::vector<Request*> stay;
		for (Request* r : sv->resident)
			if (r->planned < r->steps) stay.push_back(r);
		sv->resident.swap(stay);
	}
	sv->rounds.push_back(std::move(rd));
	std::lock_guard<std::mutex> lk(sv->mu);
	for (Request* r : sv->rounds.back().members) r->in_rounds += 1;
	sv->stats[3] += 1;
	sv->stats[6] += launches;
	for (const IoJob& j : fin) sv->down_jobs.push_back(j);
	if (!fin.empty()) sv->cv_down.notify_all();
	return 0;
}

// Retires finished rounds from the front of the queue (the oldest one is
// waited for if `block`); a round is finished when both compute streams have
// passed it.
static int reap_rounds(mmspx_server* sv, bool block)
{
	while (!sv->rounds.empty()) {
		Round& rd = sv->rounds.front();
		if (block) {
			MMSPX_HIP(hipEventSynchronize(rd.ev));
			if (rd.ev2) MMSPX_HIP(hipEventSynchronize(rd.ev2));
			block = false;
		} else {
			for (hipEvent_t e : {rd.ev, rd.ev2}) {
				if (!e) continue;
				const hipError_t q = hipEventQuery(e);
				if (q == hipErrorNotReady) return 0;
				if (q != hipSuccess) return fail(MMSPX_EDEVICE, std::string("round: ") + hipGetErrorString(q));
			}
		}
		{
			// a member fetched since this round was queued is deleted with its
			// last listing round
			std::lock_guard<std::mutex> lk(sv->mu);
			for (Request* r : rd.members) {
				r->done += rd.steps;
				r->in_rounds -= 1;
				if (r->fetched && r->in_rounds == 0) delete r;
			}
		}
		(void)hipEventDestroy(rd.ev);
		if (rd.ev2) (void)hipEventDestroy(rd.ev2);
		sv->rounds.pop_front();
	}
	return 0;
}

// Cancelled members leave before the next round; their buffers are freed behind the rounds queued so far.
static int drop_cancelled(mmspx_server* sv)
{
	std::vector<Request*> keep;
	std::vector<IoJob> drop;
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		for (Request* r : sv->resident) {
			if (r->cancel) drop.push_back({r, true});
			else keep.push_back(r);
		}
	}
	for (IoJob& j : drop) {
		MMSPX_HIP(hipEventCreateWithFlags(&j.r->ev_last, hipEventDisableTiming));
		MMSPX_HIP(hipEventRecord(j.r->ev_last, stream_of(sv, j.r)));
	}
	sv->resident.swap(keep);
	if (!drop.empty()) {
		std::lock_guard<std::mutex> lk(sv->mu);
		for (const IoJob& j : drop) sv->down_jobs.push_back(j);
		sv->cv_down.notify_all();
	}
	return 0;
}

// One pass of the worker: admit, take in uploaded requests, queue or retire a round.
// Returns 1 when the server stops, a negative code on a device error.
static int worker_pass(mmspx_server* sv)
{
	{
		std::unique_lock<std::mutex> lk(sv->mu);
		const bool idle = sv->resident.empty() && sv->rounds.empty() && sv->ready.empty();
		if (idle)
			sv->cv_work.wait_for(lk, std::chrono::milliseconds(20),
			                     [&] { return sv->stop || !sv->queue.empty() || !sv->ready.empty(); });
		if (sv->stop) return 1;
		// cancelled requests leave the queue wherever they stand
		for (auto it = sv->queue.begin(); it != sv->queue.end();) {
			if ((*it)->cancel) {
				finish(sv, *it, 4);
				it = sv->queue.erase(it);
			} else {
				++it;
			}
		}
		// admission: FIFO while slots and memory allow
		int live = (int)(sv->resident.size() + sv->ready.size()) + sv->uploading;
		bool admitted = false;
		while (!sv->queue.empty() && live < sv->cfg.max_slots) {
			Request* r = sv->queue.front();
			const size_t need = 2 * r->bytes;
			if (need > sv->cfg.hbm_budget) {
				sv->queue.pop_front();
				finish(sv, r, 3);
				continue;
			}
			if (sv->in_use + need > sv->cfg.hbm_budget) break;
			sv->queue.pop_front();
			sv->in_use += need;
			sv->stats[4] = std::max<int64_t>(sv->stats[4], (int64_t)sv->in_use);
			sv->uploading += 1;
			sv->up_jobs.push_back(r);
			live += 1;
			admitted = true;
		}
		if (admitted) sv->cv_up.notify_all();
		// uploaded requests are resident from the next round on
		for (Request* r : sv->ready) {
			sv->resident.push_back(r);
		}
		sv->ready.clear();
	}
	int rc = reap_rounds(sv, false);
	if (rc) return rc;
	rc = drop_cancelled(sv);
	if (rc) return rc;
	if (!sv->resident.empty() && (int)sv->rounds.size() < kRoundsAhead) {
		{
			std::lock_guard<std::mutex> lk(sv->mu);
			sv->stats[5] = std::max<int64_t>(sv->stats[5], (int64_t)sv->resident.size());
		}
		return enqueue_round(sv);
	}
	if (!sv->rounds.empty()) return reap_rounds(sv, true);
	return 0;
}

static void worker_main(mmspx_server* sv)
{
	(void)hipSetDevice(sv->device);
	for (;;) {
		const int rc = worker_pass(sv);
		if (rc == 0) continue;
		(void)hipStreamSynchronize(sv->st);
		(void)hipStreamSynchronize(sv->st2);
		if (rc == 1) break;
		// a device error: fail every request not yet finished and stop serving
		std::lock_guard<std::mutex> lk(sv->mu);
		if (!sv->worker_error) sv->worker_error = rc;
		for (auto& kv : sv->all)
			if (kv.second->status < 2) finish(sv, kv.second, 3);
		sv->queue.clear();
		sv->stop = true;
		sv->cv_up.notify_all();
		sv->cv_down.notify_all();
		break;
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
	if (cfg->fp64_above_steps < 0) return fail(MMSPX_EINVAL, "fp64_above_steps must be >= 0");
	auto* sv = new (std::nothrow) mmspx_server();
	if (!sv) return fail(MMSPX_ENOMEM, "host allocation failed");
	sv->cfg = *cfg;
	sv->t0 = std::chrono::steady_clock::now();
	int rc = 0;
	if (hipGetDevice(&sv->device) != hipSuccess || hipStreamCreateWithFlags(&sv->st, hipStreamNonBlocking) != hipSuccess ||
	    hipStreamCreateWithFlags(&sv->st2, hipStreamNonBlocking) != hipSuccess ||
	    hipStreamCreateWithFlags(&sv->ust, hipStreamNonBlocking) != hipSuccess ||
	    hipStreamCreateWithFlags(&sv->dst, hipStreamNonBlocking) != hipSuccess)
		rc = fail(MMSPX_EDEVICE, "HIP stream setup failed");
	for (int k = 0; k < kSlots && rc == 0; k++)
		if (hipEventCreateWithFlags(&sv->slot[k].done, hipEventDisableTiming) != hipSuccess)
			rc = fail(MMSPX_EDEVICE, "HIP event setup failed");
	if (rc == 0) {
		hipMemPoolProps props{};
		props.allocType = hipMemAllocationTypePinned;
		props.handleTypes = hipMemHandleTypeNone;
		props.location.type = hipMemLocationTypeDevice;
		props.location.id = sv->device;
		// keep freed blocks reserved (up to the budget) for the next admissions
		uint64_t keep = cfg->hbm_budget;
		if (hipMemPoolCreate(&sv->pool, &props) != hipSuccess ||
		    hipMemPoolSetAttribute(sv->pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess)
			rc = fail(MMSPX_EDEVICE, "HIP memory pool setup failed");
	}
	if (rc == 0 && (hipMalloc((void**)&sv->upart, 2 * kRedBlocks * sizeof(double)) != hipSuccess ||
	                hipMalloc((void**)&sv->dpart, 2 * kRedBlocks * sizeof(double)) != hipSuccess))
		rc = fail(MMSPX_ENOMEM, "device allocation failed");
	if (rc) {
		for (int k = 0; k < kSlots; k++)
			if (sv->slot[k].done) (void)hipEventDestroy(sv->slot[k].done);
		if (sv->upart) (void)hipFree(sv->upart);
		if (sv->dpart) (void)hipFree(sv->dpart);
		if (sv->pool) (void)hipMemPoolDestroy(sv->pool);
		if (sv->st) (void)hipStreamDestroy(sv->st);
		if (sv->st2) (void)hipStreamDestroy(sv->st2);
		if (sv->ust) (void)hipStreamDestroy(sv->ust);
		if (sv->dst) (void)hipStreamDestroy(sv->dst);
		delete sv;
		return rc;
	}
	sv->cus = device_cus();
	sv->uploader = std::thread(uploader_main, sv);
	sv->downloader = std::thread(downloader_main, sv);
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
	sv->cv_up.notify_all();
	sv->cv_down.notify_all();
	if (sv->worker.joinable()) sv->worker.join();
	if (sv->uploader.joinable()) sv->uploader.join();
	if (sv->downloader.joinable()) sv->downloader.join();
	(void)hipStreamSynchronize(sv->st);
	(void)hipStreamSynchronize(sv->st2);
	(void)hipStreamSynchronize(sv->ust);
	(void)hipStreamSynchronize(sv->dst);
	std::set<Request*> fetched;  // fetched, still listed by rounds never retired
	for (auto& rd : sv->rounds) {
		if (rd.ev) (void)hipEventDestroy(rd.ev);
		if (rd.ev2) (void)hipEventDestroy(rd.ev2);
		for (Request* r : rd.members)
			if (r->fetched) fetched.insert(r);
	}
	for (Request* r : fetched) delete r;
	for (auto& kv : sv->all) {
		Request* r = kv.second;
		if (r->dev) (void)hipFreeAsync(r->dev, sv->dst);
		if (r->ev_last) (void)hipEventDestroy(r->ev_last);
		delete r;
	}
	(void)hipStreamSynchronize(sv->dst);
	for (auto& sl : sv->slot) {
		if (sl.d_mem) (void)hipFree(sl.d_mem);
		if (sl.d_items) (void)hipFree(sl.d_items);
		if (sl.h_mem) (void)hipHostFree(sl.h_mem);
		if (sl.h_items) (void)hipHostFree(sl.h_items);
		if (sl.done) (void)hipEventDestroy(sl.done);
	}
	(void)hipFree(sv->upart);
	(void)hipFree(sv->dpart);
	(void)hipMemPoolDestroy(sv->pool);
	(void)hipStreamDestroy(sv->st);
	(void)hipStreamDestroy(sv->st2);
	(void)hipStreamDestroy(sv->ust);
	(void)hipStreamDestroy(sv->dst);
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
	r->prec = sv->cfg.precision;
	if (sv->cfg.precision == MMSPX_F32 && sv->cfg.fp64_above_steps > 0 && steps > sv->cfg.fp64_above_steps)
		r->prec = MMSPX_F64;
	r->desc = *desc;
	r->desc.precision = r->prec;
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
	if (r->g.dim == 3 && r->g.nodes >= kSoloNodes && fused_supported(r->g, r->prec, sv->cfg.model)) {
		const TuneChoice t = default_choice(r->g, r->prec, sv->cfg.model);
		if (t.variant == 2) {
			r->solo = true;
			r->tile = t;
		}
	}
	const size_t n = (size_t)r->g.nodes;
	r->bytes = (n * elem_of(r) + 255) / 256 * 256;
	r->init.resize(n * elem_of(r));
	if (r->prec == MMSPX_F64) {
		std::copy(init, init + n, (double*)r->init.data());
	} else {
		float* f = (float*)r->init.data();
		for (size_t i = 0; i < n; i++) f[i] = (float)init[i];
	}
	std::lock_guard<std::mutex> lk(sv->mu);
	if (sv->stop) {
		delete r;
		return fail(MMSPX_EBUSY, "the server has stopped");
	}
	r->id = sv->next_id++;
	r->t_submit = sv->now();
	if (r->prec != sv->cfg.precision) sv->stats[7] += 1;
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
	info->steps_done = r->status == 2 ? r->steps : r->done;
	info->t_submit = r->t_submit;
	info->t_admit = r->t_admit;
	info->t_done = r->t_done;
	info->mass0 = r->mass0;
	info->mass1 = r->mass1;
	info->precision = r->prec;
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
	Request* r = nullptr;
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		auto it = sv->all.find(id);
		if (it == sv->all.end()) return fail(MMSPX_ENOTFOUND, "unknown request id");
		r = it->second;
		if (r->status != 2) return fail(MMSPX_EBUSY, "request has not completed");
		if (count != (size_t)r->g.nodes) return fail(MMSPX_EINVAL, "output size does not match the request's grid");
		sv->all.erase(it);
	}
	// the fetching thread converts to double
	if (r->prec == MMSPX_F64) {
		const double* d = (const double*)r->out.data();
		std::copy(d, d + count, out);
	} else {
		const float* f = (const float*)r->out.data();
		for (size_t i = 0; i < count; i++) out[i] = f[i];
	}
	// a round the worker has not retired yet may still list the request (its
	// last steps are done, the round's other work may not be): then the worker
	// deletes it when it retires that round
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		r->fetched = true;
		if (r->in_rounds > 0) return 0;
	}
	delete r;
	return 0;
}

double mmspx_server_clock(mmspx_server* sv) { return sv ? sv->now() : 0.0; }

int mmspx_server_stats(mmspx_server* sv, int64_t* out, int nout)
{
	if (!sv || !out) return fail(MMSPX_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(sv->mu);
	for (int k = 0; k < nout && k < 8; k++) out[k] = sv->stats[k];
	return sv->worker_error;
}

}  // extern "C"
